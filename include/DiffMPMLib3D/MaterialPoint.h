// Drop-in header name of the reference (DiffMPMLib3D/MaterialPoint.h); everything is declared in DiffMPMLib3D.h.
#pragma once
#include "DiffMPMLib3D.h"
