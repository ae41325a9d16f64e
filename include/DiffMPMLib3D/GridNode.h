// Drop-in header name of the reference (DiffMPMLib3D/GridNode.h); everything is declared in DiffMPMLib3D.h.
#pragma once
#include "DiffMPMLib3D.h"
