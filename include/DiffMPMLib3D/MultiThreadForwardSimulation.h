// Drop-in header name of the reference (DiffMPMLib3D/MultiThreadForwardSimulation.h); everything is declared in DiffMPMLib3D.h.
#pragma once
#include "DiffMPMLib3D.h"
