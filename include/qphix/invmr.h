// Reference file name kept for drop-in includes; InvMR and InvMRSmoother live in solvers.h.
#pragma once
#include "solvers.h"
