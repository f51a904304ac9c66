// Reference file name kept for drop-in includes; the classes live in solvers.h.
#pragma once
#include "solvers.h"
