// Reference file name kept for drop-in includes; the classes live in dslash.h.
#pragma once
#include "dslash.h"
