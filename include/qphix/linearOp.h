// Reference file name kept for drop-in includes; the classes live in linear_ops.h.
#pragma once
#include "linear_ops.h"
