// Umbrella header of the drop-in QPhiX C++ API for the gfx950 library.
#pragma once
#include "geometry.h"
#include "dslash.h"
#include "linear_ops.h"
#include "solvers.h"
#include "full_spinor.h"
#include "unprec_solver_wrapper.h"
#include "syssolver_linop_clover.h"
#include "blas_new_c.h"
#include "blas_full_spinor.h"
#include "arith_type.h"
#include "clover_product.h"
#include "inv_dummy_hermtest.h"
