// Host/device declarations of the batched (segmented) solver kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace qphix_amd {

// one system's vectors for one batched BLAS step
struct SegVec {
  const void *x, *y;
  void *o0, *o1;
  long vh;
  size_t vs;
  int sys;
};

enum BatchOp { BB_NORM2 = 0, BB_COPY, BB_ZERO, BB_XMYNORM2, BB_RMAMMP, BB_AYPX };

// per-system CG state in device memory; iters is the 0-based iteration of
// convergence (invcg.h:343), loops the iterations performed
struct CgState {
  double c, d, a, b, cp, rsd_sq;
  int iters, max_iters, converged, loops;
};

// off[nseg] = total chunks of 256 sites; out[sys] receives reductions
hipError_t batch_blas_launch(int prec, int op, const SegVec *segs, const int *off, int nseg, const int *active,
                             const double *scal, double *partials, double *out, int grid, hipStream_t st);
hipError_t batch_cg_alpha(CgState *st_, const double *d, double *alpha, const int *active, int n, hipStream_t st);
hipError_t batch_cg_update(CgState *st_, const double *cp, double *beta, int *active, int n, hipStream_t st);
hipError_t batch_cg_init(CgState *st_, const double *cp, const double *rsd_sq, int *active, int max_iters, int n,
                         hipStream_t st);

}  // namespace qphix_amd
