// BLAS on full (both-checkerboard) spinors (blas_full_spinor.h of
// JeffersonLab/qphix): each function applies the single-checkerboard form of
// blas_new_c.h to checkerboards [min_cb, max_cb) and sums the reductions.
// n_blas_simt is accepted and ignored.
#pragma once

#include "blas_new_c.h"
#include "full_spinor.h"

namespace QPhiX {

#define QPHIX_FT template <typename FT, int V, int S, bool compress>
#define QPHIX_FS FullSpinor<FT, V, S, compress>
#define QPHIX_FG const Geometry<FT, V, S, compress> &geom, int n_blas_simt, const int min_cb = 0, const int max_cb = 2

QPHIX_FT void copySpinor(QPHIX_FS &result, const QPHIX_FS &src, QPHIX_FG)
{
  for (int cb = min_cb; cb < max_cb; ++cb)
    copySpinor<FT, V, S, compress>(result.getCBData(cb), src.getCBData(cb), geom, n_blas_simt);
}
QPHIX_FT void zeroSpinor(QPHIX_FS &result, QPHIX_FG)
{
  for (int cb = min_cb; cb < max_cb; ++cb) zeroSpinor<FT, V, S, compress>(result.getCBData(cb), geom, n_blas_simt);
}
QPHIX_FT void axSpinor(const double alpha, QPHIX_FS &x, QPHIX_FG)
{
  for (int cb = min_cb; cb < max_cb; ++cb) ax<FT, V, S, compress>(alpha, x.getCBData(cb), geom, n_blas_simt);
}
QPHIX_FT void caxpySpinor(double alpha[2], const QPHIX_FS &x, QPHIX_FS &y, QPHIX_FG)
{
  for (int cb = min_cb; cb < max_cb; ++cb)
    caxpySpinor<FT, V, S, compress>(alpha, x.getCBData(cb), y.getCBData(cb), geom, n_blas_simt);
}
QPHIX_FT void axpySpinor(double alpha, const QPHIX_FS &x, QPHIX_FS &y, QPHIX_FG)
{
  for (int cb = min_cb; cb < max_cb; ++cb)
    axpy<FT, V, S, compress>(alpha, x.getCBData(cb), y.getCBData(cb), geom, n_blas_simt);
}
QPHIX_FT void ypeqxSpinor(const QPHIX_FS &x, QPHIX_FS &y, QPHIX_FG)
{
  for (int cb = min_cb; cb < max_cb; ++cb)
    ypeqx<FT, V, S, compress>(x.getCBData(cb), y.getCBData(cb), geom, n_blas_simt);
}
QPHIX_FT void ymeqxSpinor(const QPHIX_FS &x, QPHIX_FS &y, QPHIX_FG)
{
  for (int cb = min_cb; cb < max_cb; ++cb)
    ymeqx<FT, V, S, compress>(x.getCBData(cb), y.getCBData(cb), geom, n_blas_simt);
}
// y = y - x, r_norm = |y|^2 over the checkerboards
QPHIX_FT void xmy2Norm2Spinor(const QPHIX_FS &x, QPHIX_FS &y, double &r_norm, QPHIX_FG)
{
  r_norm = 0;
  for (int cb = min_cb; cb < max_cb; ++cb) {
    double n = 0;
    xmy2Norm2Spinor<FT, V, S, compress>(x.getCBData(cb), y.getCBData(cb), n, geom, n_blas_simt, false);
    r_norm += n;
  }
}
QPHIX_FT void innerProductSpinor(double results[2], const QPHIX_FS &x, const QPHIX_FS &y, QPHIX_FG)
{
  results[0] = results[1] = 0;
  for (int cb = min_cb; cb < max_cb; ++cb) {
    double r[2] = {0, 0};
    innerProduct<FT, V, S, compress>(r, x.getCBData(cb), y.getCBData(cb), geom, n_blas_simt, false);
    results[0] += r[0];
    results[1] += r[1];
  }
}
QPHIX_FT void norm2Spinor(double &n2, const QPHIX_FS &x, QPHIX_FG)
{
  n2 = 0;
  for (int cb = min_cb; cb < max_cb; ++cb) {
    double n = 0;
    norm2Spinor<FT, V, S, compress>(n, x.getCBData(cb), geom, n_blas_simt, false);
    n2 += n;
  }
}

#undef QPHIX_FT
#undef QPHIX_FS
#undef QPHIX_FG

}  // namespace QPhiX
