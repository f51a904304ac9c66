// Drop-in fused BLAS for the gfx950 library (include/qphix/blas_new_c.h of
// JeffersonLab/qphix). Same names, argument order and meaning; the trailing
// n_blas_simt thread count is accepted and ignored. Reductions return FP64
// host scalars, summed over flavours in the num_flav overloads.
#pragma once

#include "geometry.h"

namespace QPhiX {

#define QPHIX_SP typename Geometry<FT, V, S, compress>::FourSpinorBlock
#define QPHIX_G const Geometry<FT, V, S, compress> &geom
#define QPHIX_T template <typename FT, int V, int S, bool compress>
#define QPHIX_C(call) detail::check(call, #call)

QPHIX_T void copySpinor(QPHIX_SP *res, const QPHIX_SP *src, QPHIX_G, int = 1) { QPHIX_C(qphix_copy_spinor(geom.handle(), res, src)); }
QPHIX_T void zeroSpinor(QPHIX_SP *res, QPHIX_G, int = 1) { QPHIX_C(qphix_zero_spinor(geom.handle(), res)); }
QPHIX_T void axy(const double alpha, const QPHIX_SP *x, QPHIX_SP *y, QPHIX_G, int = 1) { QPHIX_C(qphix_axy(geom.handle(), alpha, x, y)); }
QPHIX_T void ax(const double alpha, QPHIX_SP *x, QPHIX_G, int = 1) { QPHIX_C(qphix_ax(geom.handle(), alpha, x)); }
QPHIX_T void ypeqx(const QPHIX_SP *x, QPHIX_SP *y, QPHIX_G, int = 1) { QPHIX_C(qphix_ypeqx(geom.handle(), x, y)); }
QPHIX_T void ymeqx(const QPHIX_SP *x, QPHIX_SP *y, QPHIX_G, int = 1) { QPHIX_C(qphix_ymeqx(geom.handle(), x, y)); }
QPHIX_T void aypx(const double alpha, const QPHIX_SP *x, QPHIX_SP *y, QPHIX_G, int = 1) { QPHIX_C(qphix_aypx(geom.handle(), alpha, x, y)); }
QPHIX_T void twisted_mass(const double apimu[2], const QPHIX_SP *x, QPHIX_SP *y, QPHIX_G, int = 1)
{
  QPHIX_C(qphix_twisted_mass(geom.handle(), apimu, x, y));
}
QPHIX_T void two_flav_twisted_mass(const double apimu[2], const double epsilon, const QPHIX_SP *const *x,
                                   QPHIX_SP *const *y, QPHIX_G, int = 1)
{
  const void *xv[2] = {x[0], x[1]};
  void *yv[2] = {y[0], y[1]};
  QPHIX_C(qphix_two_flav_twisted_mass(geom.handle(), apimu, epsilon, xv, yv));
}
QPHIX_T void axpy(const double alpha, const QPHIX_SP *x, QPHIX_SP *y, QPHIX_G, int = 1) { QPHIX_C(qphix_axpy(geom.handle(), alpha, x, y)); }
QPHIX_T void axpby(const double alpha, const QPHIX_SP *x, const double beta, QPHIX_SP *y, QPHIX_G, int = 1)
{
  QPHIX_C(qphix_axpby(geom.handle(), alpha, x, beta, y));
}
QPHIX_T void norm2Spinor(double &n2, const QPHIX_SP *x, QPHIX_G, int = 1, bool = true) { QPHIX_C(qphix_norm2_spinor(geom.handle(), x, &n2)); }
QPHIX_T void axpyNorm2(const double alpha, const QPHIX_SP *x, QPHIX_SP *y, double &norm2y, QPHIX_G, int = 1)
{
  QPHIX_C(qphix_axpy_norm2(geom.handle(), alpha, x, y, &norm2y));
}
QPHIX_T void xmyNorm2Spinor(QPHIX_SP *res, const QPHIX_SP *x, QPHIX_SP *y, double &n2res, QPHIX_G, int = 1, bool = true)
{
  QPHIX_C(qphix_xmy_norm2(geom.handle(), res, x, y, &n2res));
}
QPHIX_T void xmy2Norm2Spinor(const QPHIX_SP *x, QPHIX_SP *y, double &n2res, QPHIX_G, int = 1, bool = true)
{
  QPHIX_C(qphix_xmy2_norm2(geom.handle(), x, y, &n2res));
}
QPHIX_T void rmammpNorm2rxpap(QPHIX_SP *r, const double ar, QPHIX_SP *mmp, double &cp, QPHIX_SP *x, QPHIX_SP *p, QPHIX_G, int = 1)
{
  QPHIX_C(qphix_rmammp_norm2_rxpap(geom.handle(), r, ar, mmp, x, p, &cp));
}
QPHIX_T void richardson_rxupdateNormR(QPHIX_SP *x, QPHIX_SP *r, const QPHIX_SP *delta_x, const QPHIX_SP *delta_r,
                                      double &r_norm, QPHIX_G, int = 1)
{
  QPHIX_C(qphix_richardson_rxupdate_norm_r(geom.handle(), x, r, delta_x, delta_r, &r_norm));
}
QPHIX_T void bicgstab_xmy(const QPHIX_SP *x, QPHIX_SP *y, QPHIX_G, int = 1) { QPHIX_C(qphix_bicgstab_xmy(geom.handle(), x, y)); }
QPHIX_T void innerProduct(double results[2], const QPHIX_SP *l, const QPHIX_SP *r, QPHIX_G, int = 1, bool = true)
{
  QPHIX_C(qphix_inner_product(geom.handle(), l, r, results));
}
QPHIX_T void innerProductNorm(double results[3], const QPHIX_SP *l, const QPHIX_SP *r, QPHIX_G, int = 1, bool = true)
{
  QPHIX_C(qphix_inner_product_norm(geom.handle(), l, r, results));
}
QPHIX_T void bicgstab_p_update(const QPHIX_SP *r, QPHIX_SP *p, const QPHIX_SP *v, double beta[2], double omega[2], QPHIX_G, int = 1)
{
  QPHIX_C(qphix_bicgstab_p_update(geom.handle(), r, p, v, beta, omega));
}
QPHIX_T void bicgstab_s_update(double alpha[2], QPHIX_SP *s, const QPHIX_SP *v, QPHIX_G, int = 1)
{
  QPHIX_C(qphix_bicgstab_s_update(geom.handle(), alpha, s, v));
}
QPHIX_T void caxpySpinor(double alpha[2], const QPHIX_SP *x, QPHIX_SP *y, QPHIX_G, int = 1)
{
  QPHIX_C(qphix_caxpy(geom.handle(), alpha, x, y));
}
QPHIX_T void bicgstab_rxupdate(QPHIX_SP *x, QPHIX_SP *r, const QPHIX_SP *t, const QPHIX_SP *p, double omega[2],
                               double alpha[2], double &r_norm, QPHIX_G, int = 1)
{
  QPHIX_C(qphix_bicgstab_rxupdate(geom.handle(), x, r, t, p, omega, alpha, &r_norm));
}
QPHIX_T void mr_rxupdate(QPHIX_SP *x, QPHIX_SP *r, const QPHIX_SP *Mr, double a[2], QPHIX_G, int = 1)
{
  QPHIX_C(qphix_mr_rxupdate(geom.handle(), x, r, Mr, a));
}

// precision conversion with a scale factor (blas_new_c.h:852-1031)
template <typename FTOut, int VOut, int SOut, bool COut, typename FTIn, int VIn, int SIn, bool CIn>
void convert(typename Geometry<FTOut, VOut, SOut, COut>::FourSpinorBlock *out, double scale_factor,
             const typename Geometry<FTIn, VIn, SIn, CIn>::FourSpinorBlock *in,
             const Geometry<FTOut, VOut, SOut, COut> &geom_out, const Geometry<FTIn, VIn, SIn, CIn> &geom_in, int = 1)
{
  detail::check(qphix_convert_spinor(geom_out.handle(), out, scale_factor, geom_in.handle(), in), "convert");
}

// multi-flavour forms (the num_flav template overloads of the reference)
template <typename FT, int V, int S, bool compress, int num_flav>
void norm2Spinor(double &n2, const QPHIX_SP *const x[num_flav], QPHIX_G, int = 1)
{
  n2 = 0;
  for (int f = 0; f < num_flav; f++) {
    double t;
    QPHIX_C(qphix_norm2_spinor(geom.handle(), x[f], &t));
    n2 += t;
  }
}
template <typename FT, int V, int S, bool compress, int num_flav>
void copySpinor(QPHIX_SP *const res[num_flav], const QPHIX_SP *const src[num_flav], QPHIX_G, int = 1)
{
  for (int f = 0; f < num_flav; f++) QPHIX_C(qphix_copy_spinor(geom.handle(), res[f], src[f]));
}
template <typename FT, int V, int S, bool compress, int num_flav>
void zeroSpinor(QPHIX_SP *const res[num_flav], QPHIX_G, int = 1)
{
  for (int f = 0; f < num_flav; f++) QPHIX_C(qphix_zero_spinor(geom.handle(), res[f]));
}
template <typename FT, int V, int S, bool compress, int num_flav>
void axpy(const double alpha, const QPHIX_SP *const x[num_flav], QPHIX_SP *const y[num_flav], QPHIX_G, int = 1)
{
  for (int f = 0; f < num_flav; f++) QPHIX_C(qphix_axpy(geom.handle(), alpha, x[f], y[f]));
}
template <typename FT, int V, int S, bool compress, int num_flav>
void aypx(const double alpha, const QPHIX_SP *const x[num_flav], QPHIX_SP *const y[num_flav], QPHIX_G, int = 1)
{
  for (int f = 0; f < num_flav; f++) QPHIX_C(qphix_aypx(geom.handle(), alpha, x[f], y[f]));
}
template <typename FT, int V, int S, bool compress, int num_flav>
void xmyNorm2Spinor(QPHIX_SP *const res[num_flav], const QPHIX_SP *const x[num_flav], QPHIX_SP *const y[num_flav],
                    double &n2res, QPHIX_G, int = 1)
{
  n2res = 0;
  for (int f = 0; f < num_flav; f++) {
    double t;
    QPHIX_C(qphix_xmy_norm2(geom.handle(), res[f], x[f], y[f], &t));
    n2res += t;
  }
}

#undef QPHIX_SP
#undef QPHIX_G
#undef QPHIX_T
#undef QPHIX_C

}  // namespace QPhiX
