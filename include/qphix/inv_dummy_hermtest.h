// InvDummyHermTest (inv_dummy_hermtest.h of JeffersonLab/qphix): a "solver"
// that checks the hermiticity relation of M through the solver interface,
// for callers that can reach solvers but not the operator itself. Given two
// fields x and rhs it forms <x | M^+ M rhs> and <M x | M rhs> (equal for a
// correct M^+), prints both and reports n_iters = 1, mv_apps = 3; x and rhs are
// not changed.
#pragma once

#include <cstdio>
#include <type_traits>

#include "blas_new_c.h"
#include "solvers.h"

namespace QPhiX {

template <typename FT, int veclen, int soalen, bool compress12,
          typename EvenOddLinearOperatorBase = EvenOddLinearOperator<FT, veclen, soalen, compress12>>
class InvDummyHermTest
    : public AbstractSolver<FT, veclen, soalen, compress12, EvenOddLinearOperatorBase::num_flav> {
 public:
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock Spinor;
  static constexpr int num_flav = EvenOddLinearOperatorBase::num_flav;

  InvDummyHermTest(EvenOddLinearOperatorBase &M, int MaxIters) : M_(M), geom_(M.getGeometry()), max_iters_(MaxIters)
  {
    for (int f = 0; f < num_flav; ++f) {
      m_rhs_[f] = geom_.allocCBFourSpinor();
      mdag_m_rhs_[f] = geom_.allocCBFourSpinor();
      m_x_[f] = geom_.allocCBFourSpinor();
    }
  }
  ~InvDummyHermTest()
  {
    for (int f = 0; f < num_flav; ++f) {
      geom_.free(m_rhs_[f]);
      geom_.free(mdag_m_rhs_[f]);
      geom_.free(m_x_[f]);
    }
  }
  InvDummyHermTest(const InvDummyHermTest &) = delete;
  InvDummyHermTest &operator=(const InvDummyHermTest &) = delete;

  using AbstractSolver<FT, veclen, soalen, compress12, num_flav>::operator();

  void operator()(Spinor *const x[num_flav], const Spinor *const rhs[num_flav], const double RsdTarget,
                  int &n_iters, double &rsd_sq_final, unsigned long &site_flops, unsigned long &mv_apps, int isign,
                  bool verboseP, int cb = 1, ResiduumType residType = RELATIVE) const override
  {
    (void)RsdTarget; (void)rsd_sq_final; (void)residType;
    if (verboseP) std::printf("Entering isign=%d hermiticity test with num_flav=%d and cb=%d\n", isign, num_flav, cb);
    site_flops = 0;
    mv_apps = 3;
    n_iters = 1;
    apply(m_rhs_, rhs, isign, cb);
    apply(mdag_m_rhs_, m_rhs_, -isign, cb);
    apply(m_x_, x, isign, cb);
    sum_inner(x_mdagm_rhs_, x, mdag_m_rhs_);
    sum_inner(mx_m_rhs_, m_x_, m_rhs_);
    std::printf("\n\nHERMITICITY TEST\n %20s: %lf %lf \n %20s: %lf %lf\n\n\n", "< x | Mdag M y >", x_mdagm_rhs_[0],
                x_mdagm_rhs_[1], "< M x | M y >", mx_m_rhs_[0], mx_m_rhs_[1]);
  }

  Geometry<FT, veclen, soalen, compress12> &getGeometry() override { return geom_; }
  int maxIters() const { return max_iters_; }
  // the two products of the last test (the reference only prints them)
  void lastProducts(double x_mdagm_rhs[2], double mx_m_rhs[2]) const
  {
    for (int k = 0; k < 2; ++k) {
      x_mdagm_rhs[k] = x_mdagm_rhs_[k];
      mx_m_rhs[k] = mx_m_rhs_[k];
    }
  }

 private:
  // one-flavour operators take single spinors, two-flavour ones arrays
  void apply(Spinor *const res[num_flav], const Spinor *const in[num_flav], int isign, int cb) const
  {
    apply_n(res, in, isign, cb, std::integral_constant<int, num_flav>());
  }
  void apply_n(Spinor *const res[], const Spinor *const in[], int isign, int cb, std::integral_constant<int, 1>) const
  {
    M_(res[0], in[0], isign, cb);
  }
  void apply_n(Spinor *const res[], const Spinor *const in[], int isign, int cb, std::integral_constant<int, 2>) const
  {
    M_(res, in, isign, cb);
  }
  void sum_inner(double out[2], const Spinor *const l[num_flav], const Spinor *const r[num_flav]) const
  {
    out[0] = out[1] = 0;
    for (int f = 0; f < num_flav; ++f) {
      double ip[2];
      innerProduct<FT, veclen, soalen, compress12>(ip, l[f], r[f], geom_);
      out[0] += ip[0];
      out[1] += ip[1];
    }
  }

  EvenOddLinearOperatorBase &M_;
  Geometry<FT, veclen, soalen, compress12> &geom_;
  int max_iters_;
  Spinor *m_rhs_[num_flav];
  Spinor *mdag_m_rhs_[num_flav];
  Spinor *m_x_[num_flav];
  mutable double x_mdagm_rhs_[2] = {0, 0};
  mutable double mx_m_rhs_[2] = {0, 0};
};

}  // namespace QPhiX
