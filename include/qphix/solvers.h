// Drop-in solvers for the gfx950 library: AbstractSolver / AbstractMultiSolver
// (abs_solver.h:13-267), InvCG (invcg.h:25-694), InvBiCGStab
// (invbicgstab.h:18-619), MInvCG (minvcg.h:17-452) and InvRichardsonMultiPrec
// (inv_richardson_multiprec.h:11-320) of JeffersonLab/qphix.
//
// With a library operator (native() != nullptr) each solve is one C-ABI call
// that runs the whole Krylov loop on the device stream. InvCG and InvBiCGStab
// also accept user-defined operators: the loop then runs here, calling the
// operator's operator() and the library's fused BLAS.
#pragma once

#include <cmath>
#include <cstdio>
#include <vector>

#include "linear_ops.h"

namespace QPhiX {

enum ResiduumType { ABSOLUTE, RELATIVE, INVALID };

template <typename FT, int V, int S, bool compress12, int num_flav = 1>
class AbstractSolver {
 public:
  typedef typename Geometry<FT, V, S, compress12>::FourSpinorBlock Spinor;
  virtual ~AbstractSolver() {}
  virtual void operator()(Spinor *x, const Spinor *rhs, const double RsdTarget, int &niters,
                          double &rsd_sq_final, unsigned long &site_flops, unsigned long &mv_apps, int isign,
                          bool verboseP, int target_cb = 1, ResiduumType residType = RELATIVE) const
  {
    Spinor *xa[1] = {x};
    const Spinor *ra[1] = {rhs};
    (*this)(xa, ra, RsdTarget, niters, rsd_sq_final, site_flops, mv_apps, isign, verboseP, target_cb, residType);
  }
  virtual void operator()(Spinor *const x[num_flav], const Spinor *const rhs[num_flav], const double RsdTarget,
                          int &niters, double &rsd_sq_final, unsigned long &site_flops, unsigned long &mv_apps,
                          int isign, bool verboseP, int target_cb = 1,
                          ResiduumType residType = RELATIVE) const = 0;
  virtual Geometry<FT, V, S, compress12> &getGeometry() = 0;
  // library solver backing this object (used by InvRichardsonMultiPrec)
  virtual qphix_solver *native() const { return nullptr; }
};

template <typename FT, int V, int S, bool compress12, int num_flav = 1>
class AbstractMultiSolver {
 public:
  typedef typename Geometry<FT, V, S, compress12>::FourSpinorBlock Spinor;
  virtual ~AbstractMultiSolver() {}
  virtual void operator()(Spinor **const *x, const Spinor *const *rhs, const int n_shift, const double *shifts,
                          const double *RsdTarget, int &n_iters, double *rsd_sq_final, unsigned long &site_flops,
                          unsigned long &mv_apps, int isign, bool verboseP, int target_cb = 1) const = 0;
  virtual Geometry<FT, V, S, compress12> &getGeometry() = 0;
};

namespace detail {
template <typename FT, int V, int S, bool C, int NF, class Op> class NativeSolverBase
    : public AbstractSolver<FT, V, S, C, NF> {
 public:
  typedef typename Geometry<FT, V, S, C>::FourSpinorBlock Spinor;
  using AbstractSolver<FT, V, S, C, NF>::operator();
  explicit NativeSolverBase(Op &M) : M_(M) {}
  ~NativeSolverBase()
  {
    if (s_) qphix_solver_destroy(s_);
  }
  void operator()(Spinor *const x[NF], const Spinor *const rhs[NF], const double RsdTarget, int &n_iters,
                  double &rsd_sq_final, unsigned long &site_flops, unsigned long &mv_apps, int isign, bool verboseP,
                  int cb = 1, ResiduumType residType = RELATIVE) const override
  {
    if (!s_) {
      generic(x, rhs, RsdTarget, n_iters, rsd_sq_final, site_flops, mv_apps, isign, verboseP, cb, residType);
      return;
    }
    void *xv[NF];
    const void *rv[NF];
    for (int f = 0; f < NF; f++) {
      xv[f] = x[f];
      rv[f] = rhs[f];
    }
    qphix_solve_stats st;
    check(qphix_solve(s_, xv, rv, RsdTarget, isign, verboseP ? 1 : 0, cb,
                      residType == ABSOLUTE ? QPHIX_RESID_ABSOLUTE : QPHIX_RESID_RELATIVE, &st),
          "solver");
    n_iters = st.n_iters;
    rsd_sq_final = st.rsd_sq_final;
    site_flops = st.site_flops;
    mv_apps = st.mv_apps;
    if (st.status == QPHIX_ENOCONV) std::printf("Solver did not converge in %d iterations\n", st.n_iters);
  }
  Geometry<FT, V, S, C> &getGeometry() override { return M_.getGeometry(); }
  qphix_solver *native() const override { return s_; }

 protected:
  virtual void generic(Spinor *const x[NF], const Spinor *const rhs[NF], double, int &, double &,
                       unsigned long &, unsigned long &, int, bool, int, ResiduumType) const
  {
    (void)x;
    (void)rhs;
    throw std::runtime_error("this solver needs a library operator");
  }
  Op &M_;
  qphix_solver *s_ = nullptr;
};
}  // namespace detail

template <typename FT, int veclen, int soalen, bool compress12,
          typename EvenOddLinearOperatorBase = EvenOddLinearOperator<FT, veclen, soalen, compress12>>
class InvCG : public detail::NativeSolverBase<FT, veclen, soalen, compress12, EvenOddLinearOperatorBase::num_flav,
                                              EvenOddLinearOperatorBase> {
  typedef detail::NativeSolverBase<FT, veclen, soalen, compress12, EvenOddLinearOperatorBase::num_flav,
                                   EvenOddLinearOperatorBase>
      B;

 public:
  static constexpr int num_flav = EvenOddLinearOperatorBase::num_flav;
  typedef typename B::Spinor Spinor;
  InvCG(EvenOddLinearOperatorBase &M, int MaxIters) : B(M), max_iters_(MaxIters)
  {
    if (M.native()) detail::check(qphix_cg_create(M.native(), MaxIters, &this->s_), "InvCG");
  }

 protected:
  // CG on M^+M for a user-defined operator (invcg.h:88-456), with the
  // reference's counters (site_flops at invcg.h:180,224,308,337,416, mv_apps
  // at :210,294); rsd_sq_final on convergence is the updated |r|^2
  void generic(Spinor *const x[num_flav], const Spinor *const rhs[num_flav], double RsdTarget, int &n_iters,
               double &rsd_sq_final, unsigned long &site_flops, unsigned long &mv_apps, int isign, bool verbose,
               int cb, ResiduumType rt) const override
  {
    Geometry<FT, veclen, soalen, compress12> &g = this->M_.getGeometry();
    const qphix_geom *gh = g.handle();
    std::vector<Spinor *> w;
    for (int k = 0; k < 4 * num_flav; k++) w.push_back(g.allocCBFourSpinor());
    Spinor **mp = &w[0], **mmp = &w[num_flav], **p = &w[2 * num_flav], **r = &w[3 * num_flav];
    auto M = [&](Spinor **o, Spinor *const *i, int is) { apply(o, i, is, cb); };
    auto norm2 = [&](Spinor *const *v) {
      double s = 0, t;
      for (int f = 0; f < num_flav; f++) {
        detail::check(qphix_norm2_spinor(gh, v[f], &t), "norm2");
        s += t;
      }
      return s;
    };
    const unsigned long sf = 12ul * num_flav;
    site_flops = 0;
    mv_apps = 0;
    double chi_sq = 0, t;
    for (int f = 0; f < num_flav; f++) {
      detail::check(qphix_norm2_spinor(gh, rhs[f], &t), "norm2");
      chi_sq += t;
    }
    site_flops += 4 * sf;
    double rsd_sq = RsdTarget * RsdTarget * (rt == RELATIVE ? chi_sq : 1.0);
    M(mp, x, isign);
    M(mmp, mp, -isign);
    mv_apps += 2;
    double cp = 0;
    for (int f = 0; f < num_flav; f++) {
      detail::check(qphix_xmy_norm2(gh, r[f], rhs[f], mmp[f], &t), "xmyNorm2");
      cp += t;
    }
    site_flops += 6 * sf;
    if (cp <= rsd_sq) {
      n_iters = 0;
      rsd_sq_final = cp / chi_sq;
      for (Spinor *s : w) g.free(s);
      return;
    }
    for (int f = 0; f < num_flav; f++) detail::check(qphix_copy_spinor(gh, p[f], r[f]), "copy");
    for (int k = 0; k <= max_iters_; ++k) {
      double c = cp;
      M(mp, p, isign);
      M(mmp, mp, -isign);
      mv_apps += 2;
      double d = norm2(mp), a = c / d;
      site_flops += 4 * sf;
      cp = 0;
      for (int f = 0; f < num_flav; f++) {
        detail::check(qphix_rmammp_norm2_rxpap(gh, r[f], a, mmp[f], x[f], p[f], &t), "rmammp");
        cp += t;
      }
      site_flops += 12 * sf;
      if (verbose) std::printf("CG: iter %d: r = %e target = %e\n", k, cp, rsd_sq);
      if (cp < rsd_sq) {
        n_iters = k;
        rsd_sq_final = cp;
        for (Spinor *s : w) g.free(s);
        return;
      }
      double b = cp / c;
      for (int f = 0; f < num_flav; f++) detail::check(qphix_aypx(gh, b, r[f], p[f]), "aypx");
      site_flops += 4 * sf;
    }
    n_iters = max_iters_;
    rsd_sq_final = cp;
    for (Spinor *s : w) g.free(s);
  }

 private:
  template <int NF = num_flav>
  typename std::enable_if<NF == 1>::type apply(Spinor **o, Spinor *const *i, int is, int cb) const
  {
    this->M_(o[0], i[0], is, cb);
  }
  template <int NF = num_flav>
  typename std::enable_if<NF == 2>::type apply(Spinor **o, Spinor *const *i, int is, int cb) const
  {
    const Spinor *ic[2] = {i[0], i[1]};
    this->M_(o, ic, is, cb);
  }
  int max_iters_;
};

template <typename FT, int veclen, int soalen, bool compress12,
          typename EvenOddLinearOperatorBase = EvenOddLinearOperator<FT, veclen, soalen, compress12>>
class InvBiCGStab : public detail::NativeSolverBase<FT, veclen, soalen, compress12,
                                                    EvenOddLinearOperatorBase::num_flav, EvenOddLinearOperatorBase> {
  typedef detail::NativeSolverBase<FT, veclen, soalen, compress12, EvenOddLinearOperatorBase::num_flav,
                                   EvenOddLinearOperatorBase>
      B;

 public:
  static constexpr int num_flav = EvenOddLinearOperatorBase::num_flav;
  InvBiCGStab(EvenOddLinearOperatorBase &M, int MaxIters) : B(M)
  {
    if (M.native()) detail::check(qphix_bicgstab_create(M.native(), MaxIters, &this->s_), "InvBiCGStab");
  }
};

// InvMR / InvMRSmoother (invmr.h:249-348)
template <typename FT, int V, int S, bool compress12,
          typename EvenOddLinearOperatorBase = EvenOddLinearOperator<FT, V, S, compress12>>
class InvMR : public detail::NativeSolverBase<FT, V, S, compress12, EvenOddLinearOperatorBase::num_flav,
                                              EvenOddLinearOperatorBase> {
  typedef detail::NativeSolverBase<FT, V, S, compress12, EvenOddLinearOperatorBase::num_flav,
                                   EvenOddLinearOperatorBase>
      B;

 public:
  static constexpr int num_flav = EvenOddLinearOperatorBase::num_flav;
  InvMR(EvenOddLinearOperatorBase &M, int MaxIters, double Omega) : B(M)
  {
    if (!M.native()) throw std::runtime_error("InvMR needs a library operator");
    detail::check(qphix_mr_create(M.native(), MaxIters, Omega, 0, &this->s_), "InvMR");
  }
};

template <typename FT, int V, int S, bool compress12,
          typename EvenOddLinearOperatorBase = EvenOddLinearOperator<FT, V, S, compress12>>
class InvMRSmoother : public detail::NativeSolverBase<FT, V, S, compress12, EvenOddLinearOperatorBase::num_flav,
                                                      EvenOddLinearOperatorBase> {
  typedef detail::NativeSolverBase<FT, V, S, compress12, EvenOddLinearOperatorBase::num_flav,
                                   EvenOddLinearOperatorBase>
      B;

 public:
  static constexpr int num_flav = EvenOddLinearOperatorBase::num_flav;
  InvMRSmoother(EvenOddLinearOperatorBase &M, int MaxIters, double Omega) : B(M)
  {
    if (!M.native()) throw std::runtime_error("InvMRSmoother needs a library operator");
    detail::check(qphix_mr_create(M.native(), MaxIters, Omega, 1, &this->s_), "InvMRSmoother");
  }
};

template <typename FT, int V, int S, bool Compress, typename FTInner, int VInner, int SInner, bool CompressInner,
          bool MdagM = false, typename EvenOddLinearOperatorBase = EvenOddLinearOperator<FT, V, S, Compress>>
class InvRichardsonMultiPrec
    : public detail::NativeSolverBase<FT, V, S, Compress, EvenOddLinearOperatorBase::num_flav,
                                      EvenOddLinearOperatorBase> {
  typedef detail::NativeSolverBase<FT, V, S, Compress, EvenOddLinearOperatorBase::num_flav,
                                   EvenOddLinearOperatorBase>
      B;

 public:
  static constexpr int num_flav = EvenOddLinearOperatorBase::num_flav;
  InvRichardsonMultiPrec(EvenOddLinearOperatorBase &m_outer,
                         AbstractSolver<FTInner, VInner, SInner, CompressInner, num_flav> &solver_inner,
                         const double delta, const int max_iters)
      : B(m_outer)
  {
    if (!m_outer.native() || !solver_inner.native())
      throw std::runtime_error("InvRichardsonMultiPrec needs a library operator and a library inner solver");
    detail::check(qphix_richardson_create(m_outer.native(), solver_inner.native(), delta, max_iters, MdagM ? 1 : 0,
                                          &this->s_),
                  "InvRichardsonMultiPrec");
  }
};

template <typename FT, int veclen, int soalen, bool compress12,
          typename EvenOddLinearOperatorBase = EvenOddLinearOperator<FT, veclen, soalen, compress12>>
class MInvCG : public AbstractMultiSolver<FT, veclen, soalen, compress12, EvenOddLinearOperatorBase::num_flav> {
 public:
  static constexpr int num_flav = EvenOddLinearOperatorBase::num_flav;
  typedef typename Geometry<FT, veclen, soalen, compress12>::FourSpinorBlock Spinor;
  MInvCG(EvenOddLinearOperatorBase &M, int MaxIters, int MaxShifts) : M_(M), max_shifts_(MaxShifts)
  {
    if (!M.native()) throw std::runtime_error("MInvCG needs a library operator");
    detail::check(qphix_minvcg_create(M.native(), MaxIters, MaxShifts, &s_), "MInvCG");
  }
  ~MInvCG() { qphix_solver_destroy(s_); }
  // x[shift][flavour], rhs[flavour]
  void operator()(Spinor **const *x, const Spinor *const *rhs, const int n_shift, const double *shifts,
                  const double *RsdTarget, int &n_iters, double *rsd_sq_final, unsigned long &site_flops,
                  unsigned long &mv_apps, int isign, bool verboseP, int cb = 1) const override
  {
    std::vector<void *> xv(n_shift * num_flav);
    for (int s = 0; s < n_shift; s++)
      for (int f = 0; f < num_flav; f++) xv[s * num_flav + f] = x[s][f];
    const void *rv[num_flav];
    for (int f = 0; f < num_flav; f++) rv[f] = rhs[f];
    int st = qphix_minvcg_solve(s_, xv.data(), rv, n_shift, shifts, RsdTarget, isign, verboseP ? 1 : 0, cb,
                                &n_iters, rsd_sq_final, &site_flops, &mv_apps);
    if (st != QPHIX_ENOCONV) detail::check(st, "MInvCG");
  }
  Geometry<FT, veclen, soalen, compress12> &getGeometry() override { return M_.getGeometry(); }

 private:
  EvenOddLinearOperatorBase &M_;
  int max_shifts_;
  qphix_solver *s_ = nullptr;
};

}  // namespace QPhiX
