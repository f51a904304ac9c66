// Drop-in AbstractUnprecSolver (abs_solver.h:271-376) and UnprecSolverWrapper
// (unprec_solver_wrapper.h:16-100) of JeffersonLab/qphix: an unpreconditioned
// solve through an even-odd solver, single flavour. The operator must be a
// library operator (Wilson or clover) and the even-odd solver a library solver.
#pragma once

#include "linear_ops.h"
#include "solvers.h"

namespace QPhiX {

template <typename FT, int V, int S, bool compress12, int num_flav = 1>
class AbstractUnprecSolver {
 public:
  using Spinor = FullSpinor<FT, V, S, compress12>;
  using CBSpinor = typename Spinor::CBSpinor;
  virtual ~AbstractUnprecSolver() {}
  virtual AbstractSolver<FT, V, S, compress12> &getEOSolver() const = 0;
  virtual void SourcePrepare(CBSpinor *const out[num_flav], CBSpinor const *const in_cb[num_flav],
                             CBSpinor const *const in_other_cb[num_flav], int isign, int solve_cb) const = 0;
  virtual void SolutionReconstruct(CBSpinor *const out_other_cb[num_flav], CBSpinor const *const in_other_cb[num_flav],
                                   CBSpinor const *const out_cb[num_flav], int isign, int solve_cb) const = 0;
  void operator()(Spinor *x, const Spinor *rhs, const double RsdTarget, int &niters, double &rsd_sq_final,
                  unsigned long &site_flops, unsigned long &mv_apps, int isign, bool verboseP, int solve_cb = 1,
                  ResiduumType residType = RELATIVE) const
  {
    CBSpinor *out_cb[1] = {x->getCBData(solve_cb)};
    CBSpinor *out_other[1] = {x->getCBData(1 - solve_cb)};
    CBSpinor const *in_cb[1] = {rhs->getCBData(solve_cb)};
    CBSpinor const *in_other[1] = {rhs->getCBData(1 - solve_cb)};
    // the prepared source lives in the other-checkerboard half of x
    SourcePrepare(out_other, in_cb, in_other, isign, solve_cb);
    getEOSolver()(out_cb[0], out_other[0], RsdTarget, niters, rsd_sq_final, site_flops, mv_apps, isign, verboseP,
                  solve_cb, residType);
    CBSpinor const *oc[1] = {out_cb[0]};
    SolutionReconstruct(out_other, in_other, oc, isign, solve_cb);
  }
};

template <typename FT, int V, int S, bool compress12, typename LinearOpType, int num_flav = 1>
class UnprecSolverWrapper : public AbstractUnprecSolver<FT, V, S, compress12, num_flav> {
  static_assert(num_flav == 1, "the unpreconditioned wrapper is single flavour");

 public:
  using Spinor = FullSpinor<FT, V, S, compress12>;
  using CBSpinor = typename Spinor::CBSpinor;
  UnprecSolverWrapper(AbstractSolver<FT, V, S, compress12> &theEOSolver, LinearOpType &M) : eo_(theEOSolver), M_(M)
  {
    if (!M.native()) throw std::runtime_error("UnprecSolverWrapper needs a library operator");
  }
  void SourcePrepare(CBSpinor *const out[num_flav], CBSpinor const *const in_cb[num_flav],
                     CBSpinor const *const in_other_cb[num_flav], int isign, int solve_cb) const override
  {
    void *o[1] = {out[0]};
    const void *a[1] = {in_cb[0]}, *b[1] = {in_other_cb[0]};
    detail::check(qphix_unprec_source_prepare(M_.native(), o, a, b, isign, solve_cb), "SourcePrepare");
  }
  void SolutionReconstruct(CBSpinor *const out_other_cb[num_flav], CBSpinor const *const in_other_cb[num_flav],
                           CBSpinor const *const out_cb[num_flav], int isign, int solve_cb) const override
  {
    void *o[1] = {out_other_cb[0]};
    const void *a[1] = {in_other_cb[0]}, *b[1] = {out_cb[0]};
    detail::check(qphix_unprec_solution_reconstruct(M_.native(), o, a, b, isign, solve_cb), "SolutionReconstruct");
  }
  AbstractSolver<FT, V, S, compress12> &getEOSolver() const override { return eo_; }

 private:
  AbstractSolver<FT, V, S, compress12> &eo_;
  LinearOpType &M_;
};

}  // namespace QPhiX
