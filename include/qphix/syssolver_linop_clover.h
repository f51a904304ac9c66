// The QPhiX clover LinOpSysSolver as Chroma drives it (README.markdown:163-256
// of JeffersonLab/qphix): an <InvertParam> group selects
// QPHIX_CLOVER_INVERTER or QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER, the
// solver is built once per gauge field and then solves M psi = chi on the odd
// checkerboard for each source, with the true-residual check against
// RsdTarget * RsdToleranceFactor that ends a failing solve.
//
// Fields are canonical host arrays (layouts in qphix_amd.h): the gauge field
// U_mu(x) for the full lattice, psi and chi as full-lattice spinors of which
// the odd half is used.
#pragma once

#include <stdexcept>
#include <string>

#include "geometry.h"

namespace QPhiX {

struct SystemSolverResults {
  int n_count = 0;          // solver iterations
  double resid = 0;         // |chi - M psi| on the odd checkerboard
  double rel_resid = 0;     // resid / |chi|
  unsigned long mv_apps = 0;
  double seconds = 0;
};

class CloverLinOpSysSolver {
 public:
  // invert_param_xml: a document holding an <InvertParam> group
  CloverLinOpSysSolver(const int latt[4], const std::string &invert_param_xml, const double *u)
  {
    qphix_invert_param p;
    detail::check(qphix_invert_param_read_xml(invert_param_xml.c_str(), &p), "InvertParam");
    create(latt, p, u);
  }
  CloverLinOpSysSolver(const int latt[4], const qphix_invert_param &p, const double *u) { create(latt, p, u); }
  ~CloverLinOpSysSolver() { qphix_syssolver_destroy(s_); }
  CloverLinOpSysSolver(const CloverLinOpSysSolver &) = delete;
  CloverLinOpSysSolver &operator=(const CloverLinOpSysSolver &) = delete;

  // solves from the initial guess in psi's odd half; throws std::runtime_error
  // when the true residual misses RsdTarget * RsdToleranceFactor (the
  // reference's caller aborts there)
  SystemSolverResults operator()(double *psi, const double *chi) const
  {
    qphix_syssolver_results r;
    detail::check(qphix_syssolver_solve(s_, psi, chi, &r), "CloverLinOpSysSolver");
    SystemSolverResults out;
    out.n_count = r.n_count;
    out.resid = r.resid;
    out.rel_resid = r.rel_resid;
    out.mv_apps = r.mv_apps;
    out.seconds = r.seconds;
    if (r.status != QPHIX_OK)
      throw std::runtime_error("CloverLinOpSysSolver: solve failed (" + std::string(qphix_strerror(r.status)) +
                               "), |r|/|b| = " + std::to_string(r.rel_resid));
    return out;
  }

 private:
  void create(const int latt[4], const qphix_invert_param &p, const double *u)
  {
    detail::check(qphix_syssolver_create(latt, &p, u, &s_), "CloverLinOpSysSolver");
  }
  qphix_syssolver *s_ = nullptr;
};

}  // namespace QPhiX
