// Exercises the drop-in QPhiX C++ API the way a reference caller does
// (tests/testClovDslashFull.cc, tests/testDslashFull.cc of JeffersonLab/qphix),
// compiled with plain g++ against libqphix_amd.so: no HIP headers involved.
// Prints one "check <name> <value> <limit>" line per check and exits non-zero
// on the first failure.
#include <cmath>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

#include "qphix/qphix.h"

using namespace QPhiX;

static int failures = 0;
static void check(const char *name, double v, double lim)
{
  std::printf("check %s %.3e %.1e %s\n", name, v, lim, v <= lim ? "ok" : "FAIL");
  if (!(v <= lim)) failures++;
}

// a user-defined operator: M^+ applied through another operator, so the
// solvers take their host-driven (non-native) path
template <typename FT, int V, int S, bool C>
class Wrapped : public EvenOddLinearOperator<FT, V, S, C> {
 public:
  typedef typename Geometry<FT, V, S, C>::FourSpinorBlock Spinor;
  explicit Wrapped(EvenOddLinearOperator<FT, V, S, C> &M) : M_(M) {}
  void operator()(Spinor *res, const Spinor *in, int isign, int cb = 1) const override { M_(res, in, isign, cb); }
  Geometry<FT, V, S, C> &getGeometry() override { return M_.getGeometry(); }

 private:
  EvenOddLinearOperator<FT, V, S, C> &M_;
};

int main()
{
  const int latt[4] = {8, 8, 8, 8};
  const double mass = 0.1, csw = 1.0, t_bc = -1.0;
  typedef Geometry<double, 8, 4, false> G64;
  typedef Geometry<float, 16, 8, true> G32;
  G64 geom(latt, 1, 1, 1, 1, 1, 1, 0, 1);
  G32 geom32(latt, 1, 1, 1, 1, 1, 1, 0, 1);

  // gauge and clover prepared on the device from seeds
  G64::SU3MatrixBlock *u[2] = {geom.allocCBGauge(), geom.allocCBGauge()};
  detail::check(qphix_gen_gauge(geom.handle(), u[0], u[1], 11, 0.4), "gen_gauge");
  G64::CloverBlock *clov = geom.allocCBClov(), *invclov = geom.allocCBClov();
  detail::check(qphix_clover_build(geom.handle(), u[0], u[1], csw, csw, 4 + mass, 1, QPHIX_CLOV_A, 0, 0, clov), "A");
  detail::check(qphix_clover_build(geom.handle(), u[0], u[1], csw, csw, 4 + mass, 0, QPHIX_CLOV_AINV, 0, 0, invclov),
                "Ainv");
  G32::SU3MatrixBlock *u32[2] = {geom32.allocCBGauge(), geom32.allocCBGauge()};
  G32::CloverBlock *clov32 = geom32.allocCBClov(), *invclov32 = geom32.allocCBClov();
  for (int cb = 0; cb < 2; cb++) qphix_convert_gauge(geom32.handle(), u32[cb], geom.handle(), u[cb]);
  qphix_convert_clover(geom32.handle(), clov32, geom.handle(), clov, 0);
  qphix_convert_clover(geom32.handle(), invclov32, geom.handle(), invclov, 0);

  G64::FourSpinorBlock *b = geom.allocCBFourSpinor(), *x = geom.allocCBFourSpinor();
  G64::FourSpinorBlock *r = geom.allocCBFourSpinor(), *t = geom.allocCBFourSpinor();
  detail::check(qphix_gen_spinor(geom.handle(), b, 5, 1), "gen_spinor");
  double bnorm;
  norm2Spinor(bnorm, b, geom, 1);

  // Dslash: A^-1 D psi equals clovMult(A^-1, D psi)
  {
    ClovDslash<double, 8, 4, false> D(&geom, t_bc, 1.0, 1.0);
    Dslash<double, 8, 4, false> W(&geom, t_bc, 1.0, 1.0);
    D.dslash(r, b, u[0], invclov, 1, 0);
    W.dslash(t, b, u[0], 1, 0);
    D.clovMult(x, t, invclov);
    double n;
    xmyNorm2Spinor(t, r, x, n, geom, 1);
    double rn;
    norm2Spinor(rn, r, geom, 1);
    check("clov_dslash_vs_clovMult", std::sqrt(n / rn), 1e-14);
  }

  EvenOddCloverOperator<double, 8, 4, false> M(u, clov, invclov, &geom, t_bc, 1.0, 1.0);
  // BiCGStab (native path)
  {
    zeroSpinor(x, geom, 1);
    InvBiCGStab<double, 8, 4, false> solver(M, 1000);
    int it;
    double rsd;
    unsigned long fl, mv;
    solver(x, b, 1e-10, it, rsd, fl, mv, 1, false);
    M(t, x, 1);
    double n;
    xmyNorm2Spinor(r, b, t, n, geom, 1);
    check("bicgstab_true_resid", std::sqrt(n / bnorm), 1.5e-10);
    // the reference's counters (invbicgstab.h:158-240): it = iterations + 1
    check("bicgstab_mv_apps", std::fabs((double)mv - (2.0 * it - 1)), 0);
    check("bicgstab_site_flops", std::fabs((double)fl - (72.0 + 960.0 * (it - 1))), 0);
  }
  // CG through a user-defined operator (host-driven loop)
  {
    Wrapped<double, 8, 4, false> Mw(M);
    zeroSpinor(x, geom, 1);
    InvCG<double, 8, 4, false> cg(Mw, 2000);
    int it;
    double rsd;
    unsigned long fl, mv;
    cg(x, b, 1e-10, it, rsd, fl, mv, 1, false);
    M(r, x, 1);
    M(t, r, -1);
    double n;
    xmyNorm2Spinor(r, b, t, n, geom, 1);
    check("generic_cg_true_resid", std::sqrt(n / bnorm), 1.5e-10);
    // invcg.h:180-416: converged in 0-based iteration it
    check("generic_cg_mv_apps", std::fabs((double)mv - (2.0 * it + 4)), 0);
    check("generic_cg_site_flops", std::fabs((double)fl - (312.0 + 240.0 * it)), 0);
    // the native CG on the same operator counts the same way
    InvCG<double, 8, 4, false> ncg(M, 2000);
    int it2;
    unsigned long fl2, mv2;
    zeroSpinor(x, geom, 1);
    ncg(x, b, 1e-10, it2, rsd, fl2, mv2, 1, false);
    check("native_cg_mv_apps", std::fabs((double)mv2 - (2.0 * it2 + 4)), 0);
    check("native_cg_site_flops", std::fabs((double)fl2 - (312.0 + 240.0 * it2)), 0);
  }
  // mixed precision: FP64 Richardson around FP32 BiCGStab
  {
    EvenOddCloverOperator<float, 16, 8, true> M32(u32, clov32, invclov32, &geom32, t_bc, 1.0, 1.0);
    InvBiCGStab<float, 16, 8, true> inner(M32, 1000);
    InvRichardsonMultiPrec<double, 8, 4, false, float, 16, 8, true> rich(M, inner, 1e-4, 100);
    zeroSpinor(x, geom, 1);
    int it;
    double rsd;
    unsigned long fl, mv;
    rich(x, b, 1e-11, it, rsd, fl, mv, 1, false);
    M(t, x, 1);
    double n;
    xmyNorm2Spinor(r, b, t, n, geom, 1);
    check("richardson_f64_f32_true_resid", std::sqrt(n / bnorm), 1.5e-11);
  }
  // multishift CG
  {
    const int ns = 3;
    double shifts[ns] = {0.001, 0.01, 0.1}, tgt[ns] = {1e-10, 1e-10, 1e-10}, fin[ns];
    G64::FourSpinorBlock *xs[ns][1];
    G64::FourSpinorBlock **xp[ns];
    for (int s = 0; s < ns; s++) {
      xs[s][0] = geom.allocCBFourSpinor();
      xp[s] = xs[s];
    }
    const G64::FourSpinorBlock *rhs[1] = {b};
    MInvCG<double, 8, 4, false> mcg(M, 2000, ns);
    int it;
    unsigned long fl, mv;
    mcg(xp, rhs, ns, shifts, tgt, it, fin, fl, mv, 1, false);
    double worst = 0;
    for (int s = 0; s < ns; s++) {
      M(r, xs[s][0], 1);
      M(t, r, -1);
      axpy(shifts[s], xs[s][0], t, geom, 1);
      double n;
      xmyNorm2Spinor(r, b, t, n, geom, 1);
      worst = std::fmax(worst, std::sqrt(n / bnorm));
      geom.free(xs[s][0]);
    }
    check("minvcg_worst_shift_true_resid", worst, 2e-10);
  }
  // field handles, ArithType, full-spinor BLAS, clover_product, InvDummyHermTest
  {
    static_assert(std::is_same<ArithType<half>::Type, float>::value, "half computes in float");
    static_assert(std::is_same<ArithType<double>::Type, double>::value, "double computes in double");
    FourSpinorHandle<double, 8, 4, false> h1(geom);
    auto h2 = makeFourSpinorHandle(geom);
    copySpinor(h1.get(), b, geom, 1);
    double n1;
    norm2Spinor(n1, h1.get(), geom, 1);
    check("handle_copy_norm", std::fabs(n1 - bnorm) / bnorm, 0.0);
    // full spinors: the norm is the sum over checkerboards, ax scales both
    FullSpinor<double, 8, 4, false> fs(geom), ft(geom);
    detail::check(qphix_gen_spinor(geom.handle(), fs.getCBData(0), 31, 0), "gen_spinor");
    copySpinor(fs.getCBData(1), b, geom, 1);
    double n0, nf;
    norm2Spinor(n0, fs.getCBData(0), geom, 1);
    norm2Spinor(nf, fs, geom, 1);
    check("full_norm2_sums_cbs", std::fabs(nf - n0 - bnorm) / nf, 1e-15);
    copySpinor(ft, fs, geom, 1);
    axSpinor(2.0, ft, geom, 1);
    double ip[2];
    innerProductSpinor(ip, fs, ft, geom, 1);
    check("full_inner_product", std::fabs(ip[0] - 2.0 * nf) / nf + std::fabs(ip[1]) / nf, 1e-14);
    double xr;
    xmy2Norm2Spinor(fs, ft, xr, geom, 1);  // ft = 2 fs - fs = fs
    check("full_xmy2norm2", std::fabs(xr - nf) / nf, 1e-14);
    // clover_product accumulates A in into out
    GaugeHandle<double, 8, 4, false> gh(geom);
    CloverHandle<double, 8, 4, false> ch(geom);
    FullCloverHandle<double, 8, 4, false> fh(geom);
    ClovDslash<double, 8, 4, false> D(&geom, t_bc, 1.0, 1.0);
    copySpinor(x, b, geom, 1);
    clover_product<double, 8, 4, false>(x, b, clov, geom);  // x = b + A b
    D.clovMult(t, b, clov);
    ypeqx(b, t, geom, 1);
    double n;
    xmyNorm2Spinor(r, x, t, n, geom, 1);
    check("clover_product_accumulates", std::sqrt(n / bnorm), 1e-14);
    detail::check(qphix_clover_build(geom.handle(), u[0], u[1], csw, csw, 4 + mass, 1, QPHIX_CLOV_TMINV, 0.05, 0,
                                     fh.get()),
                  "full clover");
    zeroSpinor(x, geom, 1);
    clover_product<double, 8, 4, false>(x, b, fh.get(), geom);  // (A + i mu g5)^-1 b
    // applying A + i mu g5 gives b back: A x + i mu g5 x
    D.clovMult(t, x, clov);
    const double pimu[2] = {0.0, 0.05};
    copySpinor(r, x, geom, 1);
    twisted_mass(pimu, x, r, geom, 1);  // r = i mu g5 x
    ypeqx(r, t, geom, 1);
    xmyNorm2Spinor(r, t, b, n, geom, 1);
    check("clover_product_full_block_inverse", std::sqrt(n / bnorm), 1e-13);
    (void)gh; (void)ch; (void)h2;
    // the hermiticity test sees <x|M^+M y> = <Mx|My>
    InvDummyHermTest<double, 8, 4, false> herm(M, 1);
    detail::check(qphix_gen_spinor(geom.handle(), x, 41, 1), "gen_spinor");
    int it;
    double rsd = -1;
    unsigned long fl, mv;
    herm(x, b, 1e-10, it, rsd, fl, mv, 1, false);
    double a1[2], a2[2];
    herm.lastProducts(a1, a2);
    const double sc = std::fabs(a2[0]) + std::fabs(a2[1]);
    check("dummy_hermtest_products_agree", (std::fabs(a1[0] - a2[0]) + std::fabs(a1[1] - a2[1])) / sc, 1e-12);
    check("dummy_hermtest_counts", (it == 1 && mv == 3) ? 0.0 : 1.0, 0.0);
    masterPrintf("peak host memory %.1f MiB\n", get_max_resident_mib());
    check("resident_memory_reported", get_max_resident_mib() > 1.0 ? 0.0 : 1.0, 0.0);
  }
  // errors map to the reference's exceptions
  {
    bool thrown = false;
    try {
      const int bad[4] = {8, 8, 8, 7};
      G64 g2(bad);
    } catch (const std::domain_error &) {
      thrown = true;
    }
    check("odd_extent_throws_domain_error", thrown ? 0.0 : 1.0, 0.0);
  }
  // multigrid-facing pieces and the unpreconditioned solve through the
  // even-odd solver (clover.h:174-251, unprec_solver_wrapper.h, abs_solver.h)
  {
    typedef FullSpinor<double, 8, 4, false> Full;
    Full xf(geom), bf(geom);
    detail::check(qphix_gen_spinor(geom.handle(), bf.getCBData(0), 21, 0), "gen_spinor");
    detail::check(qphix_gen_spinor(geom.handle(), bf.getCBData(1), 22, 1), "gen_spinor");
    zeroSpinor(xf.getCBData(0), geom, 1);
    zeroSpinor(xf.getCBData(1), geom, 1);
    InvBiCGStab<double, 8, 4, false> eo(M, 1000);
    UnprecSolverWrapper<double, 8, 4, false, EvenOddCloverOperator<double, 8, 4, false>> us(eo, M);
    int it;
    double rsd;
    unsigned long fl, mv;
    us(&xf, &bf, 1e-11, it, rsd, fl, mv, 1, false);
    // odd rows: A_oo x_o - 1/2 D x_e = b_o
    ClovDslash<double, 8, 4, false> D(&geom, t_bc, 1.0, 1.0);
    D.clovMult(r, xf.getCBData(1), clov);
    M.M_offdiag(t, xf.getCBData(0), 1, 1);
    ypeqx(t, r, geom, 1);
    double n, bo;
    xmyNorm2Spinor(t, bf.getCBData(1), r, n, geom, 1);
    norm2Spinor(bo, bf.getCBData(1), geom, 1);
    check("unprec_odd_rows", std::sqrt(n / bo), 1e-9);
    // even rows: x_e = A_ee^-1 (b_e + 1/2 D x_o)
    M.M_offdiag(t, xf.getCBData(1), 1, 0);
    xmyNorm2Spinor(r, bf.getCBData(0), t, n, geom, 1);
    M.M_diag_inv(x, r, 1, 0);
    double xe;
    xmyNorm2Spinor(r, x, xf.getCBData(0), n, geom, 1);
    norm2Spinor(xe, xf.getCBData(0), geom, 1);
    check("unprec_even_rows", std::sqrt(n / xe), 1e-13);
  }
  // InvMR on the clover operator (invmr.h)
  {
    zeroSpinor(x, geom, 1);
    InvMR<double, 8, 4, false> mr(M, 5000, 1.0);
    int it;
    double rsd;
    unsigned long fl, mv;
    mr(x, b, 1e-9, it, rsd, fl, mv, 1, false);
    M(t, x, 1);
    double n;
    xmyNorm2Spinor(r, b, t, n, geom, 1);
    check("mr_true_resid", std::sqrt(n / bnorm), 1.5e-9);
  }

  // the Chroma-facing system solver from an <InvertParam> group on unit links:
  // a constant odd-site source is an eigenvector of the even-odd operator with
  // eigenvalue m (8 + m) / (4 + m)
  {
    const long vh = 8 * 8 * 8 * 8 / 2;
    std::vector<double> unit(2 * vh * 4 * 18, 0.0), chi(2 * vh * 24, 0.0), psi(2 * vh * 24, 0.0);
    for (long i = 0; i < 2 * vh * 4; i++)
      for (int c = 0; c < 3; c++) unit[i * 18 + c * 8] = 1.0;
    for (long i = vh * 24; i < 2 * vh * 24; i++) chi[i] = 0.25 + 0.01 * (i % 24);
    const std::string xml =
        "<InvertParam><invType>QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER</invType>"
        "<SolverType>BICGSTAB</SolverType><MaxIter>100</MaxIter><RsdTarget>1e-10</RsdTarget>"
        "<Delta>0.1</Delta><CloverParams><Mass>0.1</Mass><clovCoeff>1.0</clovCoeff></CloverParams>"
        "<NCores>8</NCores><Tune>true</Tune></InvertParam>";
    CloverLinOpSysSolver solve(latt, xml, unit.data());
    SystemSolverResults res = solve(psi.data(), chi.data());
    const double lam = 0.1 * 8.1 / 4.1;
    double err = 0, mx = 0;
    for (long i = vh * 24; i < 2 * vh * 24; i++) {
      err = std::fmax(err, std::fabs(psi[i] - chi[i] / lam));
      mx = std::fmax(mx, std::fabs(chi[i] / lam));
    }
    check("syssolver_free_field", err / mx, 1e-9);
    check("syssolver_rel_resid", res.rel_resid, 5e-10);
    // a solve that cannot reach its target throws, as the reference's caller aborts
    qphix_invert_param p;
    detail::check(qphix_invert_param_read_xml(xml.c_str(), &p), "InvertParam");
    p.inv_type = QPHIX_INV_CLOVER;
    p.max_iter = 1;
    p.rsd_target = 1e-14;
    CloverLinOpSysSolver fails(latt, p, unit.data());
    for (long i = vh * 24; i < 2 * vh * 24; i++) chi[i] = (i * 7919 % 101) * 0.01;
    std::fill(psi.begin(), psi.end(), 0.0);
    bool threw = false;
    try {
      fails(psi.data(), chi.data());
    } catch (const std::runtime_error &) {
      threw = true;
    }
    check("syssolver_failed_solve_throws", threw ? 0 : 1, 0);
  }

  for (int cb = 0; cb < 2; cb++) {
    geom.free(u[cb]);
    geom32.free(u32[cb]);
  }
  geom.free(clov);
  geom.free(invclov);
  geom32.free(clov32);
  geom32.free(invclov32);
  geom.free(b);
  geom.free(x);
  geom.free(r);
  geom.free(t);
  std::printf("%s\n", failures ? "DROPIN FAIL" : "DROPIN OK");
  return failures ? 1 : 0;
}
