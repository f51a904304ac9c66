// Host-side declarations shared by the library translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "common.hpp"
#include <stdint.h>

namespace qphix_amd {

// lattice + precision descriptor of a geometry
struct GeomDesc {
  int L[4];
  int prec;   // Prec
  int c12;    // two-row gauge compression
  long vh;    // sites per checkerboard
  size_t vs;  // plane stride in sites
};

hipError_t fld_pack_spinor(const GeomDesc &g, void *dst, const double *dsrc, hipStream_t st);
hipError_t fld_unpack_spinor(const GeomDesc &g, double *ddst, const void *src, hipStream_t st);
hipError_t fld_gen_spinor(const GeomDesc &g, void *dst, uint64_t seed, int cb, hipStream_t st);
hipError_t fld_convert_spinor(const GeomDesc &go, void *out, double scale, const GeomDesc &gi, const void *in,
                              hipStream_t st);
hipError_t fld_pack_gauge_cb(const GeomDesc &g, void *dst, int cb, const double *dsrc, hipStream_t st);
hipError_t fld_gen_gauge_canonical(const GeomDesc &g, double *du, uint64_t seed, double eps, hipStream_t st);
hipError_t fld_convert_gauge(const GeomDesc &go, void *out, const GeomDesc &gi, const void *in, hipStream_t st);
hipError_t fld_pack_clover(const GeomDesc &g, void *dst, const double *dsrc, hipStream_t st);
hipError_t fld_unpack_clover(const GeomDesc &g, double *ddst, const void *src, hipStream_t st);
hipError_t fld_pack_full_clover(const GeomDesc &g, void *dst, const double *dsrc, hipStream_t st);
hipError_t fld_convert_clover(const GeomDesc &go, void *out, const GeomDesc &gi, const void *in, int full,
                              hipStream_t st);
hipError_t fld_clover_build(const GeomDesc &g, const void *u0, const void *u1, double cr, double ct, double dm,
                            int cb, int what, double mu, double eps, void *out, hipStream_t st);
// allocation-free clover build: FP64 scratch for kNoAllocSites sites in a
// static device array (about 63 MiB per device); fld_noalloc_init looks its
// address up on the current device (once per device, at library set-up)
constexpr long kNoAllocSites = 1l << 15;
constexpr int kNoAllocMaxDevices = 64;
hipError_t fld_noalloc_init(int device);
hipError_t fld_clover_build_noalloc(const GeomDesc &g, const void *u0, const void *u1, double cr, double ct,
                                    double dm, int cb, int what, double mu, double eps, void *out, int device,
                                    hipStream_t st);

// BLAS operations (one spinor per operand; see blas.hip)
enum BlasOp {
  B_COPY = 0, B_ZERO, B_AXY, B_AX, B_YPEQX, B_YMEQX, B_AYPX, B_AXPY, B_AXPBY,
  B_NORM2, B_AXPYNORM2, B_XMYNORM2, B_XMY2NORM2, B_RMAMMP, B_RICH, B_BICG_XMY,
  B_CDOT, B_CDOTNORM, B_PUPD, B_SUPD, B_CAXPY, B_RXUPD, B_MRUPD, B_TWM, B_TWM2,
  B_CLOVMULT, B_FULLMULT, B_TWOFLAV_INVCLOV, B_RXUPD_CDOT, NUM_BLAS_OPS
};
struct BlasArgs {
  const void *x, *y, *z, *w;  // read-only operands
  void *o0, *o1, *o2, *o3;    // written (and possibly read) operands
  double a[6];                // scalars
  // non-null: the scalars are read from device memory, dsc[0..5], when the
  // kernel starts, and the launch does nothing if dsc[7] != 0 (a solver
  // whose device state says it is done)
  const double *dsc;
};
int blas_nred(int op);
// launches op over the geometry; writes nred partial sums per block to
// partials and returns the number of blocks used
hipError_t blas_launch(const GeomDesc &g, int op, const BlasArgs &a, double *partials, int *nblocks, hipStream_t st);
hipError_t blas_finalize(const double *partials, int nblocks, int nred, double *out, hipStream_t st);
constexpr int kMaxBlasBlocks = 2048;

// Multishift CG vector updates for every shift in one sweep (minvcg.h:298-363):
//   x_s -= beta_s p_s      for s in mask_x   (uses p_s before its update)
//   p0   = r + a p0        if do_p0
//   p_s  = zeta_s r + as_s p_s   for s in mask_p
constexpr int kMaxShifts = 32;
struct MShiftArgs {
  const void *r;
  void *p0;
  void *p[kMaxShifts];
  void *x[kMaxShifts];
  double a;
  double zeta[kMaxShifts], as[kMaxShifts], beta[kMaxShifts];
  unsigned mask_x, mask_p;
  int do_p0, nshift;
};
hipError_t mshift_launch(const GeomDesc &g, const MShiftArgs &m, hipStream_t st);

// What a device-resident solver loop publishes at the end of every iteration
// into view[posts & 1] of a pinned host array: the host reads the state of a
// finished iteration without a copy command (progress: the residual over its
// target, below 1 once converged)
struct SolverView {
  int done, pad;
  double progress;
};

// Multishift CG state in device memory (minvcg.h:279-384): the FP64 scalars
// and per-shift recurrences of one solve, updated by single-thread kernels
// between the vector sweeps, so that an iteration needs no host round trip
// and can be captured in a HIP graph. cf[f] receives per-flavour reductions.
struct MsState {
  double a, b, bp, c, cp, d;
  double cf[2];
  double shifts[kMaxShifts], zeta[kMaxShifts], zeta_prev[kMaxShifts], beta[kMaxShifts], as[kMaxShifts];
  double rsd_sq[kMaxShifts], rsd_sq_final[kMaxShifts];
  int conv[kMaxShifts];
  unsigned mask_x, mask_p;
  int do_p0, nshift, nf, iters, max_iters, done, converged;
  unsigned long site_flops;
  // largest c zeta_s^2 / target_s over the shifts tested by the last
  // iteration (below 1 for a shift that converged there); lets the host
  // foresee the last iteration
  double progress;
  SolverView *hview;
  int posts;
};
// the vector sweep of mshift_launch with a, zeta, as, beta, the masks and
// do_p0 read from the device state
hipError_t mshift_dev_launch(const GeomDesc &g, const MShiftArgs &m, const MsState *st_dev, hipStream_t st);
// start of the first iteration: a = c / cp, as_s, mask_p, cp = c, bp = b
hipError_t ms_pre_launch(MsState *st_dev, hipStream_t st);
// b = -cp / d, d the sum of cf over flavours
hipError_t ms_mid_launch(MsState *st_dev, hipStream_t st);
// end of an iteration: c from cf, zeta / beta recurrences, convergence, then
// the start of the next iteration as ms_pre_launch (or, once the solve is
// done, only the pending x_s updates for the next sweep)
hipError_t ms_post_launch(MsState *st_dev, hipStream_t st);
// BiCGStab state in device memory (invbicgstab.h:150-256): the complex
// scalars, the reductions the device loop writes (<r0, v>, <t, r> with |t|^2,
// |r|^2 with <r0, r>), and the scalar blocks of the three vector updates
// (BlasArgs::dsc; [7] = skip once done). Updated by single-thread kernels,
// so one iteration needs no host round trip.
struct BiState {
  double rho[2], rho_prev[2], alpha[2], omega[2];
  double r0v[2], tr[3], ru[3];
  double r_norm, rsd_sq;
  double pupd[8], supd[8], rxupd[8];
  unsigned long site_flops, mv_apps, sf;  // sf: the reference's unit, 12 x flavours
  int iters, max_iters, done, status, n_iters, posts;
  double progress;
  SolverView *hview;
};
// start of iteration k: rho = <r0, r>, the breakdown test, beta, p-update scalars
hipError_t bi_pre_launch(BiState *st_dev, hipStream_t st);
// after v = M p and <r0, v>: the breakdown test, alpha, s-update scalars
hipError_t bi_mid_launch(BiState *st_dev, hipStream_t st);
// after t = M r and <t, r>, |t|^2: the breakdown test, omega, x/r-update scalars
hipError_t bi_mid2_launch(BiState *st_dev, hipStream_t st);
// after the x/r update: |r|^2, convergence, the iteration count
hipError_t bi_post_launch(BiState *st_dev, hipStream_t st);

// CG state in device memory (invcg.h:88-456): c = |r|^2 of the iteration,
// the reductions |M p|^2 and |r|^2, the scalar blocks of the two vector
// updates (BlasArgs::dsc; [7] = skip once done), the counters
struct CgDState {
  double c, d_raw, r_raw, rsd_sq, cp;
  double rmammp[8], aypx[8];
  unsigned long site_flops, mv_apps, sf;
  int iters, max_iters, done, status, n_iters, posts;
  double rsd_sq_final, progress;
  SolverView *hview;
};
// after M^+M p and |M p|^2: a = c / d
hipError_t cgd_mid_launch(CgDState *st_dev, hipStream_t st);
// after r -= a M^+M p, x += a p, |r|^2: convergence, b = |r|^2 / c
hipError_t cgd_post_launch(CgDState *st_dev, hipStream_t st);

// y += (*b) x and |y|^2, the scalar read from device memory; partials as
// blas_launch, the reduction finalized into *out
hipError_t axpy_norm2_dev_launch(const GeomDesc &g, const double *b, const void *x, void *y, double *partials,
                                 double *out, hipStream_t st);

// one BLAS operation on n right-hand sides in one launch, each with its own
// operands and first scalar; reductions come back per right-hand side
struct MultiBlasArgs {
  const void *x[kMaxRhs];
  const void *y[kMaxRhs];
  void *o0[kMaxRhs];
  void *o1[kMaxRhs];
  double a[kMaxRhs];
  int n;
};
// partials must hold kMaxBlasBlocks * 3 doubles; out n * nred doubles
hipError_t blas_multi_launch(const GeomDesc &g, int op, const MultiBlasArgs &m, double *partials, double *out,
                             hipStream_t st);

// streaming calibration kernel (words per site in / out: 30, 48, 60, 66 / 6; 120 / 12)
hipError_t stream_launch(const GeomDesc &g, int nwi, int nwo, const void *in, void *out, hipStream_t st);

// records msg as the thread's last error (qphix_last_error) and returns code
int api_fail(int code, const char *msg);

}  // namespace qphix_amd
