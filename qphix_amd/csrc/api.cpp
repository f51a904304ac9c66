// C ABI of the gfx950 QPhiX hot path: geometry, fields, Dslash objects,
// even-odd operators, Krylov solvers and the fused BLAS. All work is ordered on
// one HIP stream; calls that return host scalars synchronise it.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <complex>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "../../include/qphix_amd.h"
#include "batch.hpp"
#include "blockblas.hpp"
#include "fields.hpp"
#include "hop_impl.hpp"

using namespace qphix_amd;

namespace qphix_amd {
template <int P>
hipError_t launch_hop_prec(int kind, bool c12, int isign, const HopArgs &a, int block, hipStream_t st);
extern template hipError_t launch_hop_prec<P64>(int, bool, int, const HopArgs &, int, hipStream_t);
extern template hipError_t launch_hop_prec<P32>(int, bool, int, const HopArgs &, int, hipStream_t);
extern template hipError_t launch_hop_prec<P16>(int, bool, int, const HopArgs &, int, hipStream_t);
extern template hipError_t launch_hop_prec<PB16>(int, bool, int, const HopArgs &, int, hipStream_t);
template <int P>
hipError_t launch_hop_batch_prec(int kind, bool c12, int isign, const HopSeg *segs, const int *off, int nseg,
                                 const int *active, int grid, hipStream_t st);
extern template hipError_t launch_hop_batch_prec<P64>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                      int, hipStream_t);
extern template hipError_t launch_hop_batch_prec<P32>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                      int, hipStream_t);
extern template hipError_t launch_hop_batch_prec<P16>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                      int, hipStream_t);
extern template hipError_t launch_hop_batch_prec<PB16>(int, bool, int, const HopSeg *, const int *, int, const int *,
                                                       int, hipStream_t);
}  // namespace qphix_amd

struct qphix_geom {
  GeomDesc d;
  size_t bytes[4];
};
struct qphix_dslash {
  const qphix_geom *g;
  qphix_dslash_params p;
};
struct qphix_op {
  const qphix_geom *g;
  qphix_op_params p;
  qphix_dslash d;
  int nflav;
  void *tmp[2];
  void *tmp2[2];
  std::vector<void *> mtmp;  // per-right-hand-side temporaries of qphix_op_apply_multi
  std::vector<void *> utmp;  // temporaries of the even-odd source preparation / reconstruction
};
enum { S_CG = 0, S_BICGSTAB = 1, S_RICHARDSON = 2, S_MINVCG = 3, S_BLOCKCG = 4, S_MR = 5 };
struct qphix_solver {
  int kind;
  qphix_op *M;
  int max_iters;
  std::vector<void *> work;  // solver-owned spinors
  // Richardson
  qphix_solver *inner;
  double delta;
  int mdagm;
  // MInvCG
  int max_shifts;
  // MR: over-relaxation and whether to stop on the residual (InvMR) or run
  // max_iters iterations (InvMRSmoother)
  double omega = 1.0;
  int mr_terminate = 1;
  // multi-right-hand-side CG work spinors (grown on demand)
  std::vector<void *> mwork;
  // set by a caller that passes x = 0 (the defect-correction loop's inner
  // solve): the initial residual is then the right-hand side itself and the
  // operator is not applied to the zero vector
  int x_zero = 0;
  // MInvCG control: 0 host scalars (one host round trip per iteration), 1
  // FP64 scalars in device memory with eager launches, 2 the same with each
  // iteration replayed from a HIP graph
  int ms_mode = 2;
  // BiCGStab control: 0 host scalars, 1 device scalars (eager launches), 2
  // device scalars with the iteration replayed from a HIP graph
  int dev_mode = 0;
};

// ---------------------------------------------------------------------------
// state and errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
static bool g_init = false;
static int g_device = 0;
static hipStream_t g_own_stream = nullptr;
static hipStream_t g_stream = nullptr;
static double *g_partials = nullptr;  // device, kMaxBlasBlocks * 3
static double *g_dres = nullptr;      // device, kResults
static double *g_hres = nullptr;      // pinned host, kResults
static constexpr int kResults = 4 * kMaxRhs;
static hipEvent_t g_ev0, g_ev1;
static int g_hop_block = 256;

static int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}
// the error setter of this file for the other host translation units
int qphix_amd::api_fail(int code, const char *msg) { return fail(code, "%s", msg); }

#define HIPCHK(call)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (call);                                                                       \
    if (e_ != hipSuccess) return fail(QPHIX_EHIP, "%s: %s", #call, hipGetErrorString(e_));        \
  } while (0)
#define ENSURE_INIT()                 \
  do {                                \
    if (!g_init) {                    \
      int r_ = qphix_init(g_device);  \
      if (r_) return r_;              \
    }                                 \
  } while (0)
#define CHKPTR(p) \
  do {            \
    if (!(p)) return fail(QPHIX_EINVAL, "%s: null argument '%s'", __func__, #p); \
  } while (0)
#define TRY(x)            \
  do {                    \
    int r_ = (x);         \
    if (r_) return r_;    \
  } while (0)

extern "C" const char *qphix_strerror(int s) {
  switch (s) {
    case QPHIX_OK: return "ok";
    case QPHIX_EDOMAIN: return "argument outside the supported domain";
    case QPHIX_ENOMEM: return "device allocation failed";
    case QPHIX_EHIP: return "HIP runtime error";
    case QPHIX_EBREAKDOWN: return "solver breakdown";
    case QPHIX_ENOCONV: return "solver did not converge";
    case QPHIX_EINVAL: return "invalid argument";
    case QPHIX_ENODEV: return "no HIP device";
  }
  return "unknown status";
}
extern "C" const char *qphix_last_error(void) { return g_err.c_str(); }

extern "C" int qphix_init(int device) {
  if (g_init) return QPHIX_OK;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(QPHIX_ENODEV, "no HIP device visible");
  if (device < 0 || device >= n) return fail(QPHIX_EDOMAIN, "device %d out of range (%d devices)", device, n);
  HIPCHK(hipSetDevice(device));
  g_device = device;
  HIPCHK(hipStreamCreate(&g_own_stream));
  g_stream = g_own_stream;
  HIPCHK(hipMalloc(&g_partials, sizeof(double) * kMaxBlasBlocks * 3));
  HIPCHK(hipMalloc(&g_dres, sizeof(double) * kResults));
  HIPCHK(hipHostMalloc(&g_hres, sizeof(double) * kResults, hipHostMallocDefault));
  HIPCHK(hipEventCreate(&g_ev0));
  HIPCHK(hipEventCreate(&g_ev1));
  // the device address of the allocation-free clover build's static scratch
  HIPCHK(fld_noalloc_init(device));
  g_init = true;
  return QPHIX_OK;
}

extern "C" int qphix_set_stream(void *s) {
  ENSURE_INIT();
  g_stream = s ? (hipStream_t)s : g_own_stream;
  return QPHIX_OK;
}
extern "C" int qphix_sync(void) {
  ENSURE_INIT();
  HIPCHK(hipStreamSynchronize(g_stream));
  HIPCHK(hipGetLastError());
  return QPHIX_OK;
}
extern "C" int qphix_timer_start(void) {
  ENSURE_INIT();
  HIPCHK(hipEventRecord(g_ev0, g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_timer_stop(double *ms) {
  ENSURE_INIT();
  CHKPTR(ms);
  HIPCHK(hipEventRecord(g_ev1, g_stream));
  HIPCHK(hipEventSynchronize(g_ev1));
  float f = 0;
  HIPCHK(hipEventElapsedTime(&f, g_ev0, g_ev1));
  *ms = f;
  return QPHIX_OK;
}
// hop kernel work-group size (multiple of 64, at most 256)
extern "C" int qphix_set_hop_block(int b) {
  if (b <= 0 || b > 256 || (b & 63)) return fail(QPHIX_EDOMAIN, "hop block %d must be 64, 128, 192 or 256", b);
  g_hop_block = b;
  return QPHIX_OK;
}

// ---------------------------------------------------------------------------
// geometry and allocation
// ---------------------------------------------------------------------------
extern "C" int qphix_geom_create(const int latt[4], int prec, int c12, qphix_geom **out) {
  CHKPTR(latt);
  CHKPTR(out);
  for (int k = 0; k < 4; k++)
    if (latt[k] < 2 || (latt[k] & 1))
      return fail(QPHIX_EDOMAIN, "lattice extent %d in dimension %d must be even and >= 2", latt[k], k);
  if (prec != QPHIX_F64 && prec != QPHIX_F32 && prec != QPHIX_F16 && prec != QPHIX_BF16)
    return fail(QPHIX_EDOMAIN, "unknown precision %d", prec);
  if ((prec == QPHIX_F16 || prec == QPHIX_BF16) && !c12)
    return fail(QPHIX_EDOMAIN, "F16 and BF16 geometries require two-row gauge compression");
  // the kernels index sites of one parity (padded to whole 256-site blocks)
  // and their neighbours in 32 bits
  const long vh = (long)latt[0] * latt[1] * latt[2] * latt[3] / 2;
  if (vh > 0x7fffff00L) return fail(QPHIX_EDOMAIN, "a parity of %ld sites exceeds the 32-bit site index", vh);
  qphix_geom *g = new qphix_geom;
  for (int k = 0; k < 4; k++) g->d.L[k] = latt[k];
  g->d.prec = prec;
  g->d.c12 = c12 ? 1 : 0;
  g->d.vh = vh;
  {
    const long r = kSiteBlock > 256 ? kSiteBlock : 256;  // whole layout blocks and work-groups
    g->d.vs = (size_t)((g->d.vh + r - 1) / r) * r;
  }
  const int nr[4] = {kSpinorReals, gauge_reals(c12 != 0), kCloverReals, kFullCloverReals};
  for (int k = 0; k < 4; k++) {
    switch (prec) {
      case QPHIX_F64: g->bytes[k] = field_bytes<P64>(nr[k], g->d.vs); break;
      case QPHIX_F32: g->bytes[k] = field_bytes<P32>(nr[k], g->d.vs); break;
      case QPHIX_F16: g->bytes[k] = field_bytes<P16>(nr[k], g->d.vs); break;
      default: g->bytes[k] = field_bytes<PB16>(nr[k], g->d.vs); break;
    }
  }
  *out = g;
  return QPHIX_OK;
}
extern "C" int qphix_geom_destroy(qphix_geom *g) {
  delete g;
  return QPHIX_OK;
}
extern "C" int qphix_geom_info(const qphix_geom *g, int latt[4], int *prec, int *c12, long *vh, size_t bytes[4]) {
  CHKPTR(g);
  if (latt)
    for (int k = 0; k < 4; k++) latt[k] = g->d.L[k];
  if (prec) *prec = g->d.prec;
  if (c12) *c12 = g->d.c12;
  if (vh) *vh = g->d.vh;
  if (bytes)
    for (int k = 0; k < 4; k++) bytes[k] = g->bytes[k];
  return QPHIX_OK;
}

static int alloc_zeroed(size_t n, void **p) {
  ENSURE_INIT();
  CHKPTR(p);
  void *q = nullptr;
  if (hipMalloc(&q, n) != hipSuccess) {
    (void)hipGetLastError();
    return fail(QPHIX_ENOMEM, "hipMalloc of %zu bytes failed", n);
  }
  HIPCHK(hipMemsetAsync(q, 0, n, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  *p = q;
  return QPHIX_OK;
}
extern "C" int qphix_alloc_spinor(const qphix_geom *g, void **p) { CHKPTR(g); return alloc_zeroed(g->bytes[0], p); }
extern "C" int qphix_alloc_gauge(const qphix_geom *g, void **p) { CHKPTR(g); return alloc_zeroed(g->bytes[1], p); }
extern "C" int qphix_alloc_clover(const qphix_geom *g, void **p) { CHKPTR(g); return alloc_zeroed(g->bytes[2], p); }
extern "C" int qphix_alloc_full_clover(const qphix_geom *g, void **p) { CHKPTR(g); return alloc_zeroed(g->bytes[3], p); }
extern "C" int qphix_free(void *p) {
  if (!p) return QPHIX_OK;
  ENSURE_INIT();
  HIPCHK(hipStreamSynchronize(g_stream));
  HIPCHK(hipFree(p));
  return QPHIX_OK;
}

// ---------------------------------------------------------------------------
// packers, generation, conversion, clover build
// ---------------------------------------------------------------------------
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

static int stage_in(DevBuf &b, const double *src, size_t n) {
  if (hipMalloc(&b.p, n * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(QPHIX_ENOMEM, "staging allocation of %zu doubles failed", n);
  }
  HIPCHK(hipMemcpyAsync(b.p, src, n * sizeof(double), hipMemcpyHostToDevice, g_stream));
  return QPHIX_OK;
}
static int check_cb(int cb) { return (cb == 0 || cb == 1) ? QPHIX_OK : fail(QPHIX_EDOMAIN, "checkerboard %d", cb); }

extern "C" int qphix_pack_spinor(const qphix_geom *g, void *dst, const double *src, int cb) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(dst); CHKPTR(src);
  TRY(check_cb(cb));
  DevBuf b;
  TRY(stage_in(b, src, (size_t)g->d.vh * kSpinorReals));
  HIPCHK(fld_pack_spinor(g->d, dst, (const double *)b.p, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_unpack_spinor(const qphix_geom *g, double *dst, const void *src, int cb) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(dst); CHKPTR(src);
  TRY(check_cb(cb));
  DevBuf b;
  const size_t n = (size_t)g->d.vh * kSpinorReals;
  if (hipMalloc(&b.p, n * sizeof(double)) != hipSuccess) return fail(QPHIX_ENOMEM, "staging allocation failed");
  HIPCHK(fld_unpack_spinor(g->d, (double *)b.p, src, g_stream));
  HIPCHK(hipMemcpyAsync(dst, b.p, n * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_pack_gauge(const qphix_geom *g, void *u0, void *u1, const double *src) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(u0); CHKPTR(u1); CHKPTR(src);
  DevBuf b;
  TRY(stage_in(b, src, (size_t)g->d.vh * 2 * 72));
  HIPCHK(fld_pack_gauge_cb(g->d, u0, 0, (const double *)b.p, g_stream));
  HIPCHK(fld_pack_gauge_cb(g->d, u1, 1, (const double *)b.p, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_pack_clover(const qphix_geom *g, void *dst, const double *src, int cb) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(dst); CHKPTR(src);
  TRY(check_cb(cb));
  DevBuf b;
  TRY(stage_in(b, src, (size_t)g->d.vh * kCloverReals));
  HIPCHK(fld_pack_clover(g->d, dst, (const double *)b.p, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_unpack_clover(const qphix_geom *g, double *dst, const void *src, int cb) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(dst); CHKPTR(src);
  TRY(check_cb(cb));
  DevBuf b;
  const size_t n = (size_t)g->d.vh * kCloverReals;
  if (hipMalloc(&b.p, n * sizeof(double)) != hipSuccess) return fail(QPHIX_ENOMEM, "staging allocation failed");
  HIPCHK(fld_unpack_clover(g->d, (double *)b.p, src, g_stream));
  HIPCHK(hipMemcpyAsync(dst, b.p, n * sizeof(double), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_pack_full_clover(const qphix_geom *g, void *dst, const double *src, int cb) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(dst); CHKPTR(src);
  TRY(check_cb(cb));
  DevBuf b;
  TRY(stage_in(b, src, (size_t)g->d.vh * kFullCloverReals));
  HIPCHK(fld_pack_full_clover(g->d, dst, (const double *)b.p, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}

static int same_lattice(const qphix_geom *a, const qphix_geom *b) {
  for (int k = 0; k < 4; k++)
    if (a->d.L[k] != b->d.L[k]) return fail(QPHIX_EDOMAIN, "geometries describe different lattices");
  return QPHIX_OK;
}
extern "C" int qphix_convert_spinor(const qphix_geom *go, void *out, double scale, const qphix_geom *gi,
                                    const void *in) {
  ENSURE_INIT();
  CHKPTR(go); CHKPTR(gi); CHKPTR(out); CHKPTR(in);
  TRY(same_lattice(go, gi));
  HIPCHK(fld_convert_spinor(go->d, out, scale, gi->d, in, g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_convert_gauge(const qphix_geom *go, void *out, const qphix_geom *gi, const void *in) {
  ENSURE_INIT();
  CHKPTR(go); CHKPTR(gi); CHKPTR(out); CHKPTR(in);
  TRY(same_lattice(go, gi));
  HIPCHK(fld_convert_gauge(go->d, out, gi->d, in, g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_convert_clover(const qphix_geom *go, void *out, const qphix_geom *gi, const void *in, int full) {
  ENSURE_INIT();
  CHKPTR(go); CHKPTR(gi); CHKPTR(out); CHKPTR(in);
  TRY(same_lattice(go, gi));
  HIPCHK(fld_convert_clover(go->d, out, gi->d, in, full, g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_gen_gauge(const qphix_geom *g, void *u0, void *u1, uint64_t seed, double eps) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(u0); CHKPTR(u1);
  DevBuf b;
  const size_t n = (size_t)g->d.vh * 2 * 72;
  if (hipMalloc(&b.p, n * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return fail(QPHIX_ENOMEM, "gauge staging allocation failed");
  }
  HIPCHK(fld_gen_gauge_canonical(g->d, (double *)b.p, seed, eps, g_stream));
  HIPCHK(fld_pack_gauge_cb(g->d, u0, 0, (const double *)b.p, g_stream));
  HIPCHK(fld_pack_gauge_cb(g->d, u1, 1, (const double *)b.p, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_gen_spinor(const qphix_geom *g, void *dst, uint64_t seed, int cb) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(dst);
  TRY(check_cb(cb));
  HIPCHK(fld_gen_spinor(g->d, dst, seed, cb, g_stream));
  return QPHIX_OK;
}
extern "C" int qphix_clover_build(const qphix_geom *g, const void *u0, const void *u1, double csw_r, double csw_t,
                                  double diag_mass, int cb, int what, double mu, double eps, void *out) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(u0); CHKPTR(u1); CHKPTR(out);
  TRY(check_cb(cb));
  if (what < QPHIX_CLOV_A || what > QPHIX_CLOV_ND_OFF) return fail(QPHIX_EDOMAIN, "unknown clover output %d", what);
  // clover coefficients enter halved (tests/clover_term_qdp_w.h:165-182)
  HIPCHK(fld_clover_build(g->d, u0, u1, 0.5 * csw_r, 0.5 * csw_t, diag_mass, cb, what, mu, eps, out, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}
// The same build with no allocation, free or synchronisation: the FP64
// scratch is a static device array of a fixed size (kNoAllocSites sites),
// looked up by qphix_init, which must have run (it is the call that
// allocates). Lattices of any size are built in chunks of that many sites.
// Stream-ordered on the library stream; two calls that overlap on one device
// share the scratch and race.
extern "C" int qphix_clover_build_noalloc(const qphix_geom *g, const void *u0, const void *u1, double csw_r,
                                          double csw_t, double diag_mass, int cb, int what, double mu, double eps,
                                          void *out) {
  if (!g_init) return fail(QPHIX_EINVAL, "qphix_clover_build_noalloc: call qphix_init first (it allocates)");
  CHKPTR(g); CHKPTR(u0); CHKPTR(u1); CHKPTR(out);
  TRY(check_cb(cb));
  if (what < QPHIX_CLOV_A || what > QPHIX_CLOV_ND_OFF) return fail(QPHIX_EDOMAIN, "unknown clover output %d", what);
  HIPCHK(fld_clover_build_noalloc(g->d, u0, u1, 0.5 * csw_r, 0.5 * csw_t, diag_mass, cb, what, mu, eps, out,
                                  g_device, g_stream));
  return QPHIX_OK;
}

// ---------------------------------------------------------------------------
// BLAS plumbing
// ---------------------------------------------------------------------------
static int run_blas(const qphix_geom *g, int op, const BlasArgs &a, double *out) {
  int nb = 0;
  HIPCHK(blas_launch(g->d, op, a, g_partials, &nb, g_stream));
  const int nr = blas_nred(op);
  if (nr > 0) {
    HIPCHK(blas_finalize(g_partials, nb, nr, g_dres, g_stream));
    HIPCHK(hipMemcpyAsync(g_hres, g_dres, nr * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (out)
      for (int k = 0; k < nr; k++) out[k] = g_hres[k];
  }
  return QPHIX_OK;
}
extern "C" int qphix_stream_calibrate(const qphix_geom *g, int words_in, int words_out, int reps, double *ms) {
  CHKPTR(g); CHKPTR(ms);
  const bool ok = (words_out == 6 && (words_in == 30 || words_in == 48 || words_in == 60 || words_in == 66)) ||
                  (words_in == 120 && words_out == 12);
  if (!ok)
    return fail(QPHIX_EDOMAIN, "stream shape %d / %d words per site not built (30, 48, 60, 66 / 6; 120 / 12)",
                words_in, words_out);
  if (reps < 1) return fail(QPHIX_EDOMAIN, "reps %d", reps);
  ENSURE_INIT();
  void *in = nullptr, *out = nullptr;
  TRY(alloc_zeroed((size_t)words_in * g->d.vs * 16, &in));
  int rc = alloc_zeroed((size_t)words_out * g->d.vs * 16, &out);
  float f = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;  // its own events: a caller's qphix_timer_* may be running
  if (!rc) {
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = stream_launch(g->d, words_in, words_out, in, out, g_stream);
    if (e == hipSuccess) e = hipEventRecord(e0, g_stream);
    for (int k = 0; k < reps && e == hipSuccess; k++) e = stream_launch(g->d, words_in, words_out, in, out, g_stream);
    if (e == hipSuccess) e = hipEventRecord(e1, g_stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    if (e != hipSuccess) rc = fail(QPHIX_EHIP, "stream calibration: %s", hipGetErrorString(e));
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  qphix_free(in);
  qphix_free(out);
  if (!rc) *ms = f / reps;
  return rc;
}

// the reduction left in device memory: kernels only (no copy, no memset, no
// synchronisation), so the call can be captured in a HIP graph
static int run_blas_dev(const qphix_geom *g, int op, const BlasArgs &a, double *d_out) {
  int nb = 0;
  HIPCHK(blas_launch(g->d, op, a, g_partials, &nb, g_stream));
  HIPCHK(blas_finalize(g_partials, nb, blas_nred(op), d_out, g_stream));
  return QPHIX_OK;
}
static BlasArgs bargs() {
  BlasArgs a;
  memset(&a, 0, sizeof(a));
  return a;
}

#define BLAS_PRE() \
  ENSURE_INIT();   \
  CHKPTR(g)

extern "C" int qphix_copy_spinor(const qphix_geom *g, void *res, const void *src) {
  BLAS_PRE(); CHKPTR(res); CHKPTR(src);
  BlasArgs a = bargs(); a.x = src; a.o0 = res;
  return run_blas(g, B_COPY, a, nullptr);
}
extern "C" int qphix_zero_spinor(const qphix_geom *g, void *res) {
  BLAS_PRE(); CHKPTR(res);
  BlasArgs a = bargs(); a.o0 = res;
  return run_blas(g, B_ZERO, a, nullptr);
}
extern "C" int qphix_axy(const qphix_geom *g, double alpha, const void *x, void *y) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y; a.a[0] = alpha;
  return run_blas(g, B_AXY, a, nullptr);
}
extern "C" int qphix_ax(const qphix_geom *g, double alpha, void *x) {
  BLAS_PRE(); CHKPTR(x);
  BlasArgs a = bargs(); a.o0 = x; a.a[0] = alpha;
  return run_blas(g, B_AX, a, nullptr);
}
extern "C" int qphix_ypeqx(const qphix_geom *g, const void *x, void *y) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y;
  return run_blas(g, B_YPEQX, a, nullptr);
}
extern "C" int qphix_ymeqx(const qphix_geom *g, const void *x, void *y) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y;
  return run_blas(g, B_YMEQX, a, nullptr);
}
extern "C" int qphix_aypx(const qphix_geom *g, double alpha, const void *x, void *y) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y; a.a[0] = alpha;
  return run_blas(g, B_AYPX, a, nullptr);
}
extern "C" int qphix_twisted_mass(const qphix_geom *g, const double apimu[2], const void *x, void *y) {
  BLAS_PRE(); CHKPTR(apimu); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y; a.a[0] = apimu[0]; a.a[1] = apimu[1];
  return run_blas(g, B_TWM, a, nullptr);
}
extern "C" int qphix_two_flav_twisted_mass(const qphix_geom *g, const double apimu[2], double epsilon,
                                           const void *const x[2], void *const y[2]) {
  BLAS_PRE(); CHKPTR(apimu); CHKPTR(x); CHKPTR(y);
  CHKPTR(x[0]); CHKPTR(x[1]); CHKPTR(y[0]); CHKPTR(y[1]);
  BlasArgs a = bargs(); a.x = x[0]; a.y = x[1]; a.o0 = y[0]; a.o1 = y[1];
  a.a[0] = apimu[0]; a.a[1] = apimu[1]; a.a[2] = epsilon;
  return run_blas(g, B_TWM2, a, nullptr);
}
extern "C" int qphix_axpy(const qphix_geom *g, double alpha, const void *x, void *y) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y; a.a[0] = alpha;
  return run_blas(g, B_AXPY, a, nullptr);
}
extern "C" int qphix_axpby(const qphix_geom *g, double alpha, const void *x, double beta, void *y) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y; a.a[0] = alpha; a.a[1] = beta;
  return run_blas(g, B_AXPBY, a, nullptr);
}
extern "C" int qphix_norm2_spinor(const qphix_geom *g, const void *x, double *n2) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(n2);
  BlasArgs a = bargs(); a.x = x;
  return run_blas(g, B_NORM2, a, n2);
}
extern "C" int qphix_axpy_norm2(const qphix_geom *g, double alpha, const void *x, void *y, double *n2) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y); CHKPTR(n2);
  BlasArgs a = bargs(); a.x = x; a.o0 = y; a.a[0] = alpha;
  return run_blas(g, B_AXPYNORM2, a, n2);
}
extern "C" int qphix_xmy_norm2(const qphix_geom *g, void *res, const void *x, const void *y, double *n2) {
  BLAS_PRE(); CHKPTR(res); CHKPTR(x); CHKPTR(y); CHKPTR(n2);
  BlasArgs a = bargs(); a.x = x; a.y = y; a.o0 = res;
  return run_blas(g, B_XMYNORM2, a, n2);
}
extern "C" int qphix_xmy2_norm2(const qphix_geom *g, const void *x, void *y, double *n2) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y); CHKPTR(n2);
  BlasArgs a = bargs(); a.x = x; a.o0 = y;
  return run_blas(g, B_XMY2NORM2, a, n2);
}
extern "C" int qphix_rmammp_norm2_rxpap(const qphix_geom *g, void *r, double ar, const void *mmp, void *x,
                                        const void *p, double *cp) {
  BLAS_PRE(); CHKPTR(r); CHKPTR(mmp); CHKPTR(x); CHKPTR(p); CHKPTR(cp);
  BlasArgs a = bargs(); a.x = mmp; a.y = p; a.o0 = r; a.o1 = x; a.a[0] = ar;
  return run_blas(g, B_RMAMMP, a, cp);
}
extern "C" int qphix_richardson_rxupdate_norm_r(const qphix_geom *g, void *x, void *r, const void *dx,
                                                const void *dr, double *n2) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(r); CHKPTR(dx); CHKPTR(dr); CHKPTR(n2);
  BlasArgs a = bargs(); a.x = dx; a.y = dr; a.o0 = x; a.o1 = r;
  return run_blas(g, B_RICH, a, n2);
}
extern "C" int qphix_bicgstab_xmy(const qphix_geom *g, const void *x, void *y) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y;
  return run_blas(g, B_BICG_XMY, a, nullptr);
}
extern "C" int qphix_inner_product(const qphix_geom *g, const void *l, const void *r, double out[2]) {
  BLAS_PRE(); CHKPTR(l); CHKPTR(r); CHKPTR(out);
  BlasArgs a = bargs(); a.x = l; a.y = r;
  return run_blas(g, B_CDOT, a, out);
}
extern "C" int qphix_inner_product_norm(const qphix_geom *g, const void *l, const void *r, double out[3]) {
  BLAS_PRE(); CHKPTR(l); CHKPTR(r); CHKPTR(out);
  BlasArgs a = bargs(); a.x = l; a.y = r;
  return run_blas(g, B_CDOTNORM, a, out);
}

// reductions into device memory (graph-capturable forms of the above)
extern "C" int qphix_norm2_spinor_dev(const qphix_geom *g, const void *x, double *d_out) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(d_out);
  BlasArgs a = bargs(); a.x = x;
  return run_blas_dev(g, B_NORM2, a, d_out);
}
extern "C" int qphix_axpy_norm2_dev(const qphix_geom *g, double alpha, const void *x, void *y, double *d_out) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(y); CHKPTR(d_out);
  BlasArgs a = bargs(); a.x = x; a.o0 = y; a.a[0] = alpha;
  return run_blas_dev(g, B_AXPYNORM2, a, d_out);
}
extern "C" int qphix_xmy_norm2_dev(const qphix_geom *g, void *res, const void *x, const void *y, double *d_out) {
  BLAS_PRE(); CHKPTR(res); CHKPTR(x); CHKPTR(y); CHKPTR(d_out);
  BlasArgs a = bargs(); a.x = x; a.y = y; a.o0 = res;
  return run_blas_dev(g, B_XMYNORM2, a, d_out);
}
extern "C" int qphix_inner_product_dev(const qphix_geom *g, const void *l, const void *r, double *d_out) {
  BLAS_PRE(); CHKPTR(l); CHKPTR(r); CHKPTR(d_out);
  BlasArgs a = bargs(); a.x = l; a.y = r;
  return run_blas_dev(g, B_CDOT, a, d_out);
}
extern "C" int qphix_inner_product_norm_dev(const qphix_geom *g, const void *l, const void *r, double *d_out) {
  BLAS_PRE(); CHKPTR(l); CHKPTR(r); CHKPTR(d_out);
  BlasArgs a = bargs(); a.x = l; a.y = r;
  return run_blas_dev(g, B_CDOTNORM, a, d_out);
}
extern "C" int qphix_bicgstab_p_update(const qphix_geom *g, const void *r, void *p, const void *v,
                                       const double beta[2], const double omega[2]) {
  BLAS_PRE(); CHKPTR(r); CHKPTR(p); CHKPTR(v); CHKPTR(beta); CHKPTR(omega);
  BlasArgs a = bargs(); a.x = r; a.y = v; a.o0 = p;
  a.a[0] = beta[0]; a.a[1] = beta[1]; a.a[2] = omega[0]; a.a[3] = omega[1];
  return run_blas(g, B_PUPD, a, nullptr);
}
extern "C" int qphix_bicgstab_s_update(const qphix_geom *g, const double alpha[2], void *s, const void *v) {
  BLAS_PRE(); CHKPTR(alpha); CHKPTR(s); CHKPTR(v);
  BlasArgs a = bargs(); a.x = v; a.o0 = s; a.a[0] = alpha[0]; a.a[1] = alpha[1];
  return run_blas(g, B_SUPD, a, nullptr);
}
extern "C" int qphix_caxpy(const qphix_geom *g, const double alpha[2], const void *x, void *y) {
  BLAS_PRE(); CHKPTR(alpha); CHKPTR(x); CHKPTR(y);
  BlasArgs a = bargs(); a.x = x; a.o0 = y; a.a[0] = alpha[0]; a.a[1] = alpha[1];
  return run_blas(g, B_CAXPY, a, nullptr);
}
extern "C" int qphix_bicgstab_rxupdate(const qphix_geom *g, void *x, void *r, const void *t, const void *p,
                                       const double omega[2], const double alpha[2], double *r_norm) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(r); CHKPTR(t); CHKPTR(p); CHKPTR(omega); CHKPTR(alpha); CHKPTR(r_norm);
  BlasArgs a = bargs(); a.x = t; a.y = p; a.o0 = x; a.o1 = r;
  a.a[0] = omega[0]; a.a[1] = omega[1]; a.a[2] = alpha[0]; a.a[3] = alpha[1];
  return run_blas(g, B_RXUPD, a, r_norm);
}
extern "C" int qphix_mr_rxupdate(const qphix_geom *g, void *x, void *r, const void *mr, const double ac[2]) {
  BLAS_PRE(); CHKPTR(x); CHKPTR(r); CHKPTR(mr); CHKPTR(ac);
  BlasArgs a = bargs(); a.x = mr; a.o0 = x; a.o1 = r; a.a[0] = ac[0]; a.a[1] = ac[1];
  return run_blas(g, B_MRUPD, a, nullptr);
}

// ---------------------------------------------------------------------------
// Dslash objects
// ---------------------------------------------------------------------------
extern "C" int qphix_dslash_create(const qphix_geom *g, const qphix_dslash_params *p, qphix_dslash **d) {
  CHKPTR(g); CHKPTR(p); CHKPTR(d);
  qphix_dslash *x = new qphix_dslash;
  x->g = g;
  x->p = *p;
  *d = x;
  return QPHIX_OK;
}
extern "C" int qphix_dslash_destroy(qphix_dslash *d) {
  delete d;
  return QPHIX_OK;
}

// ---------------------------------------------------------------------------
// hop-kernel autotuner: per (precision, compression, kernel, isign, lattice) the
// work-group size and the traversal (natural, or z-block x t tiles per t-slab)
// are timed once on the first call, into a scratch output so that aliased
// arguments are never touched twice, and the fastest is kept for the process.
// All candidates compute bit-identical results; only the order of sites differs.
// ---------------------------------------------------------------------------
struct HopCfg {
  int block, bz, slabs;
  double ms;
  int rt;   // right-hand sides per launch (multi-RHS keys)
  int reg;  // right-hand sides per work-item (register tile; 1 = one per work-item)
  int by;   // rows per traversal tile (0: all of y)
  int lds;  // dynamic LDS bytes per work-group capping the work-groups per CU (0: none)
};
static std::map<std::string, HopCfg> g_tuned;
static int g_tune_enabled = -1;            // -1: read QPHIX_TUNE on first use
static HopCfg g_forced = {0, 0, 0, 0, 0, 1};  // block > 0: every hop uses this configuration
static int g_forced_reg = 0, g_forced_rt = 0;  // reg > 0: multi-RHS tiling set by qphix_tune_rhs

static hipError_t launch_hop_raw(const GeomDesc &g, int kind, int isign, const HopArgs &a, int block) {
  switch (g.prec) {
    case P64: return launch_hop_prec<P64>(kind, g.c12 != 0, isign, a, block, g_stream);
    case P32: return launch_hop_prec<P32>(kind, g.c12 != 0, isign, a, block, g_stream);
    case P16: return launch_hop_prec<P16>(kind, g.c12 != 0, isign, a, block, g_stream);
    default: return launch_hop_prec<PB16>(kind, g.c12 != 0, isign, a, block, g_stream);
  }
}

// launch of a.nrhs right-hand sides with `reg` of them per work-item; reg = 8
// is the matrix-core kernel (FP32 or FP16 storage), 8 right-hand sides per
// work-group
static hipError_t launch_hop_tile_raw(const GeomDesc &g, int kind, int isign, const HopArgs &a, int block, int reg) {
  if (reg == 8) {
    if (g.prec == P32) return launch_hop_mfma_f32(kind, g.c12 != 0, isign, a, g_stream);
    if (g.prec == P16) return launch_hop_mfma_f16(kind, g.c12 != 0, isign, a, g_stream);
    return hipErrorInvalidValue;
  }
  switch (g.prec) {
    case P64: return launch_hop_tile_prec<P64>(kind, g.c12 != 0, isign, reg, a, block, g_stream);
    case P32: return launch_hop_tile_prec<P32>(kind, g.c12 != 0, isign, reg, a, block, g_stream);
    case P16: return launch_hop_tile_prec<P16>(kind, g.c12 != 0, isign, reg, a, block, g_stream);
    default: return launch_hop_tile_prec<PB16>(kind, g.c12 != 0, isign, reg, a, block, g_stream);
  }
}

// a.nrhs right-hand sides (pointer tables filled) in one or two launches:
// multiples of `reg` through the register-tiled kernel, the rest one per work-item
static hipError_t launch_hop_group(const GeomDesc &g, int kind, int isign, HopArgs a, int block, int reg) {
  const int n = a.nrhs;
  const int main = reg > 1 ? n - n % reg : 0;
  if (main > 0) {
    a.nrhs = main;
    hipError_t e = launch_hop_tile_raw(g, kind, isign, a, block, reg);
    if (e != hipSuccess) return e;
  }
  const int rest = n - main;
  if (rest == 0) return hipSuccess;
  a.nrhs = rest;
  for (int r = 0; r < rest; r++) {
    a.psis[r] = a.psis[main + r];
    a.outs[r] = a.outs[main + r];
    a.chis[r] = a.chis[main + r];
  }
  if (rest == 1) {
    a.psi = a.psis[0];
    a.out = a.outs[0];
    a.chi = a.chis[0];
  }
  return launch_hop_raw(g, kind, isign, a, block);
}

static bool tuning_on(const GeomDesc &g) {
  if (g_tune_enabled < 0) {
    const char *e = getenv("QPHIX_TUNE");
    g_tune_enabled = (e && e[0] == '0') ? 0 : 1;
  }
  // small lattices: the traversal order does not matter and tuning would dominate
  return g_tune_enabled && g.vh >= (1L << 17);
}

static size_t spinor_bytes(const GeomDesc &g) {
  return g.prec == P64 ? field_bytes<P64>(kSpinorReals, g.vs)
         : g.prec == P32 ? field_bytes<P32>(kSpinorReals, g.vs) : field_bytes<P16>(kSpinorReals, g.vs);
  // (PB16 fields have the P16 size)
}

// time `reps` launches of a (reg > 0: the multi-RHS group launch); returns ms
// per launch (1e30 if it cannot launch)
static double time_hop(const GeomDesc &g, int kind, int isign, const HopArgs &a, int block, int reps, int reg = 0) {
  auto go = [&]() { return reg > 0 ? launch_hop_group(g, kind, isign, a, block, reg)
                                   : launch_hop_raw(g, kind, isign, a, block); };
  if (go() != hipSuccess) {
    (void)hipGetLastError();
    return 1e30;
  }
  if (hipEventRecord(g_ev0, g_stream) != hipSuccess) return 1e30;
  for (int r = 0; r < reps; r++)
    if (go() != hipSuccess) return 1e30;
  if (hipEventRecord(g_ev1, g_stream) != hipSuccess || hipEventSynchronize(g_ev1) != hipSuccess) return 1e30;
  float f = 0;
  if (hipEventElapsedTime(&f, g_ev0, g_ev1) != hipSuccess) return 1e30;
  return f / reps;
}

static void tune_hop(const GeomDesc &g, int kind, int isign, const HopArgs &a0, HopCfg &best) {
  best = HopCfg{g_hop_block, 0, 1, -1, 1, 1};
  if (g_forced.block > 0) {
    best = g_forced;
    if (best.bz > 0 && (g.L[2] % best.bz || g.L[3] % best.slabs)) {
      best.bz = 0;
      best.slabs = 1;
    }
    if (best.bz == 0 || (best.by > 0 && g.L[1] % best.by)) best.by = 0;
    return;
  }
  if (!tuning_on(g)) return;
  std::vector<HopCfg> cand;
  for (int blk : {128, 256}) {
    cand.push_back({blk, 0, 1, 0, 1, 1});
    for (int bz : {2, 4, 6, 8, 12, 16, 24})
      for (int sl : {1, 2, 8, 16})
        if (bz <= g.L[2] && g.L[2] % bz == 0 && g.L[3] % sl == 0 && g.L[3] / sl >= 2)
          cand.push_back({blk, bz, sl, 0, 1, 1});
  }
  DevBuf scratch;
  if (hipMalloc(&scratch.p, spinor_bytes(g)) != hipSuccess) {
    (void)hipGetLastError();
    return;  // keep the default configuration
  }
  HopArgs a = a0;
  a.nrhs = 1;
  a.out = scratch.p;
  for (HopCfg &c : cand) {
    a.bz = c.bz;
    a.slabs = c.slabs;
    c.ms = time_hop(g, kind, isign, a, c.block, 4);
    if (best.ms < 0 || c.ms < best.ms) best = c;
  }
}

// right-hand-side tiling for n right-hand sides, given the single-RHS
// traversal: `reg` sources per work-item (register tile: each link and clover
// block is loaded once for all of them) and `rt` sources per launch (the
// work-groups of one site range for all rt sources run together and share the
// remaining link and clover reads through the L2)
static void tune_rhs_tile(const GeomDesc &g, int kind, int isign, const HopArgs &a0, int n, HopCfg &cfg) {
  cfg.rt = n < kMaxRhs ? n : kMaxRhs;
  cfg.reg = 1;
  if (g_forced_reg > 0) {
    cfg.reg = g_forced_reg;
    cfg.rt = g_forced_rt;
    return;
  }
  if (g_forced.block > 0) {
    cfg.reg = g_forced.reg;
    return;
  }
  if (!tuning_on(g) || n == 1) return;
  const int rmax = cfg.rt;
  std::vector<DevBuf> outs(rmax);
  for (int r = 0; r < rmax; r++)
    if (hipMalloc(&outs[r].p, spinor_bytes(g)) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
  HopArgs a = a0;
  // ms per source of the rmax sources applied in groups of rt (the last
  // group may be smaller), as launch_hop_n applies them
  auto time_groups = [&](int rt, int reg, int blk) {
    auto go = [&]() {
      for (int r0 = 0; r0 < rmax; r0 += rt) {
        const int m = (rmax - r0) < rt ? (rmax - r0) : rt;
        a.nrhs = m;
        for (int r = 0; r < m; r++) {
          a.psis[r] = a0.psis[r0 + r];
          a.chis[r] = a0.chis[r0 + r];
          a.outs[r] = outs[r0 + r].p;
        }
        a.psi = a.psis[0];
        a.chi = a.chis[0];
        a.out = a.outs[0];
        if (launch_hop_group(g, kind, isign, a, blk, m == 1 ? 1 : reg) != hipSuccess) return false;
      }
      return true;
    };
    if (!go()) {
      (void)hipGetLastError();
      return 1e30;
    }
    if (hipEventRecord(g_ev0, g_stream) != hipSuccess) return 1e30;
    for (int k = 0; k < 2; k++)
      if (!go()) return 1e30;
    if (hipEventRecord(g_ev1, g_stream) != hipSuccess || hipEventSynchronize(g_ev1) != hipSuccess) return 1e30;
    float f = 0;
    if (hipEventElapsedTime(&f, g_ev0, g_ev1) != hipSuccess) return 1e30;
    return (double)f / (2.0 * rmax);
  };
  // first the tiling under the single-RHS traversal; the matrix-core form
  // (reg 8) is not a candidate: it runs slower per source than the register
  // tiles (qphix_tune_rhs(8, n) still forces it)
  a.bz = cfg.bz;
  a.slabs = cfg.slabs;
  std::vector<HopCfg> tiles;
  for (int reg : {1, 2, 4}) {
    for (int rt : {1, 2, 4, 8, 16, 32, rmax}) {
      if (rt > rmax || rt < reg || rt % reg) continue;
      if (rt == rmax && (rt == 1 || rt == 2 || rt == 4 || rt == 8 || rt == 16 || rt == 32)) continue;
      HopCfg c = cfg;
      c.rt = rt;
      c.reg = reg;
      c.ms = time_groups(rt, reg, cfg.block);
      tiles.push_back(c);
    }
  }
  std::sort(tiles.begin(), tiles.end(), [](const HopCfg &x, const HopCfg &y) { return x.ms < y.ms; });
  cfg = tiles[0];
  // then the traversal for the best tiling. With several right-hand sides
  // per site range the source spinors of all of them share the L2, so
  // smaller (y, z) tiles and fewer work-groups in flight per CU (an LDS
  // reservation the kernel does not use) keep the neighbours of the sites in
  // flight resident; the single-RHS traversal stays a candidate.
  const int ny = g.L[1], nz = g.L[2], nt = g.L[3];
  // (bz, by) tiles; LDS reservations of 40 KiB (three work-groups per CU) and,
  // for 256-thread work-groups, 64 KiB (two)
  static const int kTiles[][2] = {{4, 0}, {8, 0}, {2, 16}, {4, 8}, {6, 8}, {6, 16}, {8, 16}};
  for (size_t t = 0; t < tiles.size() && t < 1; t++) {
    for (int blk : {128, 256})
      for (const auto &zy : kTiles)
        for (int lds : {0, 40960, 65536}) {
          const int bz = zy[0], by = zy[1];
          if (nz % bz || (by > 0 && ny % by) || nt < 2 || (lds > 40960 && blk != 256)) continue;
          a.bz = bz;
          a.by = by;
          a.slabs = 1;
          a.lds_cap = lds;
          const double per = time_groups(tiles[t].rt, tiles[t].reg, blk);
          if (per < cfg.ms) {
            cfg = tiles[t];
            cfg.block = blk;
            cfg.bz = bz;
            cfg.by = by;
            cfg.slabs = 1;
            cfg.lds = lds;
            cfg.ms = per;
          }
        }
  }
}

// set by the device-resident solver loops while they enqueue iterations: the
// hop launches of their operator applications skip themselves once the
// device state says the solve is done
static const int *g_hop_skip = nullptr;

static HopArgs make_hop_args(const qphix_dslash *d, int cb, const void *u, const void *clov, double coef,
                             double alpha, double mu, double muinv, double rho) {
  const GeomDesc &g = d->g->d;
  HopArgs a;
  memset(&a, 0, sizeof(a));
  a.u = u;
  a.clov = clov;
  a.vs = g.vs;
  a.vh = (int)g.vh;
  a.nxh = g.L[0] / 2;
  a.ny = g.L[1];
  a.nz = g.L[2];
  a.nt = g.L[3];
  a.cb = cb;
  a.cs = coef * d->p.aniso_s;
  a.ct = coef * d->p.aniso_t;
  a.ctf = a.ct * d->p.t_boundary;
  a.ctb = a.ct * d->p.t_boundary;
  for (int k = 0; k < 4; k++) {
    if (d->p.use_tbc[k]) a.tbc_mask |= 1 << k;
    a.tbc_r[k] = d->p.tbc_phases[k][0];
    a.tbc_i[k] = d->p.tbc_phases[k][1];
  }
  a.alpha = alpha;
  a.mu = mu;
  a.muinv = muinv;
  a.rho = rho;
  a.slabs = 1;
  a.nrhs = 1;
  a.skip = g_hop_skip;
  return a;
}

// one kernel-family member applied to n right-hand sides (n = 1: the plain call)
static int launch_hop_n(const qphix_dslash *d, int kind, int isign, int cb, void *const *res,
                        const void *const *psi, const void *const *chi, int n, const void *u, const void *clov,
                        double coef, double alpha, double mu, double muinv, double rho, int dir = 0) {
  ENSURE_INIT();
  CHKPTR(d); CHKPTR(res); CHKPTR(psi); CHKPTR(u);
  if (isign != 1 && isign != -1) return fail(QPHIX_EDOMAIN, "isign must be +1 or -1, got %d", isign);
  TRY(check_cb(cb));
  if (n < 1) return fail(QPHIX_EDOMAIN, "number of right-hand sides %d", n);
  if ((kind == KD_CLOVER || kind == KA_CLOVER || kind == KD_FULL) && !clov)
    return fail(QPHIX_EINVAL, "clover field required");
  const bool needs_chi = kind == KA_WILSON || kind == KA_CLOVER || kind == KA_TM;
  for (int r = 0; r < n; r++) {
    if (!res[r] || !psi[r]) return fail(QPHIX_EINVAL, "null spinor for right-hand side %d", r);
    if (needs_chi && (!chi || !chi[r])) return fail(QPHIX_EINVAL, "chi required");
  }
  const GeomDesc &g = d->g->d;
  HopArgs a = make_hop_args(d, cb, u, clov, coef, alpha, mu, muinv, rho);
  a.dir = dir;
  char key[128];
  snprintf(key, sizeof(key), "p%d c%d k%d s%d %dx%dx%dx%d", g.prec, g.c12, kind, isign, g.L[0], g.L[1], g.L[2],
           g.L[3]);
  auto it = g_tuned.find(key);
  if (it == g_tuned.end()) {
    HopCfg c;
    HopArgs t = a;
    t.psi = psi[0];
    t.chi = needs_chi ? chi[0] : nullptr;
    tune_hop(g, kind, isign, t, c);
    it = g_tuned.emplace(key, c).first;
  }
  HopCfg cfg = it->second;
  if (n > 1) {
    // one tuning per power-of-two class of n: solvers that drop converged
    // right-hand sides apply the operator to a shrinking set
    int nb = 2;
    while (nb * 2 <= n && nb < kMaxRhs) nb *= 2;
    char mkey[160];
    snprintf(mkey, sizeof(mkey), "%s n%d", key, nb);
    auto mt = g_tuned.find(mkey);
    if (mt == g_tuned.end()) {
      HopCfg c = cfg;
      HopArgs t = a;
      for (int r = 0; r < n && r < kMaxRhs; r++) {
        t.psis[r] = psi[r];
        t.chis[r] = needs_chi ? chi[r] : nullptr;
      }
      tune_rhs_tile(g, kind, isign, t, n, c);
      mt = g_tuned.emplace(mkey, c).first;
    }
    cfg = mt->second;
  }
  a.bz = cfg.bz;
  a.slabs = cfg.slabs;
  a.by = cfg.by;
  a.lds_cap = cfg.lds;
  const int rt = n == 1 ? 1 : (cfg.rt > 0 ? cfg.rt : 1);
  const int reg = n == 1 ? 1 : cfg.reg;
  for (int r0 = 0; r0 < n; r0 += rt) {
    const int m = (n - r0) < rt ? (n - r0) : rt;
    a.nrhs = m;
    for (int r = 0; r < m; r++) {
      a.psis[r] = psi[r0 + r];
      a.outs[r] = res[r0 + r];
      a.chis[r] = needs_chi ? chi[r0 + r] : nullptr;
    }
    if (m == 1) {
      a.psi = psi[r0];
      a.out = res[r0];
      a.chi = needs_chi ? chi[r0] : nullptr;
    }
    hipError_t e = launch_hop_group(g, kind, isign, a, cfg.block, reg);
    if (e != hipSuccess) return fail(QPHIX_EHIP, "hop kernel launch: %s", hipGetErrorString(e));
  }
  return QPHIX_OK;
}

static int launch_hop(const qphix_dslash *d, int kind, int isign, int cb, void *res, const void *psi, const void *u,
                      const void *clov, const void *chi, double coef, double alpha, double mu, double muinv,
                      double rho, int dir = 0) {
  void *r[1] = {res};
  const void *p[1] = {psi};
  const void *c[1] = {chi};
  return launch_hop_n(d, kind, isign, cb, r, p, c, 1, u, clov, coef, alpha, mu, muinv, rho, dir);
}

// mode 0: no tuning (natural order, default block); 1: autotune on first use;
// 2: force (block, bz, slabs) for every lattice it divides. Clears the table.
extern "C" int qphix_tune_config(int mode, int block, int bz, int slabs) {
  if (mode < 0 || mode > 2) return fail(QPHIX_EDOMAIN, "tune mode %d", mode);
  if (mode == 2 && (block <= 0 || block > 256 || (block & 63) || bz < 0 || (bz > 0 && slabs < 1)))
    return fail(QPHIX_EDOMAIN, "forced configuration block %d bz %d slabs %d", block, bz, slabs);
  g_tuned.clear();
  g_tune_enabled = mode == 1 ? 1 : 0;
  g_forced = mode == 2 ? HopCfg{block, bz, bz > 0 ? slabs : 1, 0, kMaxRhs, 1} : HopCfg{0, 0, 0, 0, 0, 1};
  return QPHIX_OK;
}

// mode 2 of qphix_tune_config with (y, z) tiles of by rows and an LDS
// reservation of lds bytes per work-group (it caps the work-groups per CU)
extern "C" int qphix_tune_traversal(int block, int bz, int by, int slabs, int lds) {
  if (by < 0 || lds < 0 || lds > 65536) return fail(QPHIX_EDOMAIN, "tile rows %d, LDS reservation %d", by, lds);
  TRY(qphix_tune_config(2, block, bz, slabs));
  g_forced.by = bz > 0 ? by : 0;
  g_forced.lds = lds;
  return QPHIX_OK;
}

extern "C" int qphix_tune_rhs(int reg, int rt) {
  if (reg != 0 && reg != 1 && reg != 2 && reg != 4 && reg != 8)
    return fail(QPHIX_EDOMAIN, "register tile %d (1, 2, 4, or 8 for the matrix-core kernel)", reg);
  if (reg > 0 && (rt < reg || rt > kMaxRhs || rt % reg))
    return fail(QPHIX_EDOMAIN, "RHS per launch %d with register tile %d", rt, reg);
  for (auto it = g_tuned.begin(); it != g_tuned.end();)  // multi-RHS entries only
    it = it->first.find(" n") != std::string::npos ? g_tuned.erase(it) : std::next(it);
  g_forced_reg = reg;
  g_forced_rt = reg > 0 ? rt : 0;
  return QPHIX_OK;
}

// text summary of the tuned hop-kernel launch configurations
extern "C" const char *qphix_tune_summary(void) {
  static std::string out;
  out.clear();
  for (auto &kv : g_tuned) {
    char line[256];
    snprintf(line, sizeof(line), "%s: block %d bz %d by %d slabs %d rhs-tile %d reg-tile %d lds %d %.4f ms\n",
             kv.first.c_str(), kv.second.block, kv.second.bz, kv.second.by, kv.second.slabs, kv.second.rt,
             kv.second.reg, kv.second.lds, kv.second.ms);
    out += line;
  }
  return out.c_str();
}

extern "C" int qphix_wilson_dslash(qphix_dslash *d, void *res, const void *psi, const void *u, int isign, int cb) {
  return launch_hop(d, KD_WILSON, isign, cb, res, psi, u, nullptr, nullptr, 1.0, 0, 0, 0, 0);
}
// dslash_def.h:54 / dslash_body.h:350-358 (isign = +1, as the reference)
extern "C" int qphix_wilson_dslash_dir(qphix_dslash *d, void *res, const void *psi, const void *u, int cb, int dir) {
  if (dir < 0 || dir > 7) return fail(QPHIX_EDOMAIN, "direction %d (0..7)", dir);
  return launch_hop(d, KD_DIR, 1, cb, res, psi, u, nullptr, nullptr, 1.0, 0, 0, 0, 0, dir);
}
extern "C" int qphix_wilson_achimbdpsi(qphix_dslash *d, void *res, const void *psi, const void *chi, const void *u,
                                       double alpha, double beta, int isign, int cb) {
  return launch_hop(d, KA_WILSON, isign, cb, res, psi, u, nullptr, chi, -beta, alpha, 0, 0, 0);
}
extern "C" int qphix_clov_dslash(qphix_dslash *d, void *res, const void *psi, const void *u, const void *invclov,
                                 int isign, int cb) {
  return launch_hop(d, KD_CLOVER, isign, cb, res, psi, u, invclov, nullptr, 1.0, 0, 0, 0, 0);
}
extern "C" int qphix_clov_achimbdpsi(qphix_dslash *d, void *res, const void *psi, const void *chi, const void *u,
                                     const void *clov, double beta, int isign, int cb) {
  CHKPTR(d);
  return launch_hop(d, KA_CLOVER, isign, cb, res, psi, u, clov, chi, -beta, 0, 0, 0, d->p.prec_mass_rho);
}
// multi-right-hand-side forms (no reference counterpart): one launch per
// right-hand-side tile, links and clover shared through L2
extern "C" int qphix_wilson_dslash_multi(qphix_dslash *d, void *const res[], const void *const psi[], int n,
                                         const void *u, int isign, int cb) {
  return launch_hop_n(d, KD_WILSON, isign, cb, res, psi, nullptr, n, u, nullptr, 1.0, 0, 0, 0, 0);
}
extern "C" int qphix_wilson_achimbdpsi_multi(qphix_dslash *d, void *const res[], const void *const psi[],
                                             const void *const chi[], int n, const void *u, double alpha,
                                             double beta, int isign, int cb) {
  return launch_hop_n(d, KA_WILSON, isign, cb, res, psi, chi, n, u, nullptr, -beta, alpha, 0, 0, 0);
}
extern "C" int qphix_clov_dslash_multi(qphix_dslash *d, void *const res[], const void *const psi[], int n,
                                       const void *u, const void *invclov, int isign, int cb) {
  return launch_hop_n(d, KD_CLOVER, isign, cb, res, psi, nullptr, n, u, invclov, 1.0, 0, 0, 0, 0);
}
extern "C" int qphix_clov_achimbdpsi_multi(qphix_dslash *d, void *const res[], const void *const psi[],
                                           const void *const chi[], int n, const void *u, const void *clov,
                                           double beta, int isign, int cb) {
  CHKPTR(d);
  return launch_hop_n(d, KA_CLOVER, isign, cb, res, psi, chi, n, u, clov, -beta, 0, 0, 0, d->p.prec_mass_rho);
}
extern "C" int qphix_clov_mult(qphix_dslash *d, void *res, const void *chi, const void *clov) {
  ENSURE_INIT();
  CHKPTR(d); CHKPTR(res); CHKPTR(chi); CHKPTR(clov);
  BlasArgs a = bargs(); a.x = chi; a.y = clov; a.o0 = res;
  return run_blas(d->g, B_CLOVMULT, a, nullptr);
}
// res = A chi with the packed Hermitian blocks (full = 0) or the full complex
// blocks (full = 1) of one checkerboard
extern "C" int qphix_clover_apply(const qphix_geom *g, void *res, const void *chi, const void *clov, int full) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(res); CHKPTR(chi); CHKPTR(clov);
  if (full != 0 && full != 1) return fail(QPHIX_EDOMAIN, "full must be 0 or 1, got %d", full);
  BlasArgs a = bargs(); a.x = chi; a.y = clov; a.o0 = res;
  return run_blas(g, full ? B_FULLMULT : B_CLOVMULT, a, nullptr);
}
extern "C" int qphix_tm_dslash(qphix_dslash *d, void *res, const void *psi, const void *u, int isign, int cb) {
  CHKPTR(d);
  const double al = 4.0 + d->p.mass, mu = d->p.twisted_mass;
  return launch_hop(d, KD_TM, isign, cb, res, psi, u, nullptr, nullptr, 1.0, 0, mu / al, al / (mu * mu + al * al), 0);
}
extern "C" int qphix_tm_achimbdpsi(qphix_dslash *d, void *res, const void *psi, const void *chi, const void *u,
                                   double alpha, double beta, int isign, int cb, double musign) {
  CHKPTR(d);
  const double al = 4.0 + d->p.mass;
  return launch_hop(d, KA_TM, isign, cb, res, psi, u, nullptr, chi, -beta, alpha, musign * d->p.twisted_mass / al, 0,
                    0);
}
extern "C" int qphix_tm_two_flav_achimbdpsi(qphix_dslash *d, void *const res[2], const void *const psi[2],
                                            const void *const chi[2], const void *u, double alpha, double beta,
                                            double epsilon, int isign, int cb) {
  CHKPTR(d); CHKPTR(res); CHKPTR(psi); CHKPTR(chi);
  for (int f = 0; f < 2; f++) {
    TRY(qphix_tm_achimbdpsi(d, res[f], psi[f], chi[f], u, alpha, beta, isign, cb, 1.0 - 2.0 * f));
    TRY(qphix_axpy(d->g, -epsilon, chi[1 - f], res[f]));
  }
  return QPHIX_OK;
}
extern "C" int qphix_tmclov_dslash(qphix_dslash *d, void *res, const void *psi, const void *u,
                                   const void *const invclov[2], int isign, int cb, int fl) {
  CHKPTR(invclov);
  // up flavour: +mu block for M, -mu block for M^+; reversed for down
  const int clidx = (fl == 0) ? (isign == 1 ? 0 : 1) : (isign == 1 ? 1 : 0);
  return launch_hop(d, KD_FULL, isign, cb, res, psi, u, invclov[clidx], nullptr, 1.0, 0, 0, 0, 0);
}
extern "C" int qphix_tmclov_achimbdpsi(qphix_dslash *d, void *res, const void *psi, const void *chi, const void *u,
                                       const void *clov, double beta, int isign, int cb, int fl) {
  CHKPTR(d);
  const double rho = (fl == 1 ? -1.0 : 1.0) * d->p.prec_mass_rho;
  return launch_hop(d, KA_CLOVER, isign, cb, res, psi, u, clov, chi, -beta, 0, 0, 0, rho);
}
extern "C" int qphix_tmclov_two_flav_achimbdpsi(qphix_dslash *d, void *const res[2], const void *const psi[2],
                                                const void *const chi[2], const void *u, const void *clov,
                                                double beta, double epsilon, int isign, int cb) {
  CHKPTR(d); CHKPTR(res); CHKPTR(psi); CHKPTR(chi);
  for (int f = 0; f < 2; f++) {
    TRY(qphix_tmclov_achimbdpsi(d, res[f], psi[f], chi[f], u, clov, beta, isign, cb, f));
    TRY(qphix_axpy(d->g, -epsilon, chi[1 - f], res[f]));
  }
  return QPHIX_OK;
}
extern "C" int qphix_tmclov_two_flav_inverse_clover_term(qphix_dslash *d, void *const res[2],
                                                         const void *const psi[2], const void *const fcl[2],
                                                         const void *cl_offdiag, int isign) {
  ENSURE_INIT();
  CHKPTR(d); CHKPTR(res); CHKPTR(psi); CHKPTR(fcl); CHKPTR(cl_offdiag);
  CHKPTR(res[0]); CHKPTR(res[1]); CHKPTR(psi[0]); CHKPTR(psi[1]); CHKPTR(fcl[0]); CHKPTR(fcl[1]);
  BlasArgs a = bargs();
  a.x = psi[0];
  a.y = psi[1];
  a.z = isign == 1 ? fcl[0] : fcl[1];
  a.w = isign == 1 ? fcl[1] : fcl[0];
  a.o2 = const_cast<void *>(cl_offdiag);
  a.o0 = res[0];
  a.o1 = res[1];
  return run_blas(d->g, B_TWOFLAV_INVCLOV, a, nullptr);
}

// ---------------------------------------------------------------------------
// even-odd operators
// ---------------------------------------------------------------------------
extern "C" int qphix_op_create(const qphix_geom *g, const qphix_op_params *p, qphix_op **out) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(p); CHKPTR(out);
  if (p->kind < QPHIX_OP_WILSON || p->kind > QPHIX_OP_NDTMCLOVER)
    return fail(QPHIX_EDOMAIN, "unknown operator kind %d", p->kind);
  qphix_op *op = new qphix_op();
  op->g = g;
  op->p = *p;
  op->d.g = g;
  op->d.p = p->hop;
  op->d.p.mass = p->mass;
  op->d.p.twisted_mass = p->twisted_mass;
  op->d.p.prec_mass_rho = (p->kind == QPHIX_OP_NDTMCLOVER) ? p->twisted_mass : p->prec_mass_rho;
  op->nflav = (p->kind == QPHIX_OP_NDTMWILSON || p->kind == QPHIX_OP_NDTMCLOVER) ? 2 : 1;
  for (int f = 0; f < op->nflav; f++) {
    int r = qphix_alloc_spinor(g, &op->tmp[f]);
    if (!r && op->nflav == 2) r = qphix_alloc_spinor(g, &op->tmp2[f]);
    if (r) {
      qphix_op_destroy(op);
      return r;
    }
  }
  *out = op;
  return QPHIX_OK;
}
extern "C" int qphix_op_destroy(qphix_op *op) {
  if (!op) return QPHIX_OK;
  for (int f = 0; f < 2; f++) {
    qphix_free(op->tmp[f]);
    qphix_free(op->tmp2[f]);
  }
  for (void *t : op->mtmp) qphix_free(t);
  for (void *t : op->utmp) qphix_free(t);
  delete op;
  return QPHIX_OK;
}
extern "C" int qphix_op_nflav(const qphix_op *op) { return op ? op->nflav : 0; }

extern "C" int qphix_op_apply(qphix_op *op, void *const res[], const void *const in[], int isign, int t) {
  CHKPTR(op); CHKPTR(res); CHKPTR(in);
  TRY(check_cb(t));
  const int o = 1 - t;
  const qphix_op_params &p = op->p;
  qphix_dslash *D = &op->d;
  switch (p.kind) {
    case QPHIX_OP_WILSON: {
      const double al = 4.0 + p.mass;
      TRY(qphix_wilson_dslash(D, op->tmp[0], in[0], p.u[o], isign, o));
      return qphix_wilson_achimbdpsi(D, res[0], op->tmp[0], in[0], p.u[t], al, 0.25 / al, isign, t);
    }
    case QPHIX_OP_CLOVER:
      TRY(qphix_clov_dslash(D, op->tmp[0], in[0], p.u[o], p.invclov[o], isign, o));
      return qphix_clov_achimbdpsi(D, res[0], op->tmp[0], in[0], p.u[t], p.clov[t], 0.25, isign, t);
    case QPHIX_OP_TMWILSON:
      TRY(qphix_tm_dslash(D, op->tmp[0], in[0], p.u[o], isign, o));
      return qphix_tm_achimbdpsi(D, res[0], op->tmp[0], in[0], p.u[t], 4.0 + p.mass, 0.25, isign, t, 1.0);
    case QPHIX_OP_TMCLOVER:
      TRY(qphix_tmclov_dslash(D, op->tmp[0], in[0], p.u[o], p.fullclov[o], isign, o, 0));
      return qphix_tmclov_achimbdpsi(D, res[0], op->tmp[0], in[0], p.u[t], p.clov[t], 0.25, isign, t, 0);
    case QPHIX_OP_NDTMWILSON: {
      const double al = 4.0 + p.mass, mu = p.twisted_mass, eps = p.epsilon;
      const double mi = 1.0 / (al * al + mu * mu - eps * eps);
      for (int f = 0; f < 2; f++) TRY(qphix_wilson_dslash(D, op->tmp[f], in[f], p.u[o], isign, o));
      const double apimu[2] = {al * mi, -isign * mu * mi};
      TRY(qphix_two_flav_twisted_mass(op->g, apimu, eps * mi, op->tmp, op->tmp2));
      return qphix_tm_two_flav_achimbdpsi(D, res, op->tmp2, in, p.u[t], al, 0.25, eps, isign, t);
    }
    case QPHIX_OP_NDTMCLOVER: {
      for (int f = 0; f < 2; f++) TRY(qphix_wilson_dslash(D, op->tmp[f], in[f], p.u[o], isign, o));
      TRY(qphix_tmclov_two_flav_inverse_clover_term(D, op->tmp2, op->tmp, p.fullclov[o], p.clov_odiag[o], isign));
      return qphix_tmclov_two_flav_achimbdpsi(D, res, op->tmp2, in, p.u[t], p.clov[t], 0.25, p.epsilon, isign, t);
    }
  }
  return fail(QPHIX_EDOMAIN, "unknown operator kind");
}

// the operator on n single-flavour right-hand sides, two multi-RHS launches
extern "C" int qphix_op_apply_multi(qphix_op *op, void *const res[], const void *const in[], int n, int isign,
                                    int t) {
  CHKPTR(op); CHKPTR(res); CHKPTR(in);
  TRY(check_cb(t));
  if (op->nflav != 1) return fail(QPHIX_EDOMAIN, "multi-RHS apply is for single-flavour operators");
  if (n < 1) return fail(QPHIX_EDOMAIN, "number of right-hand sides %d", n);
  while ((int)op->mtmp.size() < n) {
    void *p = nullptr;
    TRY(qphix_alloc_spinor(op->g, &p));
    op->mtmp.push_back(p);
  }
  const int o = 1 - t;
  const qphix_op_params &p = op->p;
  qphix_dslash *D = &op->d;
  void *const *tmp = op->mtmp.data();
  const void *const *ctmp = (const void *const *)op->mtmp.data();
  switch (p.kind) {
    case QPHIX_OP_WILSON: {
      const double al = 4.0 + p.mass;
      TRY(launch_hop_n(D, KD_WILSON, isign, o, tmp, in, nullptr, n, p.u[o], nullptr, 1.0, 0, 0, 0, 0));
      return launch_hop_n(D, KA_WILSON, isign, t, res, ctmp, in, n, p.u[t], nullptr, -0.25 / al, al, 0, 0, 0);
    }
    case QPHIX_OP_CLOVER:
      TRY(launch_hop_n(D, KD_CLOVER, isign, o, tmp, in, nullptr, n, p.u[o], p.invclov[o], 1.0, 0, 0, 0, 0));
      return launch_hop_n(D, KA_CLOVER, isign, t, res, ctmp, in, n, p.u[t], p.clov[t], -0.25, 0, 0, 0,
                          D->p.prec_mass_rho);
    case QPHIX_OP_TMWILSON: {
      const double al = 4.0 + p.mass, mu = p.twisted_mass;
      TRY(launch_hop_n(D, KD_TM, isign, o, tmp, in, nullptr, n, p.u[o], nullptr, 1.0, 0, mu / al,
                       al / (mu * mu + al * al), 0));
      return launch_hop_n(D, KA_TM, isign, t, res, ctmp, in, n, p.u[t], nullptr, -0.25, al, mu / al, 0, 0);
    }
    case QPHIX_OP_TMCLOVER: {
      const int clidx = isign == 1 ? 0 : 1;
      TRY(launch_hop_n(D, KD_FULL, isign, o, tmp, in, nullptr, n, p.u[o], p.fullclov[o][clidx], 1.0, 0, 0, 0, 0));
      return launch_hop_n(D, KA_CLOVER, isign, t, res, ctmp, in, n, p.u[t], p.clov[t], -0.25, 0, 0, 0,
                          D->p.prec_mass_rho);
    }
  }
  return fail(QPHIX_EDOMAIN, "unknown operator kind");
}

// n right-hand sides of one BLAS operation in one launch; out gets n * nred values
static int run_blas_multi(const qphix_geom *g, int op, const MultiBlasArgs &m, double *out) {
  HIPCHK(blas_multi_launch(g->d, op, m, g_partials, g_dres, g_stream));
  const int nr = blas_nred(op);
  if (nr > 0) {
    HIPCHK(hipMemcpyAsync(g_hres, g_dres, (size_t)m.n * nr * sizeof(double), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    if (out)
      for (int k = 0; k < m.n * nr; k++) out[k] = g_hres[k];
  }
  return QPHIX_OK;
}
static MultiBlasArgs margs(int n) {
  MultiBlasArgs m;
  memset(&m, 0, sizeof(m));
  m.n = n;
  return m;
}
extern "C" int qphix_norm2_multi(const qphix_geom *g, const void *const x[], int n, double *n2) {
  ENSURE_INIT();
  CHKPTR(g); CHKPTR(x); CHKPTR(n2);
  if (n < 1 || n > kMaxRhs) return fail(QPHIX_EDOMAIN, "n = %d right-hand sides (max %d)", n, kMaxRhs);
  MultiBlasArgs m = margs(n);
  for (int r = 0; r < n; r++) {
    if (!x[r]) return fail(QPHIX_EINVAL, "null field %d", r);
    m.x[r] = x[r];
  }
  return run_blas_multi(g, B_NORM2, m, n2);
}

// ---------------------------------------------------------------------------
// solvers
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// multigrid-facing operator pieces (linearOp.h:26-35; wilson.h:67-86;
// clover.h:174-251). Fields of a full (unpreconditioned) spinor are passed as
// two checkerboard pointers [cb 0, cb 1].
// ---------------------------------------------------------------------------
static bool op_is_wc(const qphix_op *op) { return op->p.kind == QPHIX_OP_WILSON || op->p.kind == QPHIX_OP_CLOVER; }

// res = -1/2 D in on target_cb (the Wilson hop for both Wilson and clover)
extern "C" int qphix_op_m_offdiag(qphix_op *op, void *const res[], const void *const in[], int isign, int target_cb) {
  ENSURE_INIT();
  CHKPTR(op); CHKPTR(res); CHKPTR(in);
  TRY(check_cb(target_cb));
  if (!op_is_wc(op)) return fail(QPHIX_EINVAL, "M_offdiag not yet implemented for this operator");
  TRY(qphix_wilson_dslash(&op->d, op->tmp[0], in[0], op->p.u[target_cb], isign, target_cb));
  return qphix_axy(op->g, -0.5, op->tmp[0], res[0]);
}

// res = A_cb^-1 in (clover; invclov[cb] required) or in / (4 + m) (Wilson, cb 0
// as the reference asserts)
extern "C" int qphix_op_m_diag_inv(qphix_op *op, void *const res[], const void *const in[], int isign, int cb) {
  ENSURE_INIT();
  CHKPTR(op); CHKPTR(res); CHKPTR(in);
  TRY(check_cb(cb));
  (void)isign;
  if (op->p.kind == QPHIX_OP_WILSON) {
    if (cb != 0) return fail(QPHIX_EDOMAIN, "Wilson M_diag_inv is defined for cb = 0");
    return qphix_axy(op->g, 1.0 / (4.0 + op->p.mass), in[0], res[0]);
  }
  if (op->p.kind != QPHIX_OP_CLOVER) return fail(QPHIX_EINVAL, "M_diag_inv not yet implemented for this operator");
  if (!op->p.invclov[cb]) return fail(QPHIX_EINVAL, "inverse clover term of cb %d not set", cb);
  return qphix_clov_mult(&op->d, res[0], in[0], op->p.invclov[cb]);
}

// clover: res_cb = A_cb in_cb on cb (cb = -1: both checkerboards)
extern "C" int qphix_op_m_diag(qphix_op *op, void *const res[2], const void *const in[2], int isign, int cb) {
  ENSURE_INIT();
  CHKPTR(op); CHKPTR(res); CHKPTR(in);
  (void)isign;
  if (op->p.kind != QPHIX_OP_CLOVER) return fail(QPHIX_EINVAL, "M_diag is defined for the clover operator");
  if (cb != -1) TRY(check_cb(cb));
  for (int c = 0; c < 2; c++) {
    if (cb != -1 && c != cb) continue;
    if (!op->p.clov[c]) return fail(QPHIX_EINVAL, "clover term of cb %d not set", c);
    TRY(qphix_clov_mult(&op->d, res[c], in[c], op->p.clov[c]));
  }
  return QPHIX_OK;
}

// clover: res_cb = A_cb in_cb - 1/2 D in_(1-cb) (+ i rho g5 in_cb), both cb
extern "C" int qphix_op_m_unprec(qphix_op *op, void *const res[2], const void *const in[2], int isign) {
  ENSURE_INIT();
  CHKPTR(op); CHKPTR(res); CHKPTR(in);
  if (op->p.kind != QPHIX_OP_CLOVER) return fail(QPHIX_EINVAL, "M_unprec is defined for the clover operator");
  for (int c = 0; c < 2; c++) {
    if (!op->p.clov[c]) return fail(QPHIX_EINVAL, "clover term of cb %d not set", c);
    TRY(qphix_clov_achimbdpsi(&op->d, res[c], in[1 - c], in[c], op->p.u[c], op->p.clov[c], 0.5, isign, c));
  }
  return QPHIX_OK;
}

// res_cb = term `dir` of the hop (isign +1) of in_(1-cb), both cb
extern "C" int qphix_op_dslash_dir(qphix_op *op, void *const res[2], const void *const in[2], int dir) {
  ENSURE_INIT();
  CHKPTR(op); CHKPTR(res); CHKPTR(in);
  if (!op_is_wc(op)) return fail(QPHIX_EINVAL, "DslashDir is defined for the Wilson and clover operators");
  for (int c = 0; c < 2; c++) TRY(qphix_wilson_dslash_dir(&op->d, res[c], in[1 - c], op->p.u[c], c, dir));
  return QPHIX_OK;
}

typedef std::vector<void *> SV;  // one spinor per flavour

static int alloc_set(const qphix_geom *g, int nf, SV &v, qphix_solver *s) {
  v.assign(nf, nullptr);
  for (int f = 0; f < nf; f++) {
    TRY(qphix_alloc_spinor(g, &v[f]));
    s->work.push_back(v[f]);
  }
  return QPHIX_OK;
}

// multi-flavour BLAS helpers (reductions summed over flavours)
static int mf_norm2(const qphix_geom *g, int nf, const void *const *x, double &n) {
  n = 0;
  for (int f = 0; f < nf; f++) {
    double t;
    TRY(qphix_norm2_spinor(g, x[f], &t));
    n += t;
  }
  return QPHIX_OK;
}
static int mf_cdot(const qphix_geom *g, int nf, const void *const *l, const void *const *r, double o[2]) {
  o[0] = o[1] = 0;
  for (int f = 0; f < nf; f++) {
    double t[2];
    TRY(qphix_inner_product(g, l[f], r[f], t));
    o[0] += t[0];
    o[1] += t[1];
  }
  return QPHIX_OK;
}
// true when every flavour of x is exactly zero (one norm sweep): the solvers
// then take r = rhs instead of applying the operator to the zero vector, the
// common start of a solve from x = 0
static int mf_is_zero(const qphix_geom *g, int nf, const void *const *x, bool &zero) {
  double n;
  TRY(mf_norm2(g, nf, x, n));
  zero = n == 0.0;
  return QPHIX_OK;
}

// <l, r> and |l|^2 in one sweep, summed over flavours
static int mf_cdotnorm(const qphix_geom *g, int nf, const void *const *l, const void *const *r, double o[3]) {
  o[0] = o[1] = o[2] = 0;
  for (int f = 0; f < nf; f++) {
    double t[3];
    TRY(qphix_inner_product_norm(g, l[f], r[f], t));
    for (int k = 0; k < 3; k++) o[k] += t[k];
  }
  return QPHIX_OK;
}
// BiCGStab x/r update that also returns <r0, r_new>: o = {|r|^2, Re, Im}
static int mf_rxupdate_cdot(const qphix_geom *g, int nf, void *const *x, void *const *r, const void *const *t,
                            const void *const *p, const void *const *r0, const double omega[2],
                            const double alpha[2], double o[3]) {
  o[0] = o[1] = o[2] = 0;
  for (int f = 0; f < nf; f++) {
    BlasArgs a = bargs();
    a.x = t[f]; a.y = p[f]; a.z = r0[f]; a.o0 = x[f]; a.o1 = r[f];
    a.a[0] = omega[0]; a.a[1] = omega[1]; a.a[2] = alpha[0]; a.a[3] = alpha[1];
    double v[3];
    TRY(run_blas(g, B_RXUPD_CDOT, a, v));
    for (int k = 0; k < 3; k++) o[k] += v[k];
  }
  return QPHIX_OK;
}
static int op_apply_v(qphix_op *M, SV &res, const void *const *in, int isign, int cb) {
  return qphix_op_apply(M, res.data(), in, isign, cb);
}
static const void *const *cv(const SV &v) { return (const void *const *)v.data(); }

static void cdiv(double *o, const double *a, const double *b) {
  const double den = b[0] * b[0] + b[1] * b[1];
  const double re = (a[0] * b[0] + a[1] * b[1]) / den, im = (a[1] * b[0] - a[0] * b[1]) / den;
  o[0] = re;
  o[1] = im;
}
static void cmul(double *o, const double *a, const double *b) {
  const double re = a[0] * b[0] - a[1] * b[1], im = a[0] * b[1] + a[1] * b[0];
  o[0] = re;
  o[1] = im;
}

extern "C" int qphix_cg_create(qphix_op *M, int max_iters, qphix_solver **s) {
  CHKPTR(M); CHKPTR(s);
  qphix_solver *x = new qphix_solver();
  x->kind = S_CG;
  x->M = M;
  x->max_iters = max_iters;
  x->inner = nullptr;
  x->delta = 0;
  x->mdagm = 0;
  x->max_shifts = 0;
  *s = x;
  return QPHIX_OK;
}
extern "C" int qphix_bicgstab_create(qphix_op *M, int max_iters, qphix_solver **s) {
  TRY(qphix_cg_create(M, max_iters, s));
  (*s)->kind = S_BICGSTAB;
  return QPHIX_OK;
}
extern "C" int qphix_richardson_create(qphix_op *M, qphix_solver *inner, double delta, int max_iters, int mdagm,
                                       qphix_solver **s) {
  CHKPTR(inner);
  TRY(qphix_cg_create(M, max_iters, s));
  (*s)->kind = S_RICHARDSON;
  (*s)->inner = inner;
  (*s)->delta = delta;
  (*s)->mdagm = mdagm;
  if (inner->M->nflav != M->nflav) return fail(QPHIX_EDOMAIN, "inner and outer operators differ in flavours");
  return QPHIX_OK;
}
extern "C" int qphix_minvcg_create(qphix_op *M, int max_iters, int max_shifts, qphix_solver **s) {
  TRY(qphix_cg_create(M, max_iters, s));
  (*s)->kind = S_MINVCG;
  (*s)->max_shifts = max_shifts;
  return QPHIX_OK;
}
extern "C" int qphix_solver_destroy(qphix_solver *s) {
  if (!s) return QPHIX_OK;
  for (void *p : s->work) qphix_free(p);
  for (void *p : s->mwork) qphix_free(p);
  delete s;
  return QPHIX_OK;
}

// invcg.h:88-456.  Counters as the reference's default build (TIMING_CG not
// defined) counts them: site_flops at invcg.h:180,224,308,337,416 and mv_apps
// at :210,294, including the initial M^+M x, which is skipped (not its count)
// when x is exactly zero.  On convergence the reference leaves rsd_sq_final
// unset; here it is the recursively updated |r|^2 (absolute), the quantity the
// convergence test used.
static int run_blas_q(const qphix_geom *g, int op, BlasArgs a, const double *dsc, double *d_out);
static int cg_device_loop(qphix_solver *s, void *x, const SV &r, const SV &p, SV &mp, SV &mmp, int isign, int cb,
                          double cp, double rsd_sq, unsigned long sf, qphix_solve_stats *st);

static int solve_cg(qphix_solver *s, void *const *x, const void *const *rhs, double rsd_target, int isign,
                    int verbose, int cb, int rtype, qphix_solve_stats *st) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  const int nf = M->nflav;
  SV mp, mmp, p, r;
  if (s->work.empty()) {
    TRY(alloc_set(g, nf, mp, s));
    TRY(alloc_set(g, nf, mmp, s));
    TRY(alloc_set(g, nf, p, s));
    TRY(alloc_set(g, nf, r, s));
  } else {
    mp.assign(s->work.begin(), s->work.begin() + nf);
    mmp.assign(s->work.begin() + nf, s->work.begin() + 2 * nf);
    p.assign(s->work.begin() + 2 * nf, s->work.begin() + 3 * nf);
    r.assign(s->work.begin() + 3 * nf, s->work.begin() + 4 * nf);
  }
  const unsigned long sf = 12ul * (unsigned long)nf;  // the reference counts site flops in units of 12 * num_flav
  st->site_flops = 0;
  st->mv_apps = 0;
  st->status = QPHIX_OK;
  double chi_sq;
  TRY(mf_norm2(g, nf, rhs, chi_sq));
  st->site_flops += 4 * sf;
  double rsd_sq = rsd_target * rsd_target;
  if (rtype == QPHIX_RESID_RELATIVE) rsd_sq *= chi_sq;
  double cp = 0;
  bool xz = s->x_zero != 0;
  if (!xz) TRY(mf_is_zero(g, nf, (const void *const *)x, xz));
  if (xz) {
    for (int f = 0; f < nf; f++) TRY(qphix_copy_spinor(g, r[f], rhs[f]));
    cp = chi_sq;
  } else {
    TRY(op_apply_v(M, mp, x, isign, cb));
    TRY(op_apply_v(M, mmp, cv(mp), -isign, cb));
    for (int f = 0; f < nf; f++) {
      double t;
      TRY(qphix_xmy_norm2(g, r[f], rhs[f], mmp[f], &t));
      cp += t;
    }
  }
  st->mv_apps += 2;
  st->site_flops += 6 * sf;
  if (verbose) printf("CG: ||r||^2 = %e target = %e\n", cp, rsd_sq);
  if (cp <= rsd_sq) {
    st->n_iters = 0;
    st->rsd_sq_final = cp / chi_sq;
    return QPHIX_OK;
  }
  for (int f = 0; f < nf; f++) TRY(qphix_copy_spinor(g, p[f], r[f]));
  if (s->dev_mode > 0 && nf == 1 && !verbose) return cg_device_loop(s, x[0], r, p, mp, mmp, isign, cb, cp, rsd_sq, sf, st);
  for (int k = 0; k <= s->max_iters; ++k) {
    const double c = cp;
    TRY(op_apply_v(M, mp, cv(p), isign, cb));
    TRY(op_apply_v(M, mmp, cv(mp), -isign, cb));
    st->mv_apps += 2;
    double d;
    TRY(mf_norm2(g, nf, cv(mp), d));
    st->site_flops += 4 * sf;
    const double a = c / d;
    cp = 0;
    for (int f = 0; f < nf; f++) {
      double t;
      TRY(qphix_rmammp_norm2_rxpap(g, r[f], a, mmp[f], x[f], p[f], &t));
      cp += t;
    }
    st->site_flops += 12 * sf;
    if (verbose) printf("CG: iter %d: r = %e target = %e\n", k, cp, rsd_sq);
    if (cp < rsd_sq) {
      st->n_iters = k;
      st->rsd_sq_final = cp;
      return QPHIX_OK;
    }
    const double b = cp / c;
    for (int f = 0; f < nf; f++) TRY(qphix_aypx(g, b, r[f], p[f]));
    st->site_flops += 4 * sf;
  }
  st->n_iters = s->max_iters;
  st->rsd_sq_final = cp;
  st->status = QPHIX_ENOCONV;
  return QPHIX_OK;
}

// a BLAS step whose scalars come from device memory (dsc; BlasArgs::dsc), its
// reduction (if any) finalized into device memory: kernels only
static int run_blas_q(const qphix_geom *g, int op, BlasArgs a, const double *dsc, double *d_out) {
  a.dsc = dsc;
  int nb = 0;
  HIPCHK(blas_launch(g->d, op, a, g_partials, &nb, g_stream));
  if (blas_nred(op) > 0) HIPCHK(blas_finalize(g_partials, nb, blas_nred(op), d_out, g_stream));
  return QPHIX_OK;
}

struct GraphRun {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~GraphRun() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};
static SolverView *g_bi_view = nullptr;  // pinned host, 2 entries

// Runs `iteration` (kernels only, one solver iteration on g_stream) until the
// device state it publishes to view[posts & 1] says done. The host enqueues
// iteration k + 1 before it reads iteration k's view, so the GPU does not wait
// for the host, except when the residual ratio's trend says k may be the last
// (or the step limit is near): then it waits for k first. The first iteration
// runs eagerly (it tunes any kernel not tuned yet); mode 2 captures the second
// one in a HIP graph and replays it from then on.
static int device_iterations(int mode, int max_steps, const std::function<int()> &iteration,
                             SolverView *view) {
  GraphRun gr;
  HIPCHK(hipEventCreateWithFlags(&gr.ev[0], hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&gr.ev[1], hipEventDisableTiming));
  double prev = -1, prev2 = -1;  // residual ratio of the last two iterations read
  for (int k = 0;; k++) {
    if (mode == 2 && k == 1) {
      HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
      const int rc = iteration();
      const hipError_t e = hipStreamEndCapture(g_stream, &gr.graph);
      if (rc) return rc;
      HIPCHK(e);
      HIPCHK(hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0));
    }
    if (mode == 2 && k >= 1) {
      HIPCHK(hipGraphLaunch(gr.exec, g_stream));
    } else {
      TRY(iteration());
    }
    HIPCHK(hipEventRecord(gr.ev[k & 1], g_stream));
    if (k > 0) {
      HIPCHK(hipEventSynchronize(gr.ev[(k - 1) & 1]));
      if (view[(k - 1) & 1].done) break;
      prev2 = prev;
      prev = view[(k - 1) & 1].progress;
    }
    const bool last_near = (prev > 0 && prev2 > 0 && prev * (prev / prev2) < 1.5) || k + 2 >= max_steps;
    if (last_near) {
      HIPCHK(hipEventSynchronize(gr.ev[k & 1]));
      if (view[k & 1].done) break;
    }
    if (k > max_steps + 1) break;  // the device state is done by then
  }
  return QPHIX_OK;
}

// The BiCGStab iteration (single flavour) with its scalars in device memory
// (BiState): p-update, v = M p, <r0, v>, s-update, t = M r, <t, r> and |t|^2,
// the x/r update with |r|^2 and <r0, r>, between them four single-thread
// kernels with the host loop's scalar arithmetic; no host round trip
// (device_iterations runs it). Iterations after the end only read: the vector
// updates and the hops skip themselves.
static int bicgstab_device_loop(qphix_solver *s, void *x, const SV &r0, const SV &r, const SV &p, SV &v, SV &t,
                                int isign, int cb, double r_norm, double rsd_sq, const double r0r[2],
                                unsigned long sf, qphix_solve_stats *st) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  if (!g_bi_view) HIPCHK(hipHostMalloc(&g_bi_view, 2 * sizeof(SolverView), hipHostMallocCoherent));
  BiState h;
  memset(&h, 0, sizeof(h));
  h.rho_prev[0] = h.alpha[0] = h.omega[0] = 1.0;
  h.ru[1] = r0r[0];
  h.ru[2] = r0r[1];
  h.r_norm = r_norm;
  h.rsd_sq = rsd_sq;
  h.site_flops = st->site_flops;
  h.mv_apps = st->mv_apps;
  h.sf = sf;
  h.max_iters = s->max_iters;
  h.progress = r_norm / rsd_sq;
  h.hview = g_bi_view;
  DevBuf dbs;
  if (hipMalloc(&dbs.p, sizeof(BiState)) != hipSuccess) return fail(QPHIX_ENOMEM, "BiCGStab state allocation");
  BiState *bs = (BiState *)dbs.p;
  HIPCHK(hipMemcpyAsync(bs, &h, sizeof(h), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  struct SkipScope {  // hop launches of this loop skip once the solve is done
    explicit SkipScope(const int *p) { g_hop_skip = p; }
    ~SkipScope() { g_hop_skip = nullptr; }
  } skip_scope(&bs->done);
  auto iteration = [&]() -> int {
    HIPCHK(bi_pre_launch(bs, g_stream));
    BlasArgs a = bargs();
    a.x = r[0]; a.y = v[0]; a.o0 = p[0];
    TRY(run_blas_q(g, B_PUPD, a, bs->pupd, nullptr));
    TRY(op_apply_v(M, v, cv(p), isign, cb));
    a = bargs();
    a.x = r0[0]; a.y = v[0];
    TRY(run_blas_q(g, B_CDOT, a, nullptr, bs->r0v));
    HIPCHK(bi_mid_launch(bs, g_stream));
    a = bargs();
    a.x = v[0]; a.o0 = r[0];
    TRY(run_blas_q(g, B_SUPD, a, bs->supd, nullptr));
    TRY(op_apply_v(M, t, cv(r), isign, cb));
    a = bargs();
    a.x = t[0]; a.y = r[0];
    TRY(run_blas_q(g, B_CDOTNORM, a, nullptr, bs->tr));
    HIPCHK(bi_mid2_launch(bs, g_stream));
    a = bargs();
    a.x = t[0]; a.y = p[0]; a.z = r0[0]; a.o0 = x; a.o1 = r[0];
    TRY(run_blas_q(g, B_RXUPD_CDOT, a, bs->rxupd, bs->ru));
    HIPCHK(bi_post_launch(bs, g_stream));
    return QPHIX_OK;
  };
  TRY(device_iterations(s->dev_mode, s->max_iters + 1, iteration, g_bi_view));
  HIPCHK(hipMemcpyAsync(&h, bs, sizeof(h), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  st->site_flops = h.site_flops;
  st->mv_apps = h.mv_apps;
  st->rsd_sq_final = h.r_norm;
  st->n_iters = h.n_iters;
  st->status = h.status;
  return QPHIX_OK;
}

// The CG iteration on M^+M (single flavour) with its scalars in device memory
// (CgDState): M^+M p, |M p|^2, a (k_cgd_mid), r -= a M^+M p and x += a p with
// |r|^2, the convergence test and b (k_cgd_post), p = b p + r; no host round trip.
static SolverView *g_cg_view = nullptr;  // pinned host, 2 entries
static int cg_device_loop(qphix_solver *s, void *x, const SV &r, const SV &p, SV &mp, SV &mmp, int isign, int cb,
                          double cp, double rsd_sq, unsigned long sf, qphix_solve_stats *st) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  if (!g_cg_view) HIPCHK(hipHostMalloc(&g_cg_view, 2 * sizeof(SolverView), hipHostMallocCoherent));
  CgDState h;
  memset(&h, 0, sizeof(h));
  h.c = cp;
  h.cp = cp;
  h.rsd_sq = rsd_sq;
  h.site_flops = st->site_flops;
  h.mv_apps = st->mv_apps;
  h.sf = sf;
  h.max_iters = s->max_iters;
  h.progress = cp / rsd_sq;
  h.hview = g_cg_view;
  DevBuf dcs;
  if (hipMalloc(&dcs.p, sizeof(CgDState)) != hipSuccess) return fail(QPHIX_ENOMEM, "CG state allocation");
  CgDState *cs = (CgDState *)dcs.p;
  HIPCHK(hipMemcpyAsync(cs, &h, sizeof(h), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  struct SkipScope {  // hop launches of this loop skip once the solve is done
    explicit SkipScope(const int *q) { g_hop_skip = q; }
    ~SkipScope() { g_hop_skip = nullptr; }
  } skip_scope(&cs->done);
  auto iteration = [&]() -> int {
    TRY(op_apply_v(M, mp, cv(p), isign, cb));
    TRY(op_apply_v(M, mmp, cv(mp), -isign, cb));
    BlasArgs a = bargs();
    a.x = mp[0];
    TRY(run_blas_q(g, B_NORM2, a, nullptr, &cs->d_raw));
    HIPCHK(cgd_mid_launch(cs, g_stream));
    a = bargs();
    a.x = mmp[0]; a.y = p[0]; a.o0 = r[0]; a.o1 = x;
    TRY(run_blas_q(g, B_RMAMMP, a, cs->rmammp, &cs->r_raw));
    HIPCHK(cgd_post_launch(cs, g_stream));
    a = bargs();
    a.x = r[0]; a.o0 = p[0];
    TRY(run_blas_q(g, B_AYPX, a, cs->aypx, nullptr));
    return QPHIX_OK;
  };
  TRY(device_iterations(s->dev_mode, s->max_iters + 1, iteration, g_cg_view));
  HIPCHK(hipMemcpyAsync(&h, cs, sizeof(h), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  st->site_flops = h.site_flops;
  st->mv_apps = h.mv_apps;
  st->rsd_sq_final = h.rsd_sq_final;
  st->n_iters = h.n_iters;
  st->status = h.status;
  return QPHIX_OK;
}

// invbicgstab.h:86-262.  Counters as invbicgstab.h:158-240 counts them,
// including the initial M x (skipped, not its count, when x is exactly zero);
// the |r0|^2 of the iteration-0 test is not counted there either.
static int solve_bicgstab(qphix_solver *s, void *const *x, const void *const *rhs, double rsd_target, int isign,
                          int verbose, int cb, int rtype, qphix_solve_stats *st) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  const int nf = M->nflav;
  SV r0, r, p, v, t;
  if (s->work.empty()) {
    TRY(alloc_set(g, nf, r0, s));
    TRY(alloc_set(g, nf, r, s));
    TRY(alloc_set(g, nf, p, s));
    TRY(alloc_set(g, nf, v, s));
    TRY(alloc_set(g, nf, t, s));
  } else {
    SV *sets[5] = {&r0, &r, &p, &v, &t};
    for (int k = 0; k < 5; k++) sets[k]->assign(s->work.begin() + k * nf, s->work.begin() + (k + 1) * nf);
  }
  const unsigned long sf = 12ul * (unsigned long)nf;  // the reference counts site flops in units of 12 * num_flav
  st->site_flops = 0;
  st->mv_apps = 0;
  st->status = QPHIX_OK;
  double rhs_sq;
  TRY(mf_norm2(g, nf, rhs, rhs_sq));
  st->site_flops += 4 * sf;
  double rsd_sq = rsd_target * rsd_target;
  if (rtype == QPHIX_RESID_RELATIVE) rsd_sq *= rhs_sq;
  double r_norm;
  bool xz = s->x_zero != 0;
  if (!xz) TRY(mf_is_zero(g, nf, (const void *const *)x, xz));
  if (xz) {
    for (int f = 0; f < nf; f++) TRY(qphix_copy_spinor(g, r0[f], rhs[f]));
    r_norm = rhs_sq;
  } else {
    TRY(op_apply_v(M, r0, x, isign, cb));
    for (int f = 0; f < nf; f++) TRY(qphix_bicgstab_xmy(g, rhs[f], r0[f]));
    TRY(mf_norm2(g, nf, cv(r0), r_norm));
  }
  st->mv_apps++;
  st->site_flops += 2 * sf;
  if (r_norm < rsd_sq) {
    st->rsd_sq_final = r_norm;
    st->n_iters = 0;
    return QPHIX_OK;
  }
  for (int f = 0; f < nf; f++) {
    TRY(qphix_copy_spinor(g, r[f], r0[f]));
    TRY(qphix_zero_spinor(g, p[f]));
    TRY(qphix_zero_spinor(g, v[f]));
  }
  double rho_prev[2] = {1, 0}, alpha[2] = {1, 0}, omega[2] = {1, 0}, rho[2] = {0, 0};
  bool notconv = true;
  int k;
  // <r0, r>: a separate sweep on the first iteration, afterwards a by-product
  // of the previous iteration's x/r update
  double r0r[2];
  TRY(mf_cdot(g, nf, cv(r0), cv(r), r0r));
  if (s->dev_mode > 0 && nf == 1 && s->max_iters >= 1 && !verbose)
    return bicgstab_device_loop(s, x[0], r0, r, p, v, t, isign, cb, r_norm, rsd_sq, r0r, sf, st);
  for (k = 1; k <= s->max_iters && notconv; k++) {
    rho[0] = r0r[0];
    rho[1] = r0r[1];
    st->site_flops += 8 * sf;
    if (rho[0] == 0 && rho[1] == 0) {
      st->rsd_sq_final = r_norm;
      st->n_iters = k;
      st->status = QPHIX_EBREAKDOWN;
      return QPHIX_OK;
    }
    double t1[2], t2[2], beta[2];
    cdiv(t1, rho, rho_prev);
    cdiv(t2, alpha, omega);
    cmul(beta, t1, t2);
    for (int f = 0; f < nf; f++) TRY(qphix_bicgstab_p_update(g, r[f], p[f], v[f], beta, omega));
    st->site_flops += 16 * sf;
    TRY(op_apply_v(M, v, cv(p), isign, cb));
    st->mv_apps++;
    double tmp[2];
    TRY(mf_cdot(g, nf, cv(r0), cv(v), tmp));
    st->site_flops += 8 * sf;
    if (tmp[0] == 0 && tmp[1] == 0) {
      st->n_iters = k;
      st->rsd_sq_final = r_norm;
      st->status = QPHIX_EBREAKDOWN;
      return QPHIX_OK;
    }
    cdiv(alpha, rho, tmp);
    rho_prev[0] = rho[0];
    rho_prev[1] = rho[1];
    for (int f = 0; f < nf; f++) TRY(qphix_bicgstab_s_update(g, alpha, r[f], v[f]));
    st->site_flops += 8 * sf;
    TRY(op_apply_v(M, t, cv(r), isign, cb));
    st->mv_apps++;
    double tr[3];  // <t, r> and |t|^2 in one sweep
    TRY(mf_cdotnorm(g, nf, cv(t), cv(r), tr));
    st->site_flops += 4 * sf;
    const double t_norm = tr[2];
    if (t_norm == 0) {
      st->rsd_sq_final = r_norm;
      st->n_iters = k;
      st->status = QPHIX_EBREAKDOWN;
      return QPHIX_OK;
    }
    st->site_flops += 8 * sf;
    omega[0] = tr[0] / t_norm;
    omega[1] = tr[1] / t_norm;
    double ru[3];
    TRY(mf_rxupdate_cdot(g, nf, x, r.data(), cv(t), cv(p), cv(r0), omega, alpha, ru));
    st->site_flops += 28 * sf;
    r_norm = ru[0];
    r0r[0] = ru[1];
    r0r[1] = ru[2];
    if (verbose) printf("BICGSTAB: iter %d r_norm = %e target = %e\n", k, r_norm, rsd_sq);
    if (r_norm < rsd_sq) notconv = false;
  }
  st->rsd_sq_final = r_norm;
  st->n_iters = k;
  if (notconv) st->status = QPHIX_ENOCONV;
  return QPHIX_OK;
}

// invmr.h:71-221 (InvMRBase::operator(); InvMR stops on the residual,
// InvMRSmoother runs max_iters iterations and reports rsd_sq_final = -1);
// counters as invmr.h:95-182 counts them, the initial M x included
static int solve_mr(qphix_solver *s, void *const *x, const void *const *rhs, double rsd_target, int isign,
                    int verbose, int cb, int rtype, qphix_solve_stats *st) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  const int nf = M->nflav;
  SV mr, r;
  if (s->work.empty()) {
    TRY(alloc_set(g, nf, mr, s));
    TRY(alloc_set(g, nf, r, s));
  } else {
    mr.assign(s->work.begin(), s->work.begin() + nf);
    r.assign(s->work.begin() + nf, s->work.begin() + 2 * nf);
  }
  const unsigned long sf = 12ul * (unsigned long)nf;  // the reference counts site flops in units of 12 * num_flav
  st->site_flops = 0;
  st->mv_apps = 0;
  st->status = QPHIX_OK;
  if (s->max_iters < 0) return fail(QPHIX_EDOMAIN, "MR: invalid MaxIters = %d", s->max_iters);
  if (s->max_iters == 0) {
    st->n_iters = 0;
    st->rsd_sq_final = -1;
    return QPHIX_OK;
  }
  bool xz;
  TRY(mf_is_zero(g, nf, (const void *const *)x, xz));
  if (xz) {  // r = rhs - M x
    for (int f = 0; f < nf; f++) TRY(qphix_copy_spinor(g, r[f], rhs[f]));
  } else {
    TRY(op_apply_v(M, r, x, isign, cb));
    for (int f = 0; f < nf; f++) TRY(qphix_bicgstab_xmy(g, rhs[f], r[f]));
  }
  st->mv_apps++;
  st->site_flops += 2 * sf;
  const bool term = s->mr_terminate != 0;
  double norm_chi = 0, rsd_sq = 0, cp = 0;
  if (term) {
    TRY(mf_norm2(g, nf, rhs, norm_chi));
    rsd_sq = rsd_target * rsd_target * (rtype == QPHIX_RESID_RELATIVE ? norm_chi : 1.0);
    TRY(mf_norm2(g, nf, cv(r), cp));
    st->site_flops += 8 * sf;
    if (verbose) printf("QPhiX MR: iter=0: || r ||^2 = %16.8e Target || r ||^2 = %16.8e\n", cp, rsd_sq);
    if (cp < rsd_sq) {
      st->n_iters = 0;
      st->rsd_sq_final = rtype == QPHIX_RESID_RELATIVE ? cp / norm_chi : cp;
      return QPHIX_OK;
    }
  }
  int k = 0;
  bool more = true;
  while (more) {
    ++k;
    TRY(op_apply_v(M, mr, cv(r), isign, cb));
    st->mv_apps++;
    double cd[3];  // <Mr, r>, |Mr|^2
    TRY(mf_cdotnorm(g, nf, cv(mr), cv(r), cd));
    st->site_flops += 12 * sf;
    const double a[2] = {s->omega * cd[0] / cd[2], s->omega * cd[1] / cd[2]};
    for (int f = 0; f < nf; f++) TRY(qphix_mr_rxupdate(g, x[f], r[f], mr[f], a));
    st->site_flops += 16 * sf;
    if (term) {
      TRY(mf_norm2(g, nf, cv(r), cp));
      st->site_flops += 4 * sf;
      if (verbose) printf("QPhiX MR: iter=%d: || r ||^2 = %16.8e Target || r ||^2 = %16.8e\n", k, cp, rsd_sq);
      more = k < s->max_iters && cp > rsd_sq;
    } else {
      more = k < s->max_iters;
    }
  }
  st->n_iters = k;
  if (term) {
    st->rsd_sq_final = rtype == QPHIX_RESID_RELATIVE ? cp / norm_chi : cp;
    if (cp > rsd_sq) st->status = QPHIX_ENOCONV;
  } else {
    st->rsd_sq_final = -1;
  }
  return QPHIX_OK;
}

// inv_richardson_multiprec.h:42-239; counters as :55-228 counts them (outer
// site flops 47 per norm, 71 per x - y norm, 96 per x/r update, plus the inner
// solver's site flops times num_flav; the initial M x is counted also when x
// is exactly zero and its product is skipped)
static int solve_richardson(qphix_solver *s, void *const *x, const void *const *rhs, double rsd_target, int isign,
                            int verbose, int cb, int rtype, qphix_solve_stats *st) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  const qphix_geom *gi = s->inner->M->g;
  const int nf = M->nflav;
  SV r, dx, tmp, tmp2, ri, dxi;
  if (s->work.empty()) {
    TRY(alloc_set(g, nf, r, s));
    TRY(alloc_set(g, nf, dx, s));
    TRY(alloc_set(g, nf, tmp, s));
    TRY(alloc_set(g, nf, tmp2, s));
    TRY(alloc_set(gi, nf, ri, s));
    TRY(alloc_set(gi, nf, dxi, s));
  } else {
    SV *sets[6] = {&r, &dx, &tmp, &tmp2, &ri, &dxi};
    for (int k = 0; k < 6; k++) sets[k]->assign(s->work.begin() + k * nf, s->work.begin() + (k + 1) * nf);
  }
  unsigned long mv_outer = 0, mv_inner = 0, fl_inner = 0, fl_outer = 0;
  st->status = QPHIX_OK;
  double rhs_sq, r_norm_sq;
  TRY(mf_norm2(g, nf, rhs, rhs_sq));
  fl_outer += 24 + 23;
  double rsd_t = rsd_target * rsd_target;
  if (rtype == QPHIX_RESID_RELATIVE) rsd_t *= rhs_sq;
  auto apply_outer = [&](const void *const *in) -> int {
    if (s->mdagm) {
      TRY(op_apply_v(M, tmp2, in, isign, cb));
      TRY(op_apply_v(M, tmp, cv(tmp2), -isign, cb));
      mv_outer += 2;
    } else {
      TRY(op_apply_v(M, tmp, in, isign, cb));
      mv_outer++;
    }
    return QPHIX_OK;
  };
  auto true_resid = [&]() -> int {
    TRY(apply_outer(x));
    r_norm_sq = 0;
    for (int f = 0; f < nf; f++) {
      double t;
      TRY(qphix_xmy_norm2(g, r[f], rhs[f], tmp[f], &t));
      r_norm_sq += t;
    }
    fl_outer += 24 + 24 + 23;
    return QPHIX_OK;
  };
  bool xz;
  TRY(mf_is_zero(g, nf, (const void *const *)x, xz));
  if (xz) {  // r = rhs - M 0
    for (int f = 0; f < nf; f++) TRY(qphix_copy_spinor(g, r[f], rhs[f]));
    r_norm_sq = rhs_sq;
    mv_outer += s->mdagm ? 2 : 1;
    fl_outer += 24 + 24 + 23;
  } else {
    TRY(true_resid());
  }
  if (verbose) printf("RICHARDSON: Initial || r ||^2 = %16.8e Target || r ||^2 = %16.8e\n", r_norm_sq, rsd_t);
  if (r_norm_sq < rsd_t) {
    st->n_iters = 0;
    st->rsd_sq_final = r_norm_sq;
    st->mv_apps = mv_outer;
    st->site_flops = fl_outer;
    return QPHIX_OK;
  }
  int iter;
  for (iter = 1; iter < s->max_iters; iter++) {
    const double delta_finish = sqrt(rsd_t / r_norm_sq);
    const double delta_use = delta_finish > s->delta ? delta_finish : s->delta;
    const double r_norm = sqrt(r_norm_sq);
    for (int f = 0; f < nf; f++) {
      TRY(qphix_convert_spinor(gi, ri[f], 1.0 / r_norm, g, r[f]));
      TRY(qphix_zero_spinor(gi, dxi[f]));
    }
    qphix_solve_stats ist;
    s->inner->x_zero = 1;  // dxi was just zeroed
    const int ie = qphix_solve(s->inner, dxi.data(), cv(ri), delta_use, isign, 0, cb, QPHIX_RESID_RELATIVE, &ist);
    s->inner->x_zero = 0;
    TRY(ie);
    mv_inner += ist.mv_apps;
    fl_inner += ist.site_flops;
    for (int f = 0; f < nf; f++) TRY(qphix_convert_spinor(g, dx[f], r_norm, gi, dxi[f]));
    TRY(apply_outer(cv(dx)));
    r_norm_sq = 0;
    for (int f = 0; f < nf; f++) {
      double t;
      TRY(qphix_richardson_rxupdate_norm_r(g, x[f], r[f], dx[f], tmp[f], &t));
      r_norm_sq += t;
    }
    fl_outer += 8 * 12;
    if (verbose)
      printf("RICHARDSON: Iter %d inner iters %d r_norm_sq = %e Target = %e\n", iter, ist.n_iters, r_norm_sq, rsd_t);
    if (r_norm_sq < rsd_t) {
      TRY(true_resid());
      st->rsd_sq_final = r_norm_sq / rhs_sq;
      st->n_iters = iter;
      st->mv_apps = mv_outer + mv_inner;
      st->site_flops = fl_outer + fl_inner * nf;
      return QPHIX_OK;
    }
  }
  TRY(true_resid());
  st->rsd_sq_final = r_norm_sq / rhs_sq;
  st->n_iters = iter;
  st->mv_apps = mv_outer + mv_inner;
  st->site_flops = fl_outer + fl_inner * nf;
  st->status = QPHIX_ENOCONV;
  return QPHIX_OK;
}

extern "C" int qphix_solve(qphix_solver *s, void *const x[], const void *const rhs[], double rsd_target, int isign,
                           int verbose, int target_cb, int resid_type, qphix_solve_stats *st) {
  ENSURE_INIT();
  CHKPTR(s); CHKPTR(x); CHKPTR(rhs); CHKPTR(st);
  TRY(check_cb(target_cb));
  if (isign != 1 && isign != -1) return fail(QPHIX_EDOMAIN, "isign must be +1 or -1");
  memset(st, 0, sizeof(*st));
  switch (s->kind) {
    case S_CG: return solve_cg(s, x, rhs, rsd_target, isign, verbose, target_cb, resid_type, st);
    case S_BICGSTAB: return solve_bicgstab(s, x, rhs, rsd_target, isign, verbose, target_cb, resid_type, st);
    case S_RICHARDSON: return solve_richardson(s, x, rhs, rsd_target, isign, verbose, target_cb, resid_type, st);
    case S_MR: return solve_mr(s, x, rhs, rsd_target, isign, verbose, target_cb, resid_type, st);
  }
  return fail(QPHIX_EINVAL, "qphix_solve on a multishift solver; use qphix_minvcg_solve");
}

// CG on M^+M for n right-hand sides advanced in lockstep (invcg.h:88-456 per
// right-hand side): every operator application and BLAS sweep is one launch for
// all still-unconverged right-hand sides; converged ones drop out.
static int solve_cg_multi(qphix_solver *s, void *const *x, const void *const *rhs, int n, double rsd_target,
                          int isign, int verbose, int cb, int rtype, qphix_solve_stats *st) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  while ((int)s->mwork.size() < 4 * n) {
    void *p = nullptr;
    TRY(qphix_alloc_spinor(g, &p));
    s->mwork.push_back(p);
  }
  void **mp = &s->mwork[0], **mmp = &s->mwork[n], **pp = &s->mwork[2 * n], **rr = &s->mwork[3 * n];
  std::vector<double> chi_sq(n), rsd_sq(n), cp(n), c(n), d(n), tmp(n);
  std::vector<int> act;
  for (int r = 0; r < n; r++) {
    memset(&st[r], 0, sizeof(st[r]));
  }
  auto sub = [&](void *const *v, const std::vector<int> &ix, std::vector<void *> &o) {
    o.clear();
    for (int i : ix) o.push_back(v[i]);
  };
  std::vector<int> all(n);
  for (int r = 0; r < n; r++) all[r] = r;
  // chi_sq, initial residual
  {
    MultiBlasArgs m = margs(n);
    for (int r = 0; r < n; r++) m.x[r] = rhs[r];
    TRY(run_blas_multi(g, B_NORM2, m, chi_sq.data()));
  }
  for (int r = 0; r < n; r++) rsd_sq[r] = rsd_target * rsd_target * (rtype == QPHIX_RESID_RELATIVE ? chi_sq[r] : 1.0);
  auto mdagm = [&](const std::vector<int> &ix, void *const *in) -> int {
    std::vector<void *> vi, vm, vmm;
    sub(in, ix, vi);
    sub(mp, ix, vm);
    sub(mmp, ix, vmm);
    TRY(qphix_op_apply_multi(M, vm.data(), (const void *const *)vi.data(), (int)ix.size(), isign, cb));
    TRY(qphix_op_apply_multi(M, vmm.data(), (const void *const *)vm.data(), (int)ix.size(), -isign, cb));
    for (int i : ix) st[i].mv_apps += 2;
    return QPHIX_OK;
  };
  auto true_resid = [&](const std::vector<int> &ix) -> int {  // rr = rhs - M^+M x, cp
    TRY(mdagm(ix, x));
    MultiBlasArgs m = margs((int)ix.size());
    for (size_t k = 0; k < ix.size(); k++) {
      m.o0[k] = rr[ix[k]];
      m.x[k] = rhs[ix[k]];
      m.y[k] = mmp[ix[k]];
    }
    TRY(run_blas_multi(g, B_XMYNORM2, m, tmp.data()));
    for (size_t k = 0; k < ix.size(); k++) cp[ix[k]] = tmp[k];
    return QPHIX_OK;
  };
  TRY(true_resid(all));
  const unsigned long sf = 12ul * (unsigned long)M->nflav;  // per right-hand side, as solve_cg counts
  for (int r = 0; r < n; r++) {
    st[r].site_flops += (4 + 6) * sf;
    if (cp[r] <= rsd_sq[r]) {
      st[r].n_iters = 0;
      st[r].rsd_sq_final = cp[r] / chi_sq[r];
    } else {
      act.push_back(r);
    }
  }
  {
    MultiBlasArgs m = margs((int)act.size());
    for (size_t k = 0; k < act.size(); k++) {
      m.o0[k] = pp[act[k]];
      m.x[k] = rr[act[k]];
    }
    if (!act.empty()) TRY(run_blas_multi(g, B_COPY, m, nullptr));
  }
  for (int k = 0; k <= s->max_iters && !act.empty(); ++k) {
    for (int i : act) c[i] = cp[i];
    TRY(mdagm(act, pp));
    MultiBlasArgs m = margs((int)act.size());
    for (size_t q = 0; q < act.size(); q++) m.x[q] = mp[act[q]];
    TRY(run_blas_multi(g, B_NORM2, m, tmp.data()));
    m = margs((int)act.size());
    for (size_t q = 0; q < act.size(); q++) {
      const int i = act[q];
      m.o0[q] = rr[i];
      m.x[q] = mmp[i];
      m.o1[q] = x[i];
      m.y[q] = pp[i];
      m.a[q] = c[i] / tmp[q];
    }
    TRY(run_blas_multi(g, B_RMAMMP, m, tmp.data()));
    std::vector<int> still;
    for (size_t q = 0; q < act.size(); q++) {
      const int i = act[q];
      cp[i] = tmp[q];
      st[i].site_flops += (4 + 12) * sf;
      if (verbose) printf("CG[%d]: iter %d: r = %e target = %e\n", i, k, cp[i], rsd_sq[i]);
      if (cp[i] < rsd_sq[i]) {
        st[i].n_iters = k;
        st[i].rsd_sq_final = cp[i];
      } else {
        st[i].site_flops += 4 * sf;
        still.push_back(i);
      }
    }
    act = still;
    if (act.empty()) break;
    m = margs((int)act.size());
    for (size_t q = 0; q < act.size(); q++) {
      const int i = act[q];
      m.o0[q] = pp[i];
      m.x[q] = rr[i];
      m.a[q] = cp[i] / c[i];
    }
    TRY(run_blas_multi(g, B_AYPX, m, nullptr));
  }
  for (int i : act) {
    st[i].n_iters = s->max_iters;
    st[i].rsd_sq_final = cp[i];
    st[i].status = QPHIX_ENOCONV;
  }
  return QPHIX_OK;
}

// ---------------------------------------------------------------------------
// Block CG on M^+M (no reference counterpart; the reference solves right-hand
// sides one at a time with invcg.h). The n right-hand sides share one block
// Krylov space. Form: block CG with the residual block kept orthonormal by a
// Cholesky QR every iteration (Dubrulle's retooled block CG), which avoids the
// ill-conditioned Gram matrices of the plain O'Leary recurrence:
//   R = B - A X = Q C,  P = Q
//   repeat: Z = A P;  xi = (P^H Z)^-1;  X += P xi C;  Q zeta = Q - Z xi (QR);
//           P = Q + P zeta^H;  C = zeta C
// so that R = Q C throughout and the residual norm of right-hand side j is the
// norm of column j of C. The n x n algebra runs on the host in FP64; Gram
// matrices are FP64 reductions on the device.
// ---------------------------------------------------------------------------
typedef std::complex<double> cd;
struct CMat {  // n x n complex, row-major
  int n;
  std::vector<cd> v;
  explicit CMat(int n_ = 0) : n(n_), v((size_t)n_ * n_) {}
  cd &operator()(int i, int j) { return v[(size_t)i * n + j]; }
  cd operator()(int i, int j) const { return v[(size_t)i * n + j]; }
};
static CMat cm_mul(const CMat &a, const CMat &b) {
  CMat c(a.n);
  for (int i = 0; i < a.n; i++)
    for (int k = 0; k < a.n; k++) {
      const cd x = a(i, k);
      for (int j = 0; j < a.n; j++) c(i, j) += x * b(k, j);
    }
  return c;
}
static CMat cm_adj(const CMat &a) {
  CMat c(a.n);
  for (int i = 0; i < a.n; i++)
    for (int j = 0; j < a.n; j++) c(i, j) = std::conj(a(j, i));
  return c;
}
// upper-triangular U with G = U^H U (G Hermitian; its lower triangle is used
// after symmetrisation); false if G is not numerically positive definite
static bool cm_chol_upper(const CMat &g, CMat &u) {
  const int n = g.n;
  u = CMat(n);
  for (int j = 0; j < n; j++) {
    double d = g(j, j).real();
    for (int k = 0; k < j; k++) d -= std::norm(u(k, j));
    if (!(d > 0)) return false;
    const double r = sqrt(d);
    u(j, j) = r;
    for (int i = j + 1; i < n; i++) {
      cd x = 0.5 * (g(j, i) + std::conj(g(i, j)));
      for (int k = 0; k < j; k++) x -= std::conj(u(k, j)) * u(k, i);
      u(j, i) = x / r;
    }
  }
  return true;
}
static CMat cm_inv_upper(const CMat &u) {
  const int n = u.n;
  CMat x(n);
  for (int j = 0; j < n; j++) {
    x(j, j) = 1.0 / u(j, j);
    for (int i = j - 1; i >= 0; i--) {
      cd s = 0;
      for (int k = i + 1; k <= j; k++) s += u(i, k) * x(k, j);
      x(i, j) = -s / u(i, i);
    }
  }
  return x;
}

static double *g_gram_part = nullptr;  // device, kGramPartials doubles
static double *g_gram_dev = nullptr;   // device, one Gram matrix
static double *g_gram_host = nullptr;  // pinned host, one Gram matrix
static double *g_mat_dev = nullptr;    // device, 4 matrix slots
static double *g_mat_host = nullptr;   // pinned host, 4 matrix slots
static int g_mat_slot = 0;
constexpr size_t kMatDoubles = (size_t)kMaxRhs * kMaxRhs * 2;

static int block_buffers() {
  if (g_gram_part) return QPHIX_OK;
  HIPCHK(hipMalloc(&g_gram_part, kGramPartials * sizeof(double)));
  HIPCHK(hipMalloc(&g_gram_dev, kMatDoubles * sizeof(double)));
  HIPCHK(hipHostMalloc(&g_gram_host, kMatDoubles * sizeof(double), hipHostMallocDefault));
  HIPCHK(hipMalloc(&g_mat_dev, 4 * kMatDoubles * sizeof(double)));
  HIPCHK(hipHostMalloc(&g_mat_host, 4 * kMatDoubles * sizeof(double), hipHostMallocDefault));
  return QPHIX_OK;
}

// G = A^H B over the fields a[0..n), b[0..n) (synchronises)
static int block_gram(const qphix_geom *g, void *const *a, void *const *b, int n, CMat &G) {
  BlockArgs ba;
  memset(&ba, 0, sizeof(ba));
  for (int i = 0; i < n; i++) {
    ba.a[i] = a[i];
    ba.b[i] = b[i];
  }
  ba.na = ba.nb = n;
  HIPCHK(block_gram_launch(g->d, ba, g_gram_part, g_gram_dev, g_stream));
  HIPCHK(hipMemcpyAsync(g_gram_host, g_gram_dev, (size_t)n * n * 2 * sizeof(double), hipMemcpyDeviceToHost,
                        g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  G = CMat(n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) G(i, j) = cd(g_gram_host[(i * n + j) * 2], g_gram_host[(i * n + j) * 2 + 1]);
  return QPHIX_OK;
}

// o_j = y_j + sum_i a_i M_ij (y null: no addition). Matrix slots cycle over
// four pinned staging buffers; every solver step synchronises (a Gram matrix)
// before a slot is reused.
static int block_combine(const qphix_geom *g, void *const *a, const CMat &M, void *const *y, void *const *o) {
  const int n = M.n;
  double *h = g_mat_host + g_mat_slot * kMatDoubles;
  double *d = g_mat_dev + g_mat_slot * kMatDoubles;
  g_mat_slot = (g_mat_slot + 1) & 3;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      h[(i * n + j) * 2] = M(i, j).real();
      h[(i * n + j) * 2 + 1] = M(i, j).imag();
    }
  HIPCHK(hipMemcpyAsync(d, h, (size_t)n * n * 2 * sizeof(double), hipMemcpyHostToDevice, g_stream));
  BlockArgs ba;
  memset(&ba, 0, sizeof(ba));
  for (int i = 0; i < n; i++) {
    ba.a[i] = a[i];
    ba.o[i] = o[i];
    ba.y[i] = y ? y[i] : nullptr;
  }
  ba.na = ba.nb = n;
  HIPCHK(block_combine_launch(g->d, ba, d, g_stream));
  return QPHIX_OK;
}

// block BLAS entry points (host matrices, synchronous Gram)
extern "C" int qphix_block_gram(const qphix_geom *g, const void *const a[], int na, const void *const b[], int nb,
                                double *out) {
  BLAS_PRE(); CHKPTR(a); CHKPTR(b); CHKPTR(out);
  if (na < 1 || nb < 1 || na > kMaxRhs || nb > kMaxRhs) return fail(QPHIX_EDOMAIN, "block sizes %d x %d", na, nb);
  if (g->d.prec == P16 || g->d.prec == PB16) return fail(QPHIX_EDOMAIN, "block BLAS needs FP64 or FP32 fields");
  TRY(block_buffers());
  BlockArgs ba;
  memset(&ba, 0, sizeof(ba));
  for (int i = 0; i < na; i++) {
    if (!a[i]) return fail(QPHIX_EINVAL, "null field");
    ba.a[i] = a[i];
  }
  for (int j = 0; j < nb; j++) {
    if (!b[j]) return fail(QPHIX_EINVAL, "null field");
    ba.b[j] = b[j];
  }
  ba.na = na;
  ba.nb = nb;
  HIPCHK(block_gram_launch(g->d, ba, g_gram_part, g_gram_dev, g_stream));
  HIPCHK(hipMemcpyAsync(g_gram_host, g_gram_dev, (size_t)na * nb * 2 * sizeof(double), hipMemcpyDeviceToHost,
                        g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  memcpy(out, g_gram_host, (size_t)na * nb * 2 * sizeof(double));
  return QPHIX_OK;
}

extern "C" int qphix_block_combine(const qphix_geom *g, const void *const a[], int na, const double *mat,
                                   const void *const y[], void *const o[], int nb) {
  BLAS_PRE(); CHKPTR(a); CHKPTR(mat); CHKPTR(o);
  if (na < 1 || nb < 1 || na > kMaxRhs || nb > kMaxRhs) return fail(QPHIX_EDOMAIN, "block sizes %d x %d", na, nb);
  if (g->d.prec == P16 || g->d.prec == PB16) return fail(QPHIX_EDOMAIN, "block BLAS needs FP64 or FP32 fields");
  TRY(block_buffers());
  BlockArgs ba;
  memset(&ba, 0, sizeof(ba));
  for (int i = 0; i < na; i++) {
    if (!a[i]) return fail(QPHIX_EINVAL, "null field");
    ba.a[i] = a[i];
  }
  for (int j = 0; j < nb; j++) {
    if (!o[j]) return fail(QPHIX_EINVAL, "null field");
    ba.o[j] = o[j];
    ba.y[j] = y ? y[j] : nullptr;
  }
  ba.na = na;
  ba.nb = nb;
  // synchronous staging of the matrix (the call may reuse mat at once)
  HIPCHK(hipStreamSynchronize(g_stream));
  memcpy(g_mat_host, mat, (size_t)na * nb * 2 * sizeof(double));
  HIPCHK(hipMemcpyAsync(g_mat_dev, g_mat_host, (size_t)na * nb * 2 * sizeof(double), hipMemcpyHostToDevice,
                        g_stream));
  HIPCHK(block_combine_launch(g->d, ba, g_mat_dev, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return QPHIX_OK;
}

static int solve_block_cg(qphix_solver *s, void *const *x, const void *const *rhs, int n, double rsd_target,
                          int isign, int verbose, int cb, int rtype, qphix_solve_stats *st) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  TRY(block_buffers());
  while ((int)s->mwork.size() < 4 * n) {
    void *p = nullptr;
    TRY(qphix_alloc_spinor(g, &p));
    s->mwork.push_back(p);
  }
  void **mp = &s->mwork[0], **zz = &s->mwork[n], **pp = &s->mwork[2 * n], **qq = &s->mwork[3 * n];
  for (int r = 0; r < n; r++) memset(&st[r], 0, sizeof(st[r]));
  std::vector<double> chi_sq(n), rsd_sq(n), tmp(n);
  {
    MultiBlasArgs m = margs(n);
    for (int r = 0; r < n; r++) m.x[r] = rhs[r];
    TRY(run_blas_multi(g, B_NORM2, m, chi_sq.data()));
  }
  for (int r = 0; r < n; r++) {
    if (chi_sq[r] == 0) return fail(QPHIX_EDOMAIN, "right-hand side %d is zero", r);
    rsd_sq[r] = rsd_target * rsd_target * (rtype == QPHIX_RESID_RELATIVE ? chi_sq[r] : 1.0);
  }
  auto mdagm = [&](void *const *in, void **out) -> int {
    TRY(qphix_op_apply_multi(M, mp, (const void *const *)in, n, isign, cb));
    TRY(qphix_op_apply_multi(M, out, (const void *const *)mp, n, -isign, cb));
    for (int r = 0; r < n; r++) st[r].mv_apps += 2;
    return QPHIX_OK;
  };
  // R = B - A X  (into qq)
  TRY(mdagm(x, zz));
  {
    MultiBlasArgs m = margs(n);
    for (int r = 0; r < n; r++) {
      m.o0[r] = qq[r];
      m.x[r] = rhs[r];
      m.y[r] = zz[r];
    }
    TRY(run_blas_multi(g, B_XMYNORM2, m, tmp.data()));
  }
  bool all_done = true;
  for (int r = 0; r < n; r++) all_done = all_done && tmp[r] <= rsd_sq[r];
  if (all_done) {
    for (int r = 0; r < n; r++) st[r].rsd_sq_final = tmp[r] / chi_sq[r];
    return QPHIX_OK;
  }
  // R = Q C
  CMat G, C, U;
  TRY(block_gram(g, qq, qq, n, G));
  if (!cm_chol_upper(G, C)) return fail(QPHIX_EBREAKDOWN, "block CG: residual block is rank deficient");
  TRY(block_combine(g, qq, cm_inv_upper(C), nullptr, qq));
  {
    MultiBlasArgs m = margs(n);
    for (int r = 0; r < n; r++) {
      m.o0[r] = pp[r];
      m.x[r] = qq[r];
    }
    TRY(run_blas_multi(g, B_COPY, m, nullptr));
  }
  int k;
  bool conv = false, broke = false;
  std::vector<double> rn(n);
  for (k = 1; k <= s->max_iters; k++) {
    TRY(mdagm(pp, zz));
    TRY(block_gram(g, pp, zz, n, G));  // P^H A P
    CMat L;
    if (!cm_chol_upper(G, L)) {
      broke = true;
      break;
    }
    const CMat Li = cm_inv_upper(L);
    const CMat xi = cm_mul(Li, cm_adj(Li));  // (P^H A P)^-1 = L^-1 L^-H
    TRY(block_combine(g, pp, cm_mul(xi, C), x, x));
    CMat mxi = xi;
    for (cd &e : mxi.v) e = -e;
    TRY(block_combine(g, zz, mxi, qq, qq));  // W = Q - Z xi
    TRY(block_gram(g, qq, qq, n, G));
    CMat zeta;
    if (!cm_chol_upper(G, zeta)) {
      broke = true;
      break;
    }
    TRY(block_combine(g, qq, cm_inv_upper(zeta), nullptr, qq));  // Q = W zeta^-1
    TRY(block_combine(g, pp, cm_adj(zeta), qq, pp));              // P = Q + P zeta^H
    C = cm_mul(zeta, C);
    conv = true;
    for (int j = 0; j < n; j++) {
      double c2 = 0;
      for (int i = 0; i < n; i++) c2 += std::norm(C(i, j));
      rn[j] = c2;
      conv = conv && c2 <= rsd_sq[j];
    }
    if (verbose) {
      double worst = 0;
      for (int j = 0; j < n; j++) worst = fmax(worst, rn[j] / rsd_sq[j]);
      printf("BLOCK CG: iter %d worst r^2 / target %e\n", k, worst);
    }
    if (conv) break;
  }
  for (int r = 0; r < n; r++) {
    st[r].n_iters = k;
    st[r].rsd_sq_final = rn[r] / chi_sq[r];
    st[r].status = broke ? QPHIX_EBREAKDOWN : (conv ? QPHIX_OK : QPHIX_ENOCONV);
  }
  return QPHIX_OK;
}

// invmr.h:258-348: smoother = 0 is InvMR, 1 is InvMRSmoother
extern "C" int qphix_mr_create(qphix_op *M, int max_iters, double omega, int smoother, qphix_solver **s) {
  TRY(qphix_cg_create(M, max_iters, s));
  (*s)->kind = S_MR;
  (*s)->omega = omega;
  (*s)->mr_terminate = smoother ? 0 : 1;
  return QPHIX_OK;
}

// ---------------------------------------------------------------------------
// unpreconditioned solves through the even-odd solver
// (unprec_solver_wrapper.h:16-100, abs_solver.h:271-376), single flavour.
// The other-checkerboard inverse diagonal is taken on cb = 1 - solve_cb (the
// reference calls M_diag_inv with its default cb = 0, the same for the
// default solve_cb = 1).
// ---------------------------------------------------------------------------
static int unprec_tmp(qphix_op *M) {
  while (M->utmp.size() < 2) {
    void *p = nullptr;
    TRY(qphix_alloc_spinor(M->g, &p));
    M->utmp.push_back(p);
  }
  return QPHIX_OK;
}

// out = in_cb - M_offdiag(solve_cb) M_diag_inv(1 - solve_cb) in_other_cb
extern "C" int qphix_unprec_source_prepare(qphix_op *M, void *const out[], const void *const in_cb[],
                                           const void *const in_other_cb[], int isign, int solve_cb) {
  ENSURE_INIT();
  CHKPTR(M); CHKPTR(out); CHKPTR(in_cb); CHKPTR(in_other_cb);
  TRY(check_cb(solve_cb));
  if (M->nflav != 1) return fail(QPHIX_EDOMAIN, "even-odd source preparation is for single-flavour operators");
  TRY(unprec_tmp(M));
  void *t1[1] = {M->utmp[0]}, *t2[1] = {M->utmp[1]};
  TRY(qphix_copy_spinor(M->g, out[0], in_cb[0]));
  TRY(qphix_op_m_diag_inv(M, t1, in_other_cb, isign, 1 - solve_cb));
  TRY(qphix_op_m_offdiag(M, t2, (const void *const *)t1, isign, solve_cb));
  return qphix_axpy(M->g, -1.0, t2[0], out[0]);
}

// out_other_cb = M_diag_inv(1 - solve_cb) (in_other_cb - M_offdiag(1 - solve_cb) out_cb)
extern "C" int qphix_unprec_solution_reconstruct(qphix_op *M, void *const out_other_cb[],
                                                 const void *const in_other_cb[], const void *const out_cb[],
                                                 int isign, int solve_cb) {
  ENSURE_INIT();
  CHKPTR(M); CHKPTR(out_other_cb); CHKPTR(in_other_cb); CHKPTR(out_cb);
  TRY(check_cb(solve_cb));
  if (M->nflav != 1) return fail(QPHIX_EDOMAIN, "even-odd solution reconstruction is for single-flavour operators");
  TRY(unprec_tmp(M));
  void *t1[1] = {M->utmp[0]}, *t2[1] = {M->utmp[1]};
  TRY(qphix_copy_spinor(M->g, t2[0], in_other_cb[0]));
  TRY(qphix_op_m_offdiag(M, t1, out_cb, isign, 1 - solve_cb));
  TRY(qphix_axpy(M->g, -1.0, t1[0], t2[0]));
  return qphix_op_m_diag_inv(M, out_other_cb, (const void *const *)t2, isign, 1 - solve_cb);
}

// x and rhs are full spinors [cb 0, cb 1]; the even-odd solver works on solve_cb
extern "C" int qphix_unprec_solve(qphix_solver *eo, void *const x[2], const void *const rhs[2], double rsd_target,
                                  int isign, int verbose, int solve_cb, int resid_type, qphix_solve_stats *st) {
  ENSURE_INIT();
  CHKPTR(eo); CHKPTR(x); CHKPTR(rhs); CHKPTR(st);
  TRY(check_cb(solve_cb));
  qphix_op *M = eo->M;
  void *out_other[1] = {x[1 - solve_cb]}, *out_cb[1] = {x[solve_cb]};
  const void *in_cb[1] = {rhs[solve_cb]}, *in_other[1] = {rhs[1 - solve_cb]};
  // the prepared source goes into the other-checkerboard half of x
  TRY(qphix_unprec_source_prepare(M, out_other, in_cb, in_other, isign, solve_cb));
  TRY(qphix_solve(eo, out_cb, (const void *const *)out_other, rsd_target, isign, verbose, solve_cb, resid_type, st));
  return qphix_unprec_solution_reconstruct(M, out_other, in_other, (const void *const *)out_cb, isign, solve_cb);
}

extern "C" int qphix_block_cg_create(qphix_op *M, int max_iters, qphix_solver **s) {
  TRY(qphix_cg_create(M, max_iters, s));
  (*s)->kind = S_BLOCKCG;
  return QPHIX_OK;
}

extern "C" int qphix_solve_multi(qphix_solver *s, void *const x[], const void *const rhs[], int n, double rsd_target,
                                 int isign, int verbose, int target_cb, int resid_type, qphix_solve_stats *st) {
  ENSURE_INIT();
  CHKPTR(s); CHKPTR(x); CHKPTR(rhs); CHKPTR(st);
  TRY(check_cb(target_cb));
  if (isign != 1 && isign != -1) return fail(QPHIX_EDOMAIN, "isign must be +1 or -1");
  if (s->kind != S_CG && s->kind != S_BLOCKCG)
    return fail(QPHIX_EINVAL, "multi-right-hand-side solves use a CG or block CG solver");
  if (s->M->nflav != 1) return fail(QPHIX_EDOMAIN, "multi-right-hand-side CG is for single-flavour operators");
  if (n < 1) return fail(QPHIX_EDOMAIN, "number of right-hand sides %d", n);
  if (s->kind == S_BLOCKCG) {
    if (n > kMaxRhs) return fail(QPHIX_EDOMAIN, "block CG takes at most %d right-hand sides, got %d", kMaxRhs, n);
    if (s->M->g->d.prec == P16 || s->M->g->d.prec == PB16)
      return fail(QPHIX_EDOMAIN, "block CG needs FP64 or FP32 fields");
    return solve_block_cg(s, x, rhs, n, rsd_target, isign, verbose, target_cb, resid_type, st);
  }
  for (int r0 = 0; r0 < n; r0 += kMaxRhs) {
    const int m = (n - r0) < kMaxRhs ? (n - r0) : kMaxRhs;
    TRY(solve_cg_multi(s, x + r0, rhs + r0, m, rsd_target, isign, verbose, target_cb, resid_type, st + r0));
  }
  return QPHIX_OK;
}

// The multishift iteration with its scalars in device memory (MsState): per
// iteration the kernels a / as_s (k_ms_pre), the vector sweep, M^+M p0, |M p0|^2,
// b (k_ms_mid), r += b M^+M p0 with |r|^2, then the zeta / beta recurrences and
// the convergence test (k_ms_post); no host round trip (device_iterations runs
// it, eagerly or replayed from a HIP graph). Once done the state is frozen; a
// closing sweep applies the x_s updates left pending. Returns 0, -1 (not
// converged) or a status.
static SolverView *g_ms_view = nullptr;  // pinned host, 2 entries
static int device_iterations(int mode, int max_steps, const std::function<int()> &iteration,
                             SolverView *view);

static int minvcg_device_loop(qphix_solver *s, void *const x[], int n_shift, const double *shifts, int isign, int cb,
                              const std::vector<void *> &r, const std::vector<void *> &p0,
                              std::vector<void *> &mp, std::vector<void *> &mmp,
                              const std::vector<std::vector<void *>> &p, const std::vector<double> &zeta,
                              const std::vector<double> &zeta_prev, const std::vector<double> &beta,
                              const std::vector<double> &rsd_sq, double b, double c, double cp, unsigned long fl,
                              double *rsd_sq_final, int *n_iters, unsigned long *site_flops, unsigned long *mv_apps,
                              int *k_out) {
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  const int nf = M->nflav;
  if (!g_ms_view) HIPCHK(hipHostMalloc(&g_ms_view, 2 * sizeof(SolverView), hipHostMallocCoherent));
  MsState h;
  memset(&h, 0, sizeof(h));
  h.b = b;
  h.c = c;
  h.cp = cp;
  h.nshift = n_shift;
  h.nf = nf;
  h.max_iters = s->max_iters;
  h.done = s->max_iters <= 1;
  h.site_flops = fl;
  h.hview = g_ms_view;
  for (int q = 0; q < n_shift; q++) {
    h.shifts[q] = shifts[q];
    h.zeta[q] = zeta[q];
    h.zeta_prev[q] = zeta_prev[q];
    h.beta[q] = beta[q];
    h.rsd_sq[q] = rsd_sq[q];
  }
  DevBuf dms;
  if (hipMalloc(&dms.p, sizeof(MsState)) != hipSuccess) return fail(QPHIX_ENOMEM, "multishift state allocation");
  MsState *ms = (MsState *)dms.p;
  HIPCHK(hipMemcpyAsync(ms, &h, sizeof(h), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  auto sweep = [&]() -> int {
    for (int f = 0; f < nf; f++) {
      MShiftArgs m;
      memset(&m, 0, sizeof(m));
      m.r = r[f];
      m.p0 = p0[f];
      m.nshift = n_shift;
      for (int q = 0; q < n_shift; q++) {
        m.p[q] = p[q][f];
        m.x[q] = x[q * nf + f];
      }
      HIPCHK(mshift_dev_launch(g->d, m, ms, g_stream));
    }
    return QPHIX_OK;
  };
  struct SkipScope {  // hop launches of this loop skip once the solve is done
    explicit SkipScope(const int *p) { g_hop_skip = p; }
    ~SkipScope() { g_hop_skip = nullptr; }
  } skip_scope(&ms->done);
  auto iteration = [&]() -> int {
    TRY(sweep());
    TRY(op_apply_v(M, mp, (const void *const *)p0.data(), isign, cb));
    TRY(op_apply_v(M, mmp, cv(mp), -isign, cb));
    for (int f = 0; f < nf; f++) {
      BlasArgs a = bargs();
      a.x = mp[f];
      TRY(run_blas_dev(g, B_NORM2, a, &ms->cf[f]));
    }
    HIPCHK(ms_mid_launch(ms, g_stream));
    for (int f = 0; f < nf; f++)
      HIPCHK(axpy_norm2_dev_launch(g->d, &ms->b, mmp[f], r[f], g_partials, &ms->cf[f], g_stream));
    HIPCHK(ms_post_launch(ms, g_stream));
    return QPHIX_OK;
  };
  HIPCHK(ms_pre_launch(ms, g_stream));  // the first iteration's a, as_s
  TRY(device_iterations(s->ms_mode, s->max_iters, iteration, g_ms_view));
  TRY(sweep());  // the pending x_s updates (a no-op if the last iteration applied them)
  HIPCHK(ms_post_launch(ms, g_stream));
  HIPCHK(hipMemcpyAsync(&h, ms, sizeof(h), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  *n_iters = h.iters;
  for (int q = 0; q < n_shift; q++)
    if (h.conv[q]) rsd_sq_final[q] = h.rsd_sq_final[q];
  if (site_flops) *site_flops = h.site_flops;
  if (mv_apps) *mv_apps = 2ul + 2ul * (unsigned long)h.iters;
  *k_out = h.iters + 1;
  return h.converged ? QPHIX_OK : -1;
}

extern "C" int qphix_solver_set_mode(qphix_solver *s, int mode) {
  CHKPTR(s);
  if (mode < 0 || mode > 2) return fail(QPHIX_EDOMAIN, "solver mode %d (0, 1 or 2)", mode);
  if (s->kind == S_MINVCG) {
    s->ms_mode = mode;
  } else if (s->kind == S_BICGSTAB || s->kind == S_CG) {
    s->dev_mode = mode;
  } else {
    return fail(QPHIX_EINVAL, "device-resident control exists for CG, BiCGStab and multishift CG solvers");
  }
  return QPHIX_OK;
}

extern "C" int qphix_minvcg_set_mode(qphix_solver *s, int mode) {
  CHKPTR(s);
  if (s->kind != S_MINVCG) return fail(QPHIX_EINVAL, "not a multishift solver");
  if (mode < 0 || mode > 2) return fail(QPHIX_EDOMAIN, "multishift mode %d (0, 1 or 2)", mode);
  s->ms_mode = mode;
  return QPHIX_OK;
}

// minvcg.h:77-387; site_flops as minvcg.h:139-307 counts them (47 per norm,
// 95 per axpy + norm, 24 per shift for the x_s start, then per iteration 190
// plus 72 + 48 per shift not yet converged; not scaled by the flavours)
extern "C" int qphix_minvcg_solve(qphix_solver *s, void *const x[], const void *const rhs[], int n_shift,
                                  const double *shifts, const double *rsd_target, int isign, int verbose, int cb,
                                  int *n_iters, double *rsd_sq_final, unsigned long *site_flops,
                                  unsigned long *mv_apps) {
  ENSURE_INIT();
  CHKPTR(s); CHKPTR(x); CHKPTR(rhs); CHKPTR(shifts); CHKPTR(rsd_target); CHKPTR(n_iters); CHKPTR(rsd_sq_final);
  if (s->kind != S_MINVCG) return fail(QPHIX_EINVAL, "not a multishift solver");
  if (n_shift < 1 || n_shift > s->max_shifts)
    return fail(QPHIX_EDOMAIN, "n_shift = %d exceeds MaxShifts = %d", n_shift, s->max_shifts);
  if (n_shift > kMaxShifts) return fail(QPHIX_EDOMAIN, "n_shift = %d exceeds the library limit %d", n_shift, kMaxShifts);
  qphix_op *M = s->M;
  const qphix_geom *g = M->g;
  const int nf = M->nflav;
  SV r, p0, mp, mmp;
  std::vector<SV> p(n_shift);
  for (void *w : s->work) qphix_free(w);
  s->work.clear();
  TRY(alloc_set(g, nf, r, s));
  TRY(alloc_set(g, nf, p0, s));
  TRY(alloc_set(g, nf, mp, s));
  TRY(alloc_set(g, nf, mmp, s));
  for (int q = 0; q < n_shift; q++) TRY(alloc_set(g, nf, p[q], s));
  std::vector<double> zeta(n_shift), zeta_prev(n_shift), beta(n_shift), rsd_sq(n_shift);
  std::vector<char> conv(n_shift, 0);
  unsigned long mv = 0, fl = 0;
  double cp;
  TRY(mf_norm2(g, nf, rhs, cp));
  fl += 47;
  for (int q = 0; q < n_shift; q++) rsd_sq[q] = rsd_target[q] * rsd_target[q] * cp;
  for (int f = 0; f < nf; f++) {
    TRY(qphix_copy_spinor(g, r[f], rhs[f]));
    TRY(qphix_copy_spinor(g, p0[f], rhs[f]));
    for (int q = 0; q < n_shift; q++) TRY(qphix_copy_spinor(g, p[q][f], rhs[f]));
  }
  TRY(op_apply_v(M, mp, cv(p0), isign, cb));
  TRY(op_apply_v(M, mmp, cv(mp), -isign, cb));
  mv += 2;
  double d;
  TRY(mf_norm2(g, nf, cv(mp), d));
  double b = -cp / d, c = 0;
  for (int f = 0; f < nf; f++) {
    double t;
    TRY(qphix_axpy_norm2(g, b, mmp[f], r[f], &t));
    c += t;
  }
  fl += 47 + 95 + 24ul * (unsigned long)n_shift;
  for (int q = 0; q < n_shift; q++) {
    zeta_prev[q] = 1;
    zeta[q] = 1.0 / (1.0 - shifts[q] * b);
    beta[q] = b / (1.0 - shifts[q] * b);
    for (int f = 0; f < nf; f++) TRY(qphix_axy(g, -beta[q], rhs[f], x[q * nf + f]));
  }
  if (s->ms_mode > 0) {
    int kd = 0;
    const int rc = minvcg_device_loop(s, x, n_shift, shifts, isign, cb, r, p0, mp, mmp, p, zeta, zeta_prev, beta,
                                      rsd_sq, b, c, cp, fl, rsd_sq_final, n_iters, site_flops, mv_apps, &kd);
    for (void *w : s->work) qphix_free(w);
    s->work.clear();
    if (rc < 0) return fail(QPHIX_ENOCONV, "multishift CG did not converge in %d iterations", kd);
    return rc;
  }
  // one fused sweep per iteration updates p0 and every shift's p_s, and applies
  // the x_s update left pending by the previous iteration (minvcg.h:298-363)
  auto fused = [&](int do_p0, double a, unsigned mask_p, unsigned mask_x, const double *as) -> int {
    for (int f = 0; f < nf; f++) {
      MShiftArgs m;
      memset(&m, 0, sizeof(m));
      m.r = r[f];
      m.p0 = p0[f];
      m.a = a;
      m.do_p0 = do_p0;
      m.nshift = n_shift;
      m.mask_p = mask_p;
      m.mask_x = mask_x;
      for (int q = 0; q < n_shift; q++) {
        m.p[q] = p[q][f];
        m.x[q] = x[q * nf + f];
        m.zeta[q] = zeta[q];
        m.as[q] = as ? as[q] : 0.0;
        m.beta[q] = beta[q];
      }
      HIPCHK(mshift_launch(g->d, m, g_stream));
    }
    return QPHIX_OK;
  };
  bool convP = false;
  int k;
  unsigned pending_x = 0;
  std::vector<double> as(n_shift, 0.0);
  *n_iters = 0;
  for (k = 1; k < s->max_iters && !convP; ++k) {
    const double a = c / cp;
    unsigned mask_p = 0;
    for (int q = 0; q < n_shift; q++)
      if (!conv[q]) {
        as[q] = a * zeta[q] * beta[q] / (zeta_prev[q] * b);
        mask_p |= 1u << q;
        fl += 72 + 48;
      }
    fl += 190;
    TRY(fused(1, a, mask_p, pending_x, as.data()));
    cp = c;
    const double bp = b;
    TRY(op_apply_v(M, mp, cv(p0), isign, cb));
    TRY(op_apply_v(M, mmp, cv(mp), -isign, cb));
    mv += 2;
    TRY(mf_norm2(g, nf, cv(mp), d));
    b = -cp / d;
    c = 0;
    for (int f = 0; f < nf; f++) {
      double t;
      TRY(qphix_axpy_norm2(g, b, mmp[f], r[f], &t));
      c += t;
    }
    convP = true;
    pending_x = 0;
    for (int q = 0; q < n_shift; q++) {
      if (!conv[q]) {
        const double z0 = zeta[q], z1 = zeta_prev[q];
        zeta_prev[q] = z0;
        zeta[q] = z0 * z1 * bp;
        zeta[q] /= (b * a * (z1 - z0) + z1 * bp * (1.0 - shifts[q] * b));
        beta[q] = b * zeta[q] / z0;
        pending_x |= 1u << q;
        const double css = c * zeta[q] * zeta[q];
        if (verbose) printf("MINVCG: iter=%d shift=%d css=%16.8e target=%16.8e\n", k, q, css, rsd_sq[q]);
        conv[q] = css < rsd_sq[q];
        if (conv[q]) rsd_sq_final[q] = css;
      }
      convP &= (conv[q] != 0);
    }
    *n_iters = k;
  }
  if (pending_x) TRY(fused(0, 0.0, 0u, pending_x, nullptr));
  if (site_flops) *site_flops = fl;
  if (mv_apps) *mv_apps = mv;
  for (void *w : s->work) qphix_free(w);
  s->work.clear();
  return convP ? QPHIX_OK : fail(QPHIX_ENOCONV, "multishift CG did not converge in %d iterations", k);
}

// ---------------------------------------------------------------------------
// batched solves: many single-flavour systems (possibly on different lattices)
// advanced together, one launch per step for all of them, with device-resident
// FP64 scalars and (optionally) the iteration captured once in a HIP graph
// ---------------------------------------------------------------------------
struct qphix_batch {
  std::vector<qphix_op *> ops;
  int n = 0, prec = 0, c12 = 0, kind = 0;
  std::vector<void *> tmp, mp, mmp, p, r;  // per-system work spinors
  // mixed-precision Richardson work spinors, kept between calls while the
  // outer spinor sizes stay the same: r, dx, tmp, tmp2 (outer), ri, dxi (inner)
  std::vector<size_t> rich_bytes;
  std::vector<void *> rich[6];
  std::vector<int> off;                    // chunk / work-group offsets (256 sites)
  int *d_off = nullptr, *d_active = nullptr;
  // hop segments: one per system, or (tile > 1) one per group of `tile`
  // consecutive systems that share one operator, applied as a register tile
  int tile = 1, nhseg = 0;
  std::vector<int> hoff;
  int *d_hoff = nullptr;
  HopSeg *d_hop = nullptr;                 // 4 stages x n
  SegVec *d_seg = nullptr;                 // 8 step kinds x n
  CgState *d_state = nullptr;
  double *d_alpha = nullptr, *d_beta = nullptr, *d_d = nullptr, *d_cp = nullptr, *d_rsd = nullptr;
  double *d_partials = nullptr;
};

static int batch_free_dev(qphix_batch *b) {
  void *ptrs[] = {b->d_off, b->d_active, b->d_hop, b->d_seg, b->d_state, b->d_alpha, b->d_beta,
                  b->d_d, b->d_cp, b->d_rsd, b->d_partials, b->d_hoff};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  return QPHIX_OK;
}

extern "C" int qphix_batch_destroy(qphix_batch *b) {
  if (!b) return QPHIX_OK;
  if (g_init) (void)hipStreamSynchronize(g_stream);
  std::vector<void *> *sets[5] = {&b->tmp, &b->mp, &b->mmp, &b->p, &b->r};
  for (auto *v : sets)
    for (void *q : *v) qphix_free(q);
  for (auto &v : b->rich)
    for (void *q : v) qphix_free(q);
  batch_free_dev(b);
  delete b;
  return QPHIX_OK;
}

extern "C" int qphix_batch_create(qphix_op *const ops[], int nsys, qphix_batch **out) {
  ENSURE_INIT();
  CHKPTR(ops); CHKPTR(out);
  if (nsys < 1) return fail(QPHIX_EDOMAIN, "batch of %d systems", nsys);
  qphix_batch *b = new qphix_batch();
  b->n = nsys;
  for (int i = 0; i < nsys; i++) {
    if (!ops[i]) { delete b; return fail(QPHIX_EINVAL, "null operator %d", i); }
    if (ops[i]->nflav != 1) { delete b; return fail(QPHIX_EDOMAIN, "batched solves take single-flavour operators"); }
    b->ops.push_back(ops[i]);
  }
  const GeomDesc &g0 = ops[0]->g->d;
  b->prec = g0.prec;
  b->c12 = g0.c12;
  b->kind = ops[0]->p.kind;
  for (int i = 1; i < nsys; i++) {
    const GeomDesc &gi = ops[i]->g->d;
    if (gi.prec != b->prec || gi.c12 != b->c12 || ops[i]->p.kind != b->kind) {
      delete b;
      return fail(QPHIX_EDOMAIN, "batched operators must share precision, compression and kind");
    }
  }
  b->off.assign(nsys + 1, 0);
  for (int i = 0; i < nsys; i++) b->off[i + 1] = b->off[i] + (int)((ops[i]->g->d.vh + 255) / 256);
  // Systems that come in groups of kBatchTile sharing one operator (several
  // right-hand sides per lattice) can run their hops as a register tile, which
  // reads links and clover once per group. That pays where the batched hops
  // are bandwidth-bound, i.e. when the batch holds enough sites; a batch of a
  // few small lattices is latency-bound, and one source per work-item (twice
  // the work-items) is faster. Default: tile when the batch has at least 2^18
  // sites per parity in all; QPHIX_BATCH_TILE=0 / 1 forces it off / on (where
  // the grouping allows).
  const char *te = getenv("QPHIX_BATCH_TILE");
  long sites = 0;
  for (int i = 0; i < nsys; i++) sites += ops[i]->g->d.vh;
  bool want = sites >= (1L << 18);
  if (te && te[0] == '0') want = false;
  if (te && te[0] == '1') want = true;
  b->tile = want && nsys % kBatchTile == 0 ? kBatchTile : 1;
  for (int i = 0; i < nsys && b->tile > 1; i++)
    if (ops[i] != ops[i - i % kBatchTile]) b->tile = 1;
  b->nhseg = nsys / b->tile;
  b->hoff.assign(b->nhseg + 1, 0);
  for (int h = 0; h < b->nhseg; h++)
    b->hoff[h + 1] = b->hoff[h] + (int)((ops[h * b->tile]->g->d.vh + 255) / 256);
  std::vector<void *> *sets[5] = {&b->tmp, &b->mp, &b->mmp, &b->p, &b->r};
  for (auto *v : sets)
    for (int i = 0; i < nsys; i++) {
      void *q = nullptr;
      int rc = qphix_alloc_spinor(ops[i]->g, &q);
      if (rc) { qphix_batch_destroy(b); return rc; }
      v->push_back(q);
    }
  const size_t nn = nsys;
  bool ok = hipMalloc(&b->d_off, sizeof(int) * (nn + 1)) == hipSuccess &&
            hipMalloc(&b->d_active, sizeof(int) * nn) == hipSuccess &&
            hipMalloc(&b->d_hop, sizeof(HopSeg) * 4 * nn) == hipSuccess &&
            hipMalloc(&b->d_seg, sizeof(SegVec) * 8 * nn) == hipSuccess &&
            hipMalloc(&b->d_state, sizeof(CgState) * nn) == hipSuccess &&
            hipMalloc(&b->d_alpha, sizeof(double) * nn) == hipSuccess &&
            hipMalloc(&b->d_beta, sizeof(double) * nn) == hipSuccess &&
            hipMalloc(&b->d_d, sizeof(double) * nn) == hipSuccess &&
            hipMalloc(&b->d_cp, sizeof(double) * nn) == hipSuccess &&
            hipMalloc(&b->d_rsd, sizeof(double) * nn) == hipSuccess &&
            hipMalloc(&b->d_partials, sizeof(double) * (size_t)b->off[nsys]) == hipSuccess &&
            hipMalloc(&b->d_hoff, sizeof(int) * (size_t)(b->nhseg + 1)) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    qphix_batch_destroy(b);
    return fail(QPHIX_ENOMEM, "batch device allocation failed");
  }
  if (hipMemcpy(b->d_off, b->off.data(), sizeof(int) * (nn + 1), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_hoff, b->hoff.data(), sizeof(int) * (size_t)(b->nhseg + 1), hipMemcpyHostToDevice) != hipSuccess) {
    qphix_batch_destroy(b);
    return fail(QPHIX_EHIP, "batch offsets upload failed");
  }
  *out = b;
  return QPHIX_OK;
}

// the two hop launches of op (isign, target t): in -> tmp (other cb) -> out
static int op_stages(qphix_op *op, int isign, int t, const void *in, void *tmp, void *out, int kinds[2],
                     HopArgs args[2]) {
  const int o = 1 - t;
  const qphix_op_params &p = op->p;
  qphix_dslash *D = &op->d;
  switch (p.kind) {
    case QPHIX_OP_WILSON: {
      const double al = 4.0 + p.mass;
      kinds[0] = KD_WILSON;
      args[0] = make_hop_args(D, o, p.u[o], nullptr, 1.0, 0, 0, 0, 0);
      kinds[1] = KA_WILSON;
      args[1] = make_hop_args(D, t, p.u[t], nullptr, -0.25 / al, al, 0, 0, 0);
    } break;
    case QPHIX_OP_CLOVER:
      kinds[0] = KD_CLOVER;
      args[0] = make_hop_args(D, o, p.u[o], p.invclov[o], 1.0, 0, 0, 0, 0);
      kinds[1] = KA_CLOVER;
      args[1] = make_hop_args(D, t, p.u[t], p.clov[t], -0.25, 0, 0, 0, D->p.prec_mass_rho);
      break;
    case QPHIX_OP_TMWILSON: {
      const double al = 4.0 + p.mass, mu = p.twisted_mass;
      kinds[0] = KD_TM;
      args[0] = make_hop_args(D, o, p.u[o], nullptr, 1.0, 0, mu / al, al / (mu * mu + al * al), 0);
      kinds[1] = KA_TM;
      args[1] = make_hop_args(D, t, p.u[t], nullptr, -0.25, al, mu / al, 0, 0);
    } break;
    case QPHIX_OP_TMCLOVER: {
      const int clidx = isign == 1 ? 0 : 1;
      kinds[0] = KD_FULL;
      args[0] = make_hop_args(D, o, p.u[o], p.fullclov[o][clidx], 1.0, 0, 0, 0, 0);
      kinds[1] = KA_CLOVER;
      args[1] = make_hop_args(D, t, p.u[t], p.clov[t], -0.25, 0, 0, 0, D->p.prec_mass_rho);
    } break;
    default: return fail(QPHIX_EDOMAIN, "batched solves support Wilson, clover and twisted-mass kinds");
  }
  if ((kinds[0] == KD_CLOVER || kinds[0] == KD_FULL) && !args[0].clov) return fail(QPHIX_EINVAL, "clover field missing");
  if (kinds[1] == KA_CLOVER && !args[1].clov) return fail(QPHIX_EINVAL, "clover field missing");
  args[0].psi = in;
  args[0].out = tmp;
  args[1].psi = tmp;
  args[1].chi = in;
  args[1].out = out;
  return QPHIX_OK;
}

static int batch_grid(const qphix_batch *b) {
  const int total = b->off[b->n];
  return total < 2048 ? total : 2048;
}

static hipError_t batch_hop(const qphix_batch *b, int stage, int kind, int isign) {
  const HopSeg *segs = b->d_hop + (size_t)stage * b->n;
  if (b->tile > 1) {
    const int total = b->hoff[b->nhseg], grid = total < 2048 ? total : 2048;
    switch (b->prec) {
      case P64: return launch_hop_batch_tile_prec<P64>(kind, b->c12 != 0, isign, segs, b->d_hoff, b->nhseg, b->d_active, grid, g_stream);
      case P32: return launch_hop_batch_tile_prec<P32>(kind, b->c12 != 0, isign, segs, b->d_hoff, b->nhseg, b->d_active, grid, g_stream);
      case P16: return launch_hop_batch_tile_prec<P16>(kind, b->c12 != 0, isign, segs, b->d_hoff, b->nhseg, b->d_active, grid, g_stream);
      default: return launch_hop_batch_tile_prec<PB16>(kind, b->c12 != 0, isign, segs, b->d_hoff, b->nhseg, b->d_active, grid, g_stream);
    }
  }
  switch (b->prec) {
    case P64: return launch_hop_batch_prec<P64>(kind, b->c12 != 0, isign, segs, b->d_off, b->n, b->d_active, batch_grid(b), g_stream);
    case P32: return launch_hop_batch_prec<P32>(kind, b->c12 != 0, isign, segs, b->d_off, b->n, b->d_active, batch_grid(b), g_stream);
    case P16: return launch_hop_batch_prec<P16>(kind, b->c12 != 0, isign, segs, b->d_off, b->n, b->d_active, batch_grid(b), g_stream);
    default: return launch_hop_batch_prec<PB16>(kind, b->c12 != 0, isign, segs, b->d_off, b->n, b->d_active, batch_grid(b), g_stream);
  }
}
static hipError_t batch_blas(const qphix_batch *b, int step, int op, const double *scal, double *out, bool gate) {
  return batch_blas_launch(b->prec, op, b->d_seg + (size_t)step * b->n, b->d_off, b->n, gate ? b->d_active : nullptr,
                           scal, b->d_partials, out, batch_grid(b), g_stream);
}

// one CG iteration on M^+M for every active system (no host interaction)
static hipError_t batch_cg_iteration(const qphix_batch *b, const int kinds[4], int isign) {
  hipError_t e;
  if ((e = batch_hop(b, 0, kinds[0], isign)) != hipSuccess) return e;
  if ((e = batch_hop(b, 1, kinds[1], isign)) != hipSuccess) return e;
  if ((e = batch_hop(b, 2, kinds[2], -isign)) != hipSuccess) return e;
  if ((e = batch_hop(b, 3, kinds[3], -isign)) != hipSuccess) return e;
  if ((e = batch_blas(b, 0, BB_NORM2, nullptr, b->d_d, true)) != hipSuccess) return e;         // d = |mp|^2
  if ((e = batch_cg_alpha(b->d_state, b->d_d, b->d_alpha, b->d_active, b->n, g_stream)) != hipSuccess) return e;
  if ((e = batch_blas(b, 1, BB_RMAMMP, b->d_alpha, b->d_cp, true)) != hipSuccess) return e;   // r -= a mmp, x += a p
  if ((e = batch_cg_update(b->d_state, b->d_cp, b->d_beta, b->d_active, b->n, g_stream)) != hipSuccess) return e;
  return batch_blas(b, 2, BB_AYPX, b->d_beta, nullptr, true);                                   // p = r + b p
}

extern "C" int qphix_batch_cg(qphix_batch *b, void *const x[], const void *const rhs[], const double rsd_target[],
                              int max_iters, int isign, int t, int graph_steps, qphix_solve_stats st[]) {
  ENSURE_INIT();
  CHKPTR(b); CHKPTR(x); CHKPTR(rhs); CHKPTR(rsd_target); CHKPTR(st);
  TRY(check_cb(t));
  if (isign != 1 && isign != -1) return fail(QPHIX_EDOMAIN, "isign must be +1 or -1");
  const int n = b->n;
  // segment tables for this solve's vectors
  std::vector<HopSeg> hop(4 * (size_t)n);
  std::vector<SegVec> seg(8 * (size_t)n);
  int kinds[4];
  for (int i = 0; i < n; i++) {
    if (!x[i] || !rhs[i]) return fail(QPHIX_EINVAL, "null vector for system %d", i);
    const GeomDesc &g = b->ops[i]->g->d;
    HopArgs a[2];
    int k2[2];
    // hop segment of system i: its own (tile 1), or slot j of its group's
    // register tile (the group's first system carries the operator)
    const int h = i / b->tile, j = i % b->tile;
    auto put = [&](int stage, const HopArgs &ai) {
      HopSeg &sg = hop[(size_t)stage * n + (b->tile > 1 ? h : i)];
      if (j == 0) sg = HopSeg{ai, i, b->tile};
      if (b->tile > 1) {
        sg.a.nrhs = b->tile;
        sg.a.psis[j] = ai.psi;
        sg.a.outs[j] = ai.out;
        sg.a.chis[j] = ai.chi;
      }
    };
    TRY(op_stages(b->ops[i], isign, t, b->p[i], b->tmp[i], b->mp[i], k2, a));   // mp = M p
    put(0, a[0]);
    put(1, a[1]);
    kinds[0] = k2[0];
    kinds[1] = k2[1];
    TRY(op_stages(b->ops[i], -isign, t, b->mp[i], b->tmp[i], b->mmp[i], k2, a));  // mmp = M^+ mp
    put(2, a[0]);
    put(3, a[1]);
    kinds[2] = k2[0];
    kinds[3] = k2[1];
    auto sv = [&](const void *xx, const void *yy, void *o0, void *o1) { return SegVec{xx, yy, o0, o1, g.vh, g.vs, i}; };
    seg[0 * n + i] = sv(b->mp[i], nullptr, nullptr, nullptr);        // |mp|^2
    seg[1 * n + i] = sv(b->mmp[i], b->p[i], b->r[i], x[i]);          // r -= a mmp, x += a p
    seg[2 * n + i] = sv(b->r[i], nullptr, b->p[i], nullptr);         // p = r + b p
    seg[3 * n + i] = sv(rhs[i], nullptr, nullptr, nullptr);          // |rhs|^2
    seg[4 * n + i] = sv(rhs[i], b->mmp[i], b->r[i], nullptr);        // r = rhs - mmp
    seg[5 * n + i] = sv(b->r[i], nullptr, b->p[i], nullptr);         // p = r
    seg[6 * n + i] = sv(x[i], nullptr, nullptr, nullptr);            // |x|^2
  }
  HIPCHK(hipMemcpyAsync(b->d_hop, hop.data(), sizeof(HopSeg) * hop.size(), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));  // the host tables are local
  HIPCHK(hipMemcpyAsync(b->d_seg, seg.data(), sizeof(SegVec) * seg.size(), hipMemcpyHostToDevice, g_stream));
  // set-up: r = rhs - M^+M x (the operator only where x != 0), p = r, targets
  std::vector<double> rn(n), xn(n), rsd(n);
  HIPCHK(batch_blas(b, 3, BB_NORM2, nullptr, b->d_d, false));
  HIPCHK(hipMemcpyAsync(rn.data(), b->d_d, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(batch_blas(b, 6, BB_NORM2, nullptr, b->d_d, false));
  HIPCHK(hipMemcpyAsync(xn.data(), b->d_d, sizeof(double) * n, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  std::vector<int> act(n);
  for (int i = 0; i < n; i++) {
    rsd[i] = rsd_target[i] * rsd_target[i] * rn[i];
    act[i] = xn[i] != 0.0;
  }
  HIPCHK(hipMemcpyAsync(b->d_active, act.data(), sizeof(int) * n, hipMemcpyHostToDevice, g_stream));
  // hop stages 0-3 read p; start from p = x for the systems with a guess
  std::vector<SegVec> cpx(n);
  for (int i = 0; i < n; i++) cpx[i] = seg[6 * n + i], cpx[i].o0 = b->p[i];
  HIPCHK(hipMemcpyAsync(b->d_seg + 7 * (size_t)n, cpx.data(), sizeof(SegVec) * n, hipMemcpyHostToDevice, g_stream));
  HIPCHK(batch_blas(b, 7, BB_COPY, nullptr, nullptr, true));
  for (int s = 0; s < 4; s++) HIPCHK(batch_hop(b, s, kinds[s], s < 2 ? isign : -isign));
  // systems without a guess: mmp = 0 so that r = rhs
  std::vector<SegVec> zero(n);
  for (int i = 0; i < n; i++) zero[i] = SegVec{nullptr, nullptr, b->mmp[i], nullptr, seg[i].vh, seg[i].vs, i};
  for (int i = 0; i < n; i++) act[i] = !act[i];
  HIPCHK(hipMemcpyAsync(b->d_seg + 7 * (size_t)n, zero.data(), sizeof(SegVec) * n, hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(b->d_active, act.data(), sizeof(int) * n, hipMemcpyHostToDevice, g_stream));
  HIPCHK(batch_blas(b, 7, BB_ZERO, nullptr, nullptr, true));
  HIPCHK(batch_blas(b, 4, BB_XMYNORM2, nullptr, b->d_cp, false));
  HIPCHK(batch_blas(b, 5, BB_COPY, nullptr, nullptr, false));
  HIPCHK(hipMemcpyAsync(b->d_rsd, rsd.data(), sizeof(double) * n, hipMemcpyHostToDevice, g_stream));
  HIPCHK(batch_cg_init(b->d_state, b->d_cp, b->d_rsd, b->d_active, max_iters, n, g_stream));
  const int nhops = 4;
  (void)nhops;
  // iterations
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (graph_steps > 0) {
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = batch_cg_iteration(b, kinds, isign);
    hipError_t e2 = hipStreamEndCapture(g_stream, &graph);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      return fail(QPHIX_EHIP, "graph capture of the CG iteration failed: %s",
                  hipGetErrorString(e != hipSuccess ? e : e2));
    }
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(QPHIX_EHIP, "graph instantiation failed");
    }
  }
  const int check = graph_steps > 0 ? graph_steps : 1;
  int done_iters = 0;
  int rc = QPHIX_OK;
  while (done_iters <= max_iters + 1) {
    for (int k = 0; k < check; k++) {
      hipError_t e = exec ? hipGraphLaunch(exec, g_stream) : batch_cg_iteration(b, kinds, isign);
      if (e != hipSuccess) {
        rc = fail(QPHIX_EHIP, "batched CG iteration: %s", hipGetErrorString(e));
        break;
      }
    }
    if (rc) break;
    done_iters += check;
    if (hipMemcpyAsync(act.data(), b->d_active, sizeof(int) * n, hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
        hipStreamSynchronize(g_stream) != hipSuccess) {
      rc = fail(QPHIX_EHIP, "reading convergence flags failed");
      break;
    }
    bool any = false;
    for (int i = 0; i < n; i++) any |= act[i] != 0;
    if (!any) break;
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc) return rc;
  std::vector<CgState> hs(n);
  HIPCHK(hipMemcpy(hs.data(), b->d_state, sizeof(CgState) * n, hipMemcpyDeviceToHost));
  for (int i = 0; i < n; i++) {
    memset(&st[i], 0, sizeof(st[i]));
    st[i].n_iters = hs[i].iters;
    st[i].rsd_sq_final = hs[i].cp;
    // the counters of invcg.h:180-416 for a solve that ran hs.loops iterations:
    // the initial M^+M x (counted also when x = 0 skipped it) and 2 per
    // iteration; norm2 + x - y norm, then norm2 + r/x update per iteration and
    // aypx on every iteration that did not converge
    const unsigned long lp = (unsigned long)hs[i].loops;
    const unsigned long conv_in_loop = hs[i].converged && lp > 0 ? 1 : 0;
    st[i].mv_apps = 2 + 2 * lp;
    st[i].site_flops = 12 * (4 + 6) + 12 * (4 + 12) * lp + 12 * 4 * (lp - conv_in_loop);
    st[i].status = hs[i].converged ? QPHIX_OK : QPHIX_ENOCONV;
  }
  return QPHIX_OK;
}

// FP64 defect correction around a batched inner CG on M^+M
// (inv_richardson_multiprec.h:42-239 per system with MdagM = true)
extern "C" int qphix_batch_richardson(qphix_batch *inner, qphix_op *const outer[], void *const x[],
                                      const void *const rhs[], double rsd_target, double delta, int max_outer,
                                      int max_inner, int isign, int t, int graph_steps, qphix_solve_stats st[]) {
  ENSURE_INIT();
  CHKPTR(inner); CHKPTR(outer); CHKPTR(x); CHKPTR(rhs); CHKPTR(st);
  const int n = inner->n;
  for (int i = 0; i < n; i++) {
    CHKPTR(outer[i]);
    TRY(same_lattice(outer[i]->g, inner->ops[i]->g));
  }
  // work spinors: allocated on the first call (or when an outer spinor size
  // changed) and kept in the batch object until qphix_batch_destroy
  bool reuse = (int)inner->rich_bytes.size() == n;
  for (int i = 0; reuse && i < n; i++) reuse = inner->rich_bytes[i] == outer[i]->g->bytes[0];
  if (!reuse) {
    for (auto &v : inner->rich) {
      for (void *q : v) qphix_free(q);
      v.assign(n, nullptr);
    }
    inner->rich_bytes.assign(n, 0);
    for (int i = 0; i < n; i++) {
      for (int k = 0; k < 6; k++)
        TRY(qphix_alloc_spinor(k < 4 ? outer[i]->g : inner->ops[i]->g, &inner->rich[k][i]));
      inner->rich_bytes[i] = outer[i]->g->bytes[0];
    }
  }
  const std::vector<void *> &r = inner->rich[0], &dx = inner->rich[1], &tmp = inner->rich[2],
                            &tmp2 = inner->rich[3], &ri = inner->rich[4], &dxi = inner->rich[5];
  std::vector<double> rhs_sq(n), rn2(n), rsd_t(n), tgt(n);
  std::vector<int> outer_it(n, 0);
  std::vector<unsigned long> mv(n, 0);
  // outer site flops as inv_richardson_multiprec.h:63,87,171,191 counts them
  std::vector<unsigned long> fl(n, (24 + 23) + (24 + 24 + 23));
  auto mdagm = [&](int i, const void *in) -> int {
    void *a[1] = {tmp2[i]}, *c[1] = {tmp[i]};
    const void *ii[1] = {in}, *mm[1] = {tmp2[i]};
    TRY(qphix_op_apply(outer[i], a, ii, isign, t));
    TRY(qphix_op_apply(outer[i], c, mm, -isign, t));
    mv[i] += 2;
    return QPHIX_OK;
  };
  for (int i = 0; i < n; i++) {
    TRY(qphix_norm2_spinor(outer[i]->g, rhs[i], &rhs_sq[i]));
    rsd_t[i] = rsd_target * rsd_target * rhs_sq[i];
    TRY(mdagm(i, x[i]));
    TRY(qphix_xmy_norm2(outer[i]->g, r[i], rhs[i], tmp[i], &rn2[i]));
  }
  for (int it = 1; it <= max_outer; it++) {
    bool any = false;
    for (int i = 0; i < n; i++) {
      const bool a = rn2[i] >= rsd_t[i];
      any |= a;
      const double rn = sqrt(rn2[i]);
      if (a) {
        const double dfin = sqrt(rsd_t[i] / rn2[i]);
        tgt[i] = dfin > delta ? dfin : delta;
        TRY(qphix_convert_spinor(inner->ops[i]->g, ri[i], 1.0 / rn, outer[i]->g, r[i]));
        outer_it[i] = it;
      } else {
        tgt[i] = 1e150;
        TRY(qphix_zero_spinor(inner->ops[i]->g, ri[i]));
      }
      TRY(qphix_zero_spinor(inner->ops[i]->g, dxi[i]));
    }
    if (!any) break;
    std::vector<qphix_solve_stats> ist(n);
    TRY(qphix_batch_cg(inner, dxi.data(), (const void *const *)ri.data(), tgt.data(), max_inner, isign, t, graph_steps,
                       ist.data()));
    for (int i = 0; i < n; i++) {
      if (rn2[i] < rsd_t[i]) continue;
      mv[i] += ist[i].mv_apps;
      fl[i] += ist[i].site_flops + 8 * 12;
      TRY(qphix_convert_spinor(outer[i]->g, dx[i], sqrt(rn2[i]), inner->ops[i]->g, dxi[i]));
      TRY(mdagm(i, dx[i]));
      TRY(qphix_richardson_rxupdate_norm_r(outer[i]->g, x[i], r[i], dx[i], tmp[i], &rn2[i]));
    }
  }
  for (int i = 0; i < n; i++) {
    TRY(mdagm(i, x[i]));
    TRY(qphix_xmy_norm2(outer[i]->g, r[i], rhs[i], tmp[i], &rn2[i]));
    memset(&st[i], 0, sizeof(st[i]));
    st[i].n_iters = outer_it[i];
    st[i].rsd_sq_final = rn2[i] / rhs_sq[i];
    st[i].mv_apps = mv[i];
    st[i].site_flops = fl[i] + (24 + 24 + 23);
    st[i].status = rn2[i] < rsd_t[i] ? QPHIX_OK : QPHIX_ENOCONV;
  }
  return QPHIX_OK;
}
