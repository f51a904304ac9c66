// Batched (segmented) BLAS and device-resident CG scalars for solving many
// systems at once, possibly on lattices of different sizes.
//
// A batch step walks a flat space of 256-site chunks; segment g (one system's
// vectors) owns chunks [off[g], off[g+1]). A fixed grid strides over the chunks
// (persistent form), so every launch has the same shape whatever the batch and
// the whole CG iteration can be captured once in a HIP graph. All scalars
// (alpha, beta, residual norms, convergence flags) live in device memory and
// are produced by one-work-group update kernels in FP64; nothing returns to the
// host inside an iteration. Reductions are per chunk, summed per segment in a
// fixed order, so results do not depend on the grid.
#include "batch.hpp"
#include "hop_impl.hpp"

namespace qphix_amd {

constexpr int kChunk = 256;

template <int NR> __device__ inline void chunk_reduce(const double *red, double *out) {
  __shared__ double lds[kChunk / 64][NR];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double v = red[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) lds[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NR) {
    double v = 0;
    for (int w = 0; w < kChunk / 64; w++) v += lds[w][threadIdx.x];
    out[threadIdx.x] = v;
  }
  __syncthreads();
}

template <int P, int OP>
__global__ void __launch_bounds__(kChunk) k_batch_blas(const SegVec *segs, const int *off, int nseg, const int *active,
                                                        const double *scal, double *partials) {
  typedef typename PT<P>::calc T;
  constexpr int N = kSpinorReals;
  constexpr int NR = (OP == BB_NORM2 || OP == BB_RMAMMP || OP == BB_XMYNORM2) ? 1 : 0;
  const int total = off[nseg];
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    int g = 0;
    while (g + 1 < nseg && off[g + 1] <= b) g++;
    const SegVec &v = segs[g];
    if (active && !active[v.sys]) continue;
    const long s = (long)(b - off[g]) * kChunk + threadIdx.x;
    double red[1] = {0.0};
    if (s < v.vh) {
      T x[N], y[N];
      if constexpr (OP == BB_NORM2) {
        load_spinor<P, true>(v.x, v.vs, s, x);
#pragma unroll
        for (int k = 0; k < N; k++) red[0] += (double)x[k] * (double)x[k];
      } else if constexpr (OP == BB_COPY) {
        load_spinor<P, true>(v.x, v.vs, s, x);
        store_spinor<P, true>(v.o0, v.vs, s, x);
      } else if constexpr (OP == BB_ZERO) {
#pragma unroll
        for (int k = 0; k < N; k++) x[k] = 0;
        store_spinor<P, true>(v.o0, v.vs, s, x);
      } else if constexpr (OP == BB_XMYNORM2) {  // o0 = x - y
        load_spinor<P, true>(v.x, v.vs, s, x);
        load_spinor<P, true>(v.y, v.vs, s, y);
#pragma unroll
        for (int k = 0; k < N; k++) {
          x[k] = x[k] - y[k];
          red[0] += (double)x[k] * (double)x[k];
        }
        store_spinor<P, true>(v.o0, v.vs, s, x);
      } else if constexpr (OP == BB_RMAMMP) {  // r(o0) -= a mmp(x) ; x(o1) += a p(y)
        const T a = (T)scal[v.sys];
        load_spinor<P, true>(v.x, v.vs, s, x);
        load_spinor<P, true>(v.o0, v.vs, s, y);
#pragma unroll
        for (int k = 0; k < N; k++) {
          y[k] = y[k] - a * x[k];
          red[0] += (double)y[k] * (double)y[k];
        }
        store_spinor<P, true>(v.o0, v.vs, s, y);
        load_spinor<P, true>(v.y, v.vs, s, x);
        load_spinor<P, true>(v.o1, v.vs, s, y);
#pragma unroll
        for (int k = 0; k < N; k++) y[k] = y[k] + a * x[k];
        store_spinor<P, true>(v.o1, v.vs, s, y);
      } else if constexpr (OP == BB_AYPX) {  // p(o0) = r(x) + b p
        const T bb = (T)scal[v.sys];
        load_spinor<P, true>(v.x, v.vs, s, x);
        load_spinor<P, true>(v.o0, v.vs, s, y);
#pragma unroll
        for (int k = 0; k < N; k++) y[k] = x[k] + bb * y[k];
        store_spinor<P, true>(v.o0, v.vs, s, y);
      }
    }
    if constexpr (NR > 0) chunk_reduce<1>(red, partials + b);
  }
}

// out[sys] = sum of the segment's chunk partials (active segments only)
__global__ void __launch_bounds__(kChunk) k_batch_finalize(const SegVec *segs, const int *off, int nseg,
                                                           const int *active, const double *partials, double *out) {
  const int g = blockIdx.x;
  if (g >= nseg) return;
  const SegVec &v = segs[g];
  if (active && !active[v.sys]) return;
  double red[1] = {0.0};
  for (int b = off[g] + threadIdx.x; b < off[g + 1]; b += blockDim.x) red[0] += partials[b];
  double o[1];
  __shared__ double res[1];
  chunk_reduce<1>(red, res);
  if (threadIdx.x == 0) out[v.sys] = res[0];
  (void)o;
}

// CG scalars (invcg.h:281-417 per system): alpha = c / d
__global__ void k_cg_alpha(CgState *st, const double *d, double *alpha, const int *active, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !active[s]) return;
  st[s].d = d[s];
  st[s].a = st[s].c / st[s].d;
  alpha[s] = st[s].a;
}

// after r -= a mmp: convergence test, beta = cp / c, c = cp
__global__ void k_cg_update(CgState *st, const double *cp, double *beta, int *active, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || !active[s]) return;
  CgState &c = st[s];
  c.cp = cp[s];
  c.loops++;
  if (c.cp < c.rsd_sq) {
    c.converged = 1;
    active[s] = 0;
    return;
  }
  c.iters++;
  if (c.iters > c.max_iters) {
    active[s] = 0;
    return;
  }
  c.b = c.cp / c.c;
  c.c = c.cp;
  beta[s] = c.b;
}

// set-up: c = |r|^2 (already in cp), targets, flags
__global__ void k_cg_init(CgState *st, const double *cp, const double *rsd_sq, int *active, int max_iters, int n) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  CgState &c = st[s];
  c.cp = cp[s];
  c.c = cp[s];
  c.rsd_sq = rsd_sq[s];
  c.iters = 0;
  c.loops = 0;
  c.max_iters = max_iters;
  c.converged = c.cp <= c.rsd_sq ? 1 : 0;
  c.a = c.b = c.d = 0;
  active[s] = c.converged ? 0 : 1;
}

template <int P>
static hipError_t blas_prec(int op, const SegVec *segs, const int *off, int nseg, const int *active,
                            const double *scal, double *partials, int grid, hipStream_t st) {
  switch (op) {
#define BC(O) \
  case O: k_batch_blas<P, O><<<grid, kChunk, 0, st>>>(segs, off, nseg, active, scal, partials); break;
    BC(BB_NORM2) BC(BB_COPY) BC(BB_ZERO) BC(BB_XMYNORM2) BC(BB_RMAMMP) BC(BB_AYPX)
#undef BC
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t batch_blas_launch(int prec, int op, const SegVec *segs, const int *off, int nseg, const int *active,
                             const double *scal, double *partials, double *out, int grid, hipStream_t st) {
  hipError_t e;
  switch (prec) {
    case P64: e = blas_prec<P64>(op, segs, off, nseg, active, scal, partials, grid, st); break;
    case P32: e = blas_prec<P32>(op, segs, off, nseg, active, scal, partials, grid, st); break;
    case P16: e = blas_prec<P16>(op, segs, off, nseg, active, scal, partials, grid, st); break;
    default: e = blas_prec<PB16>(op, segs, off, nseg, active, scal, partials, grid, st); break;
  }
  if (e != hipSuccess) return e;
  if (out && (op == BB_NORM2 || op == BB_RMAMMP || op == BB_XMYNORM2)) {
    k_batch_finalize<<<nseg, kChunk, 0, st>>>(segs, off, nseg, active, partials, out);
    return hipGetLastError();
  }
  return hipSuccess;
}

hipError_t batch_cg_alpha(CgState *st_, const double *d, double *alpha, const int *active, int n, hipStream_t st) {
  k_cg_alpha<<<(n + 63) / 64, 64, 0, st>>>(st_, d, alpha, active, n);
  return hipGetLastError();
}
hipError_t batch_cg_update(CgState *st_, const double *cp, double *beta, int *active, int n, hipStream_t st) {
  k_cg_update<<<(n + 63) / 64, 64, 0, st>>>(st_, cp, beta, active, n);
  return hipGetLastError();
}
hipError_t batch_cg_init(CgState *st_, const double *cp, const double *rsd_sq, int *active, int max_iters, int n,
                         hipStream_t st) {
  k_cg_init<<<(n + 63) / 64, 64, 0, st>>>(st_, cp, rsd_sq, active, max_iters, n);
  return hipGetLastError();
}

}  // namespace qphix_amd
