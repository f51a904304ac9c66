// Streaming BLAS over checkerboard spinors with fused FP64 reductions
// (include/qphix/blas_new_c.h, real_functors.h, complex_functors.h).
//
// Every operation is one grid-stride sweep: each work-item loads its sites'
// operands, updates them in the geometry's arithmetic type (float for F32 and
// F16 storage, as ArithType<half> = float in the reference) and accumulates up
// to three reductions in double. Reductions are combined by a wavefront
// shuffle tree, then across the waves of the block through LDS, and the
// per-block partials are summed in a fixed order by a one-block finalize
// kernel, so results are bitwise reproducible for a given lattice.
#include "fields.hpp"
#include "hop_impl.hpp"

namespace qphix_amd {

constexpr int kBlasThreads = 256;

__host__ __device__ constexpr int nred_of(int op) {
  return (op == B_NORM2 || op == B_AXPYNORM2 || op == B_XMYNORM2 || op == B_XMY2NORM2 || op == B_RMAMMP ||
          op == B_RICH || op == B_RXUPD)
             ? 1
             : (op == B_CDOT) ? 2 : (op == B_CDOTNORM || op == B_RXUPD_CDOT) ? 3 : 0;
}
int blas_nred(int op) { return nred_of(op); }

template <typename T> __device__ inline double sq(const T *v) {
  double s = 0;
#pragma unroll
  for (int k = 0; k < kSpinorReals; k++) s += (double)v[k] * (double)v[k];
  return s;
}

template <int P, int OP>
__device__ inline void blas_site(const BlasArgs &a, size_t vs, long s, double *red) {
  typedef typename PT<P>::calc T;
  constexpr int N = kSpinorReals;
  T x[N], y[N];
  const T a0 = (T)a.a[0], a1 = (T)a.a[1];
  if constexpr (OP == B_COPY) {
    load_spinor<P, true>(a.x, vs, s, x);
    store_spinor<P, true>(a.o0, vs, s, x);
  } else if constexpr (OP == B_ZERO) {
#pragma unroll
    for (int k = 0; k < N; k++) x[k] = 0;
    store_spinor<P, true>(a.o0, vs, s, x);
  } else if constexpr (OP == B_AXY) {  // y = a x
    load_spinor<P, true>(a.x, vs, s, x);
#pragma unroll
    for (int k = 0; k < N; k++) y[k] = a0 * x[k];
    store_spinor<P, true>(a.o0, vs, s, y);
  } else if constexpr (OP == B_AX) {  // x = a x
    load_spinor<P, true>(a.o0, vs, s, x);
#pragma unroll
    for (int k = 0; k < N; k++) x[k] = a0 * x[k];
    store_spinor<P, true>(a.o0, vs, s, x);
  } else if constexpr (OP == B_YPEQX || OP == B_YMEQX || OP == B_AYPX || OP == B_AXPY || OP == B_AXPBY ||
                       OP == B_AXPYNORM2 || OP == B_XMY2NORM2 || OP == B_BICG_XMY) {
    load_spinor<P, true>(a.x, vs, s, x);
    load_spinor<P, true>(a.o0, vs, s, y);
#pragma unroll
    for (int k = 0; k < N; k++) {
      if constexpr (OP == B_YPEQX) y[k] = y[k] + x[k];
      if constexpr (OP == B_YMEQX || OP == B_XMY2NORM2) y[k] = y[k] - x[k];
      if constexpr (OP == B_AYPX) y[k] = a0 * y[k] + x[k];
      if constexpr (OP == B_AXPY || OP == B_AXPYNORM2) y[k] = a0 * x[k] + y[k];
      if constexpr (OP == B_AXPBY) y[k] = a0 * x[k] + a1 * y[k];
      if constexpr (OP == B_BICG_XMY) y[k] = x[k] - y[k];
    }
    if constexpr (OP == B_AXPYNORM2 || OP == B_XMY2NORM2) red[0] += sq(y);
    store_spinor<P, true>(a.o0, vs, s, y);
  } else if constexpr (OP == B_NORM2) {
    load_spinor<P, true>(a.x, vs, s, x);
    red[0] += sq(x);
  } else if constexpr (OP == B_XMYNORM2) {  // o0 = x - y
    load_spinor<P, true>(a.x, vs, s, x);
    load_spinor<P, true>(a.y, vs, s, y);
#pragma unroll
    for (int k = 0; k < N; k++) x[k] = x[k] - y[k];
    red[0] += sq(x);
    store_spinor<P, true>(a.o0, vs, s, x);
  } else if constexpr (OP == B_RMAMMP) {  // r -= a mmp ; x += a p
    T r[N], xx[N];
    load_spinor<P, true>(a.x, vs, s, x);   // mmp
    load_spinor<P, true>(a.o0, vs, s, r);  // r
#pragma unroll
    for (int k = 0; k < N; k++) r[k] = r[k] - a0 * x[k];
    red[0] += sq(r);
    store_spinor<P, true>(a.o0, vs, s, r);
    load_spinor<P, true>(a.y, vs, s, y);    // p
    load_spinor<P, true>(a.o1, vs, s, xx);  // x
#pragma unroll
    for (int k = 0; k < N; k++) xx[k] = xx[k] + a0 * y[k];
    store_spinor<P, true>(a.o1, vs, s, xx);
  } else if constexpr (OP == B_RICH) {  // x += dx ; r -= dr
    T t[N];
    load_spinor<P, true>(a.x, vs, s, x);   // dx
    load_spinor<P, true>(a.o0, vs, s, t);  // x
#pragma unroll
    for (int k = 0; k < N; k++) t[k] = t[k] + x[k];
    store_spinor<P, true>(a.o0, vs, s, t);
    load_spinor<P, true>(a.y, vs, s, y);   // dr
    load_spinor<P, true>(a.o1, vs, s, t);  // r
#pragma unroll
    for (int k = 0; k < N; k++) t[k] = t[k] - y[k];
    red[0] += sq(t);
    store_spinor<P, true>(a.o1, vs, s, t);
  } else if constexpr (OP == B_CDOT || OP == B_CDOTNORM) {  // <x, y> = sum conj(x) y
    load_spinor<P, true>(a.x, vs, s, x);
    load_spinor<P, true>(a.y, vs, s, y);
    double re = 0, im = 0;
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      re += (double)x[k] * (double)y[k] + (double)x[k + 1] * (double)y[k + 1];
      im += (double)x[k] * (double)y[k + 1] - (double)x[k + 1] * (double)y[k];
    }
    red[0] += re;
    red[1] += im;
    if constexpr (OP == B_CDOTNORM) red[2] += sq(x);
  } else if constexpr (OP == B_PUPD) {  // p = r + beta (p - omega v); x=r, y=v, o0=p, a = beta, omega
    T v[N];
    load_spinor<P, true>(a.x, vs, s, x);
    load_spinor<P, true>(a.y, vs, s, v);
    load_spinor<P, true>(a.o0, vs, s, y);
    const T br = a0, bi = a1, wr = (T)a.a[2], wi = (T)a.a[3];
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      const T tr = y[k] - (wr * v[k] - wi * v[k + 1]);
      const T ti = y[k + 1] - (wr * v[k + 1] + wi * v[k]);
      y[k] = x[k] + (br * tr - bi * ti);
      y[k + 1] = x[k + 1] + (br * ti + bi * tr);
    }
    store_spinor<P, true>(a.o0, vs, s, y);
  } else if constexpr (OP == B_SUPD || OP == B_CAXPY) {  // s -= alpha v ; y += alpha x
    load_spinor<P, true>(a.x, vs, s, x);
    load_spinor<P, true>(a.o0, vs, s, y);
    const T sg = (OP == B_SUPD) ? (T)-1 : (T)1;
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      const T pr = a0 * x[k] - a1 * x[k + 1], pi = a0 * x[k + 1] + a1 * x[k];
      y[k] = y[k] + sg * pr;
      y[k + 1] = y[k + 1] + sg * pi;
    }
    store_spinor<P, true>(a.o0, vs, s, y);
  } else if constexpr (OP == B_RXUPD || OP == B_RXUPD_CDOT) {
    // x += omega r + alpha p ; r -= omega t  (x=t, y=p, o0=x, o1=r); the _CDOT
    // form also returns <r0, r_new> (z = r0) for the next BiCGStab iteration
    T r[N], xx[N];
    const T wr = a0, wi = a1, ar = (T)a.a[2], ai = (T)a.a[3];
    load_spinor<P, true>(a.o1, vs, s, r);
    load_spinor<P, true>(a.y, vs, s, y);
    load_spinor<P, true>(a.o0, vs, s, xx);
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      xx[k] = xx[k] + (wr * r[k] - wi * r[k + 1]) + (ar * y[k] - ai * y[k + 1]);
      xx[k + 1] = xx[k + 1] + (wr * r[k + 1] + wi * r[k]) + (ar * y[k + 1] + ai * y[k]);
    }
    store_spinor<P, true>(a.o0, vs, s, xx);
    load_spinor<P, true>(a.x, vs, s, x);
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      const T tr = wr * x[k] - wi * x[k + 1], ti = wr * x[k + 1] + wi * x[k];
      r[k] = r[k] - tr;
      r[k + 1] = r[k + 1] - ti;
    }
    red[0] += sq(r);
    store_spinor<P, true>(a.o1, vs, s, r);
    if constexpr (OP == B_RXUPD_CDOT) {
      load_spinor<P, true>(a.z, vs, s, y);
#pragma unroll
      for (int k = 0; k < N; k += 2) {
        red[1] += (double)y[k] * (double)r[k] + (double)y[k + 1] * (double)r[k + 1];
        red[2] += (double)y[k] * (double)r[k + 1] - (double)y[k + 1] * (double)r[k];
      }
    }
  } else if constexpr (OP == B_MRUPD) {  // x += a r ; r -= a Mr   (x = Mr, o0 = x, o1 = r)
    T r[N], xx[N];
    load_spinor<P, true>(a.o1, vs, s, r);
    load_spinor<P, true>(a.o0, vs, s, xx);
    load_spinor<P, true>(a.x, vs, s, x);
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      xx[k] = xx[k] + (a0 * r[k] - a1 * r[k + 1]);
      xx[k + 1] = xx[k + 1] + (a0 * r[k + 1] + a1 * r[k]);
      r[k] = r[k] - (a0 * x[k] - a1 * x[k + 1]);
      r[k + 1] = r[k + 1] - (a0 * x[k + 1] + a1 * x[k]);
    }
    store_spinor<P, true>(a.o0, vs, s, xx);
    store_spinor<P, true>(a.o1, vs, s, r);
  } else if constexpr (OP == B_TWM) {  // y = (a0 + i a1 g5) x
    load_spinor<P, true>(a.x, vs, s, x);
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      const T g = (k < 12) ? a1 : -a1;
      y[k] = a0 * x[k] - g * x[k + 1];
      y[k + 1] = a0 * x[k + 1] + g * x[k];
    }
    store_spinor<P, true>(a.o0, vs, s, y);
  } else if constexpr (OP == B_TWM2) {  // (a0 + i a1 g5 tau3 + eps tau1): x=up, y=dn, o0, o1
    T u[N], d[N], ou[N], od[N];
    const T e = (T)a.a[2];
    load_spinor<P, true>(a.x, vs, s, u);
    load_spinor<P, true>(a.y, vs, s, d);
#pragma unroll
    for (int k = 0; k < N; k += 2) {
      const T g = (k < 12) ? a1 : -a1;
      ou[k] = a0 * u[k] - g * u[k + 1] + e * d[k];
      ou[k + 1] = a0 * u[k + 1] + g * u[k] + e * d[k + 1];
      od[k] = a0 * d[k] + g * d[k + 1] + e * u[k];
      od[k + 1] = a0 * d[k + 1] - g * d[k] + e * u[k + 1];
    }
    store_spinor<P, true>(a.o0, vs, s, ou);
    store_spinor<P, true>(a.o1, vs, s, od);
  } else if constexpr (OP == B_CLOVMULT || OP == B_FULLMULT) {  // o0 = A x  (y = clover)
    cx<T> in[4][3], out[4][3];
    load_spinor<P, true>(a.x, vs, s, x);
    to_spinor(x, in);
    if constexpr (OP == B_CLOVMULT) {
      clover_block_apply<P, 0>(a.y, vs, s, in, out);
      clover_block_apply<P, 1>(a.y, vs, s, in, out);
    } else {
      full_block_apply<P, 0>(a.y, vs, s, in, out);
      full_block_apply<P, 1>(a.y, vs, s, in, out);
    }
    from_spinor(out, y);
    store_spinor<P, true>(a.o0, vs, s, y);
  } else if constexpr (OP == B_TWOFLAV_INVCLOV) {
    // o0 = Fu x + C y ; o1 = Fd y + C x   (x = psi_up, y = psi_dn, z = Fu, w = Fd, o2 = C)
    cx<T> pu[4][3], pd[4][3], r[4][3], c[4][3];
    load_spinor<P, true>(a.x, vs, s, x);
    load_spinor<P, true>(a.y, vs, s, y);
    to_spinor(x, pu);
    to_spinor(y, pd);
    full_block_apply<P, 0>(a.z, vs, s, pu, r);
    full_block_apply<P, 1>(a.z, vs, s, pu, r);
    clover_block_apply<P, 0>(a.o2, vs, s, pd, c);
    clover_block_apply<P, 1>(a.o2, vs, s, pd, c);
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int cc = 0; cc < 3; cc++) r[sp][cc] = r[sp][cc] + c[sp][cc];
    T o[N];
    from_spinor(r, o);
    store_spinor<P, true>(a.o0, vs, s, o);
    full_block_apply<P, 0>(a.w, vs, s, pd, r);
    full_block_apply<P, 1>(a.w, vs, s, pd, r);
    clover_block_apply<P, 0>(a.o2, vs, s, pu, c);
    clover_block_apply<P, 1>(a.o2, vs, s, pu, c);
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int cc = 0; cc < 3; cc++) r[sp][cc] = r[sp][cc] + c[sp][cc];
    from_spinor(r, o);
    store_spinor<P, true>(a.o1, vs, s, o);
  }
}

template <int NR> __device__ inline void block_reduce_store(double *red, double *partials) {
  // partials: this launch's (or this right-hand side's) partial array
  __shared__ double lds[kBlasThreads / 64][NR > 0 ? NR : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double v = red[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) lds[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NR) {
    double v = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) v += lds[w][threadIdx.x];
    partials[blockIdx.x * NR + threadIdx.x] = v;
  }
}

template <int P, int OP>
__global__ void __launch_bounds__(kBlasThreads) k_blas(const BlasArgs a0, size_t vs, long vh, double *partials) {
  constexpr int NR = nred_of(OP);
  BlasArgs a = a0;
  if (a0.dsc) {
    if (a0.dsc[7] != 0.0) return;
#pragma unroll
    for (int k = 0; k < 6; k++) a.a[k] = a0.dsc[k];
  }
  double red[NR > 0 ? NR : 1] = {0};
  const long stride = (long)gridDim.x * blockDim.x;
  for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < vh; s += stride) blas_site<P, OP>(a, vs, s, red);
  if constexpr (NR > 0) block_reduce_store<NR>(red, partials);
}

template <int NR>
__global__ void __launch_bounds__(kBlasThreads) k_finalize(const double *partials, int nblocks, double *out) {
  double red[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) red[k] = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NR; k++) red[k] += partials[b * NR + k];
  __shared__ double lds[kBlasThreads / 64][NR];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double v = red[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) lds[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NR) {
    double v = 0;
    for (int w = 0; w < kBlasThreads / 64; w++) v += lds[w][threadIdx.x];
    out[threadIdx.x] = v;
  }
}

template <int P, int OP>
static void launch_one(const GeomDesc &g, const BlasArgs &a, double *partials, int nb, hipStream_t st) {
  k_blas<P, OP><<<nb, kBlasThreads, 0, st>>>(a, g.vs, g.vh, partials);
}

template <int P>
static hipError_t launch_prec(const GeomDesc &g, int op, const BlasArgs &a, double *partials, int nb, hipStream_t st) {
#define OPCASE(o) \
  case o: launch_one<P, o>(g, a, partials, nb, st); break;
  switch (op) {
    OPCASE(B_COPY) OPCASE(B_ZERO) OPCASE(B_AXY) OPCASE(B_AX) OPCASE(B_YPEQX) OPCASE(B_YMEQX)
    OPCASE(B_AYPX) OPCASE(B_AXPY) OPCASE(B_AXPBY) OPCASE(B_NORM2) OPCASE(B_AXPYNORM2) OPCASE(B_XMYNORM2)
    OPCASE(B_XMY2NORM2) OPCASE(B_RMAMMP) OPCASE(B_RICH) OPCASE(B_BICG_XMY) OPCASE(B_CDOT) OPCASE(B_CDOTNORM)
    OPCASE(B_PUPD) OPCASE(B_SUPD) OPCASE(B_CAXPY) OPCASE(B_RXUPD) OPCASE(B_RXUPD_CDOT) OPCASE(B_MRUPD) OPCASE(B_TWM)
    OPCASE(B_TWM2) OPCASE(B_CLOVMULT) OPCASE(B_FULLMULT) OPCASE(B_TWOFLAV_INVCLOV)
    default: return hipErrorInvalidValue;
  }
#undef OPCASE
  return hipGetLastError();
}

hipError_t blas_launch(const GeomDesc &g, int op, const BlasArgs &a, double *partials, int *nblocks, hipStream_t st) {
  long nb = (g.vh + kBlasThreads - 1) / kBlasThreads;
  if (nb > kMaxBlasBlocks) nb = kMaxBlasBlocks;
  if (nb < 1) nb = 1;
  *nblocks = (int)nb;
  switch (g.prec) {
    case P64: return launch_prec<P64>(g, op, a, partials, (int)nb, st);
    case P32: return launch_prec<P32>(g, op, a, partials, (int)nb, st);
    case P16: return launch_prec<P16>(g, op, a, partials, (int)nb, st);
    case PB16: return launch_prec<PB16>(g, op, a, partials, (int)nb, st);
  }
  return hipErrorInvalidValue;
}

template <int P>
__global__ void __launch_bounds__(kBlasThreads) k_mshift(const MShiftArgs m, size_t vs, long vh) {
  typedef typename PT<P>::calc T;
  constexpr int N = kSpinorReals;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < vh; s += stride) {
    T r[N], v[N], w[N];
    load_spinor<P, true>(m.r, vs, s, r);
    if (m.do_p0) {
      load_spinor<P, true>(m.p0, vs, s, v);
      const T a = (T)m.a;
#pragma unroll
      for (int k = 0; k < N; k++) v[k] = r[k] + a * v[k];
      store_spinor<P, true>(m.p0, vs, s, v);
    }
    for (int q = 0; q < m.nshift; q++) {
      const unsigned bit = 1u << q;
      if (!((m.mask_x | m.mask_p) & bit)) continue;
      load_spinor<P, true>(m.p[q], vs, s, v);
      if (m.mask_x & bit) {
        load_spinor<P, true>(m.x[q], vs, s, w);
        const T b = (T)m.beta[q];
#pragma unroll
        for (int k = 0; k < N; k++) w[k] = w[k] - b * v[k];
        store_spinor<P, true>(m.x[q], vs, s, w);
      }
      if (m.mask_p & bit) {
        const T z = (T)m.zeta[q], as = (T)m.as[q];
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = z * r[k] + as * v[k];
        store_spinor<P, true>(m.p[q], vs, s, v);
      }
    }
  }
}

hipError_t mshift_launch(const GeomDesc &g, const MShiftArgs &m, hipStream_t st) {
  long nb = (g.vh + kBlasThreads - 1) / kBlasThreads;
  if (nb > kMaxBlasBlocks) nb = kMaxBlasBlocks;
  switch (g.prec) {
    case P64: k_mshift<P64><<<(int)nb, kBlasThreads, 0, st>>>(m, g.vs, g.vh); break;
    case P32: k_mshift<P32><<<(int)nb, kBlasThreads, 0, st>>>(m, g.vs, g.vh); break;
    case P16: k_mshift<P16><<<(int)nb, kBlasThreads, 0, st>>>(m, g.vs, g.vh); break;
    case PB16: k_mshift<PB16><<<(int)nb, kBlasThreads, 0, st>>>(m, g.vs, g.vh); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// multishift CG with its scalars in device memory
//
// The scalar kernels evaluate the host loop's FP64 expressions in the same
// order with contraction off, so each operation rounds as the host's does and
// the solve is bit-identical to the one with host scalars.
// ---------------------------------------------------------------------------
template <int P>
__global__ void __launch_bounds__(kBlasThreads) k_mshift_dev(const MShiftArgs m, const MsState *ms, size_t vs,
                                                              long vh) {
  typedef typename PT<P>::calc T;
  constexpr int N = kSpinorReals;
  const unsigned mask_x = ms->mask_x, mask_p = ms->mask_p;
  const int do_p0 = ms->do_p0;
  if (!do_p0 && !mask_x && !mask_p) return;
  const T a = (T)ms->a;
  const long stride = (long)gridDim.x * blockDim.x;
  for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < vh; s += stride) {
    T r[N], v[N], w[N];
    load_spinor<P, true>(m.r, vs, s, r);
    if (do_p0) {
      load_spinor<P, true>(m.p0, vs, s, v);
#pragma unroll
      for (int k = 0; k < N; k++) v[k] = r[k] + a * v[k];
      store_spinor<P, true>(m.p0, vs, s, v);
    }
    for (int q = 0; q < m.nshift; q++) {
      const unsigned bit = 1u << q;
      if (!((mask_x | mask_p) & bit)) continue;
      load_spinor<P, true>(m.p[q], vs, s, v);
      if (mask_x & bit) {
        load_spinor<P, true>(m.x[q], vs, s, w);
        const T b = (T)ms->beta[q];
#pragma unroll
        for (int k = 0; k < N; k++) w[k] = w[k] - b * v[k];
        store_spinor<P, true>(m.x[q], vs, s, w);
      }
      if (mask_p & bit) {
        const T z = (T)ms->zeta[q], as = (T)ms->as[q];
#pragma unroll
        for (int k = 0; k < N; k++) v[k] = z * r[k] + as * v[k];
        store_spinor<P, true>(m.p[q], vs, s, v);
      }
    }
  }
}

hipError_t mshift_dev_launch(const GeomDesc &g, const MShiftArgs &m, const MsState *st_dev, hipStream_t st) {
  long nb = (g.vh + kBlasThreads - 1) / kBlasThreads;
  if (nb > kMaxBlasBlocks) nb = kMaxBlasBlocks;
  switch (g.prec) {
    case P64: k_mshift_dev<P64><<<(int)nb, kBlasThreads, 0, st>>>(m, st_dev, g.vs, g.vh); break;
    case P32: k_mshift_dev<P32><<<(int)nb, kBlasThreads, 0, st>>>(m, st_dev, g.vs, g.vh); break;
    case P16: k_mshift_dev<P16><<<(int)nb, kBlasThreads, 0, st>>>(m, st_dev, g.vs, g.vh); break;
    case PB16: k_mshift_dev<PB16><<<(int)nb, kBlasThreads, 0, st>>>(m, st_dev, g.vs, g.vh); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// minvcg.h:279-297 (a, as_s; the counters as the reference counts them)
__device__ inline void ms_prepare(MsState *ms) {
#pragma clang fp contract(off)
  const double a = ms->c / ms->cp, b = ms->b;
  unsigned mask = 0;
  for (int q = 0; q < ms->nshift; q++)
    if (!ms->conv[q]) {
      ms->as[q] = a * ms->zeta[q] * ms->beta[q] / (ms->zeta_prev[q] * b);
      mask |= 1u << q;
      ms->site_flops += 72 + 48;
    }
  ms->site_flops += 190;
  ms->a = a;
  ms->mask_p = mask;
  ms->do_p0 = 1;
  ms->cp = ms->c;
  ms->bp = b;
}

__global__ void k_ms_pre(MsState *ms) {
  if (ms->done) {
    ms->do_p0 = 0;
    ms->mask_p = 0;
    return;
  }
  ms_prepare(ms);
}

// minvcg.h:316-320
__global__ void k_ms_mid(MsState *ms) {
#pragma clang fp contract(off)
  if (ms->done) return;
  double d = 0;
  for (int f = 0; f < ms->nf; f++) d += ms->cf[f];
  ms->d = d;
  ms->b = -ms->cp / d;
}

__device__ inline void ms_publish(MsState *ms) {
  SolverView *v = ms->hview + (ms->posts & 1);
  v->done = ms->done;
  v->progress = ms->progress;
  __threadfence_system();
  ms->posts++;
}

// minvcg.h:322-384
__global__ void k_ms_post(MsState *ms) {
#pragma clang fp contract(off)
  if (ms->done) {
    // the x_s updates left pending were applied by this iteration's sweep
    ms->mask_x = 0;
    ms->mask_p = 0;
    ms->do_p0 = 0;
    ms_publish(ms);
    return;
  }
  double c = 0;
  for (int f = 0; f < ms->nf; f++) c += ms->cf[f];
  ms->c = c;
  const double a = ms->a, b = ms->b, bp = ms->bp;
  unsigned pending = 0;
  bool all = true;
  double prog = 0;
  for (int q = 0; q < ms->nshift; q++) {
    if (!ms->conv[q]) {
      const double z0 = ms->zeta[q], z1 = ms->zeta_prev[q];
      ms->zeta_prev[q] = z0;
      double z = z0 * z1 * bp;
      z /= (b * a * (z1 - z0) + z1 * bp * (1.0 - ms->shifts[q] * b));
      ms->zeta[q] = z;
      ms->beta[q] = b * z / z0;
      pending |= 1u << q;
      const double css = c * z * z;
      const double ratio = css / ms->rsd_sq[q];
      if (ratio > prog) prog = ratio;
      ms->conv[q] = css < ms->rsd_sq[q];
      if (ms->conv[q]) ms->rsd_sq_final[q] = css;
    }
    all = all && ms->conv[q] != 0;
  }
  ms->mask_x = pending;
  ms->progress = prog;
  const int k = ++ms->iters;
  if (all) {
    ms->done = 1;
    ms->converged = 1;
  } else if (k + 1 >= ms->max_iters) {
    ms->done = 1;
  }
  if (ms->done) {
    ms->mask_p = 0;
    ms->do_p0 = 0;
  } else {
    ms_prepare(ms);
  }
  ms_publish(ms);
}

hipError_t ms_pre_launch(MsState *st_dev, hipStream_t st) {
  k_ms_pre<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}
hipError_t ms_mid_launch(MsState *st_dev, hipStream_t st) {
  k_ms_mid<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}
hipError_t ms_post_launch(MsState *st_dev, hipStream_t st) {
  k_ms_post<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// BiCGStab with its scalars in device memory (invbicgstab.h:150-256)
//
// Each kernel evaluates the host loop's expressions in the same order with
// contraction off (the host's flavour sums start from 0, so do these), so
// the solve is bit-identical to the one with host scalars. A breakdown or the
// end of the solve freezes the state and sets the skip flag of the vector
// updates.
// ---------------------------------------------------------------------------
__device__ inline void bi_cdiv(double *o, const double *a, const double *b) {
#pragma clang fp contract(off)
  const double den = b[0] * b[0] + b[1] * b[1];
  const double re = (a[0] * b[0] + a[1] * b[1]) / den, im = (a[1] * b[0] - a[0] * b[1]) / den;
  o[0] = re;
  o[1] = im;
}
__device__ inline void bi_cmul(double *o, const double *a, const double *b) {
#pragma clang fp contract(off)
  const double re = a[0] * b[0] - a[1] * b[1], im = a[0] * b[1] + a[1] * b[0];
  o[0] = re;
  o[1] = im;
}
__device__ inline void bi_freeze(BiState *b, int status, int n_iters) {
  b->status = status;
  b->n_iters = n_iters;
  b->done = 1;
  b->pupd[7] = b->supd[7] = b->rxupd[7] = 1.0;
}
__device__ inline void bi_publish(BiState *b) {
  SolverView *v = b->hview + (b->posts & 1);
  v->done = b->done;
  v->progress = b->progress;
  __threadfence_system();
  b->posts++;
}

__global__ void k_bi_pre(BiState *b) {
#pragma clang fp contract(off)
  if (b->done) return;
  const int k = b->iters + 1;
  double rho[2] = {0, 0};
  rho[0] += b->ru[1];
  rho[1] += b->ru[2];
  b->rho[0] = rho[0];
  b->rho[1] = rho[1];
  b->site_flops += 8 * b->sf;
  if (rho[0] == 0 && rho[1] == 0) {
    bi_freeze(b, 4 /* QPHIX_EBREAKDOWN */, k);
    return;
  }
  double t1[2], t2[2], beta[2];
  bi_cdiv(t1, rho, b->rho_prev);
  bi_cdiv(t2, b->alpha, b->omega);
  bi_cmul(beta, t1, t2);
  b->pupd[0] = beta[0];
  b->pupd[1] = beta[1];
  b->pupd[2] = b->omega[0];
  b->pupd[3] = b->omega[1];
  b->site_flops += 16 * b->sf;
}

__global__ void k_bi_mid(BiState *b) {
#pragma clang fp contract(off)
  if (b->done) return;
  const int k = b->iters + 1;
  b->mv_apps++;
  double tmp[2] = {0, 0};
  tmp[0] += b->r0v[0];
  tmp[1] += b->r0v[1];
  b->site_flops += 8 * b->sf;
  if (tmp[0] == 0 && tmp[1] == 0) {
    bi_freeze(b, 4, k);
    return;
  }
  bi_cdiv(b->alpha, b->rho, tmp);
  b->rho_prev[0] = b->rho[0];
  b->rho_prev[1] = b->rho[1];
  b->supd[0] = b->alpha[0];
  b->supd[1] = b->alpha[1];
  b->site_flops += 8 * b->sf;
}

__global__ void k_bi_mid2(BiState *b) {
#pragma clang fp contract(off)
  if (b->done) return;
  const int k = b->iters + 1;
  b->mv_apps++;
  double tr[3] = {0, 0, 0};
  for (int q = 0; q < 3; q++) tr[q] += b->tr[q];
  b->site_flops += 4 * b->sf;
  const double t_norm = tr[2];
  if (t_norm == 0) {
    bi_freeze(b, 4, k);
    return;
  }
  b->site_flops += 8 * b->sf;
  b->omega[0] = tr[0] / t_norm;
  b->omega[1] = tr[1] / t_norm;
  b->rxupd[0] = b->omega[0];
  b->rxupd[1] = b->omega[1];
  b->rxupd[2] = b->alpha[0];
  b->rxupd[3] = b->alpha[1];
}

__global__ void k_bi_post(BiState *b) {
#pragma clang fp contract(off)
  if (b->done) {
    bi_publish(b);
    return;
  }
  const int k = ++b->iters;
  b->site_flops += 28 * b->sf;
  double r_norm = 0;
  r_norm += b->ru[0];
  b->r_norm = r_norm;
  b->progress = r_norm / b->rsd_sq;
  if (r_norm < b->rsd_sq) {
    bi_freeze(b, 0, k + 1);
  } else if (k >= b->max_iters) {
    bi_freeze(b, 5 /* QPHIX_ENOCONV */, k + 1);
  }
  bi_publish(b);
}

hipError_t bi_pre_launch(BiState *st_dev, hipStream_t st) {
  k_bi_pre<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}
hipError_t bi_mid_launch(BiState *st_dev, hipStream_t st) {
  k_bi_mid<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}
hipError_t bi_mid2_launch(BiState *st_dev, hipStream_t st) {
  k_bi_mid2<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}
hipError_t bi_post_launch(BiState *st_dev, hipStream_t st) {
  k_bi_post<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}

// CG with its scalars in device memory (invcg.h:293-343): the host loop's
// expressions, contraction off
__device__ inline void cgd_freeze(CgDState *c, int status, int n_iters, double rsd) {
  c->status = status;
  c->n_iters = n_iters;
  c->rsd_sq_final = rsd;
  c->done = 1;
  c->rmammp[7] = c->aypx[7] = 1.0;
}
__global__ void k_cgd_mid(CgDState *c) {
#pragma clang fp contract(off)
  if (c->done) return;
  c->mv_apps += 2;
  double d = 0;
  d += c->d_raw;
  c->site_flops += 4 * c->sf;
  c->rmammp[0] = c->c / d;
}
__global__ void k_cgd_post(CgDState *c) {
#pragma clang fp contract(off)
  if (!c->done) {
    double cp = 0;
    cp += c->r_raw;
    c->cp = cp;
    c->site_flops += 12 * c->sf;
    c->progress = cp / c->rsd_sq;
    const int k = c->iters;
    if (cp < c->rsd_sq) {
      cgd_freeze(c, 0, k, cp);
    } else {
      c->aypx[0] = cp / c->c;
      c->site_flops += 4 * c->sf;
      c->c = cp;
      if (k >= c->max_iters)
        cgd_freeze(c, 5 /* QPHIX_ENOCONV */, c->max_iters, cp);
      else
        c->iters = k + 1;
    }
  }
  SolverView *v = c->hview + (c->posts & 1);
  v->done = c->done;
  v->progress = c->progress;
  __threadfence_system();
  c->posts++;
}
hipError_t cgd_mid_launch(CgDState *st_dev, hipStream_t st) {
  k_cgd_mid<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}
hipError_t cgd_post_launch(CgDState *st_dev, hipStream_t st) {
  k_cgd_post<<<1, 1, 0, st>>>(st_dev);
  return hipGetLastError();
}

// blas_site<P, B_AXPYNORM2> with the scalar read from device memory
template <int P>
__global__ void __launch_bounds__(kBlasThreads) k_axpy_norm2_dev(const double *b, const void *x, void *y, size_t vs,
                                                                  long vh, double *partials) {
  BlasArgs a;
  a.x = x;
  a.y = nullptr;
  a.z = nullptr;
  a.w = nullptr;
  a.o0 = y;
  a.o1 = a.o2 = a.o3 = nullptr;
  a.a[0] = *b;
  for (int k = 1; k < 6; k++) a.a[k] = 0;
  double red[1] = {0};
  const long stride = (long)gridDim.x * blockDim.x;
  for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < vh; s += stride) blas_site<P, B_AXPYNORM2>(a, vs, s, red);
  block_reduce_store<1>(red, partials);
}

hipError_t axpy_norm2_dev_launch(const GeomDesc &g, const double *b, const void *x, void *y, double *partials,
                                 double *out, hipStream_t st) {
  long nb = (g.vh + kBlasThreads - 1) / kBlasThreads;
  if (nb > kMaxBlasBlocks) nb = kMaxBlasBlocks;
  switch (g.prec) {
    case P64: k_axpy_norm2_dev<P64><<<(int)nb, kBlasThreads, 0, st>>>(b, x, y, g.vs, g.vh, partials); break;
    case P32: k_axpy_norm2_dev<P32><<<(int)nb, kBlasThreads, 0, st>>>(b, x, y, g.vs, g.vh, partials); break;
    case P16: k_axpy_norm2_dev<P16><<<(int)nb, kBlasThreads, 0, st>>>(b, x, y, g.vs, g.vh, partials); break;
    case PB16: k_axpy_norm2_dev<PB16><<<(int)nb, kBlasThreads, 0, st>>>(b, x, y, g.vs, g.vh, partials); break;
    default: return hipErrorInvalidValue;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return blas_finalize(partials, (int)nb, 1, out, st);
}

template <int P, int OP>
__global__ void __launch_bounds__(kBlasThreads) k_blas_multi(const MultiBlasArgs m, size_t vs, long vh,
                                                              double *partials) {
  constexpr int NR = nred_of(OP);
  const int r = blockIdx.y;
  BlasArgs a;
  a.x = m.x[r];
  a.y = m.y[r];
  a.z = nullptr;
  a.w = nullptr;
  a.o0 = m.o0[r];
  a.o1 = m.o1[r];
  a.o2 = nullptr;
  a.o3 = nullptr;
  a.a[0] = m.a[r];
  for (int k = 1; k < 6; k++) a.a[k] = 0;
  double red[NR > 0 ? NR : 1] = {0};
  const long stride = (long)gridDim.x * blockDim.x;
  for (long s = (long)blockIdx.x * blockDim.x + threadIdx.x; s < vh; s += stride) blas_site<P, OP>(a, vs, s, red);
  if constexpr (NR > 0) block_reduce_store<NR>(red, partials + (size_t)r * gridDim.x * NR);
}

template <int NR>
__global__ void __launch_bounds__(kBlasThreads) k_finalize_multi(const double *partials, int nblocks, double *out) {
  const int r = blockIdx.x;
  const double *p = partials + (size_t)r * nblocks * NR;
  double red[NR];
#pragma unroll
  for (int k = 0; k < NR; k++) red[k] = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
    for (int k = 0; k < NR; k++) red[k] += p[b * NR + k];
  __shared__ double lds[kBlasThreads / 64][NR];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NR; k++) {
    double v = red[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) lds[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NR) {
    double v = 0;
    for (int w = 0; w < kBlasThreads / 64; w++) v += lds[w][threadIdx.x];
    out[r * NR + threadIdx.x] = v;
  }
}

template <int P>
static hipError_t multi_prec(const GeomDesc &g, int op, const MultiBlasArgs &m, double *partials, int nb,
                             hipStream_t st) {
  const dim3 grid(nb, m.n);
#define MOP(o) \
  case o: k_blas_multi<P, o><<<grid, kBlasThreads, 0, st>>>(m, g.vs, g.vh, partials); break;
  switch (op) {
    MOP(B_COPY) MOP(B_ZERO) MOP(B_NORM2) MOP(B_XMYNORM2) MOP(B_RMAMMP) MOP(B_AYPX) MOP(B_AXPY) MOP(B_AXPYNORM2)
    default: return hipErrorInvalidValue;
  }
#undef MOP
  return hipGetLastError();
}

hipError_t blas_multi_launch(const GeomDesc &g, int op, const MultiBlasArgs &m, double *partials, double *out,
                             hipStream_t st) {
  if (m.n < 1 || m.n > kMaxRhs) return hipErrorInvalidValue;
  long nb = (g.vh + kBlasThreads - 1) / kBlasThreads;
  if (nb > kMaxBlasBlocks / m.n) nb = kMaxBlasBlocks / m.n;
  if (nb < 1) nb = 1;
  hipError_t e;
  switch (g.prec) {
    case P64: e = multi_prec<P64>(g, op, m, partials, (int)nb, st); break;
    case P32: e = multi_prec<P32>(g, op, m, partials, (int)nb, st); break;
    case P16: e = multi_prec<P16>(g, op, m, partials, (int)nb, st); break;
    default: e = multi_prec<PB16>(g, op, m, partials, (int)nb, st); break;
  }
  if (e != hipSuccess) return e;
  const int nr = nred_of(op);
  if (nr == 0) return hipSuccess;
  switch (nr) {
    case 1: k_finalize_multi<1><<<m.n, kBlasThreads, 0, st>>>(partials, (int)nb, out); break;
    case 2: k_finalize_multi<2><<<m.n, kBlasThreads, 0, st>>>(partials, (int)nb, out); break;
    case 3: k_finalize_multi<3><<<m.n, kBlasThreads, 0, st>>>(partials, (int)nb, out); break;
  }
  return hipGetLastError();
}

hipError_t blas_finalize(const double *partials, int nblocks, int nred, double *out, hipStream_t st) {
  switch (nred) {
    case 1: k_finalize<1><<<1, kBlasThreads, 0, st>>>(partials, nblocks, out); break;
    case 2: k_finalize<2><<<1, kBlasThreads, 0, st>>>(partials, nblocks, out); break;
    case 3: k_finalize<3><<<1, kBlasThreads, 0, st>>>(partials, nblocks, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace qphix_amd

// ---------------------------------------------------------------------------
// Streaming calibration: one work-item per site reads NWI 16-byte words of one
// field and writes NWO words of another, both in the site-blocked layout of
// every field, with the hop kernels' load / store policies. It moves a hop
// kernel's per-site traffic without the stencil, which measures the bandwidth
// that traffic shape can reach on the device.
// ---------------------------------------------------------------------------
namespace qphix_amd {

template <int NWI, int NWO>
__global__ void __launch_bounds__(256) k_stream(const f4_t *in, f4_t *out, long n) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  f4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int w = 0; w < NWI; w++) acc += __builtin_nontemporal_load(in + widx(s, w, NWI));
#pragma unroll
  for (int w = 0; w < NWO; w++) __builtin_nontemporal_store(acc * (float)(w + 1), out + widx(s, w, NWO));
}

hipError_t stream_launch(const GeomDesc &g, int nwi, int nwo, const void *in, void *out, hipStream_t st) {
  const unsigned grid = (unsigned)((g.vh + 255) / 256);
  const f4_t *i = (const f4_t *)in;
  f4_t *o = (f4_t *)out;
  if (nwi == 30 && nwo == 6) k_stream<30, 6><<<grid, 256, 0, st>>>(i, o, g.vh);
  else if (nwi == 48 && nwo == 6) k_stream<48, 6><<<grid, 256, 0, st>>>(i, o, g.vh);
  else if (nwi == 66 && nwo == 6) k_stream<66, 6><<<grid, 256, 0, st>>>(i, o, g.vh);
  else if (nwi == 60 && nwo == 6) k_stream<60, 6><<<grid, 256, 0, st>>>(i, o, g.vh);
  else if (nwi == 120 && nwo == 12) k_stream<120, 12><<<grid, 256, 0, st>>>(i, o, g.vh);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace qphix_amd
