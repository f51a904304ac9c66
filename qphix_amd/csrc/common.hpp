// Shared device/host definitions for the gfx950 Dslash library.
//
// Field layout in HBM (site-blocked structure of arrays):
//   A field is W "words" of 16 bytes per site. Sites are grouped in blocks of
//   B = kSiteBlock consecutive checkerboard sites; within a block the W words
//   are stored word-major, so word w of site s lives at 16-byte index
//   ((s / B) * W + w) * B + s % B. One wavefront (64 lanes, 64 consecutive
//   sites) reads each word as one coalesced 1 KiB transaction, and the
//   wavefronts in flight on the device, which take consecutive sites, read
//   each word plane of a block as one contiguous stream: a stream of 48 words
//   in and 6 out per site measured on MI355X runs at 76% of 8 TB/s with
//   B = 4096, 71% with B = 64 (each wavefront's words in a region of their
//   own) and 64% with whole-field word planes (B = all sites).
//   A word holds 2 doubles (F64), 4 floats (F32) or 8 halves (F16). F16 fields
//   carry one float scale per site after the last block
//   (byte offset W * vs * 16 + s * 4); the stored halves are value / scale.
//   vs (sites rounded up to a multiple of B) sizes the allocation.
//
// Per-site real counts (same real order as the canonical host layout):
//   spinor        24   [spin 4][colour 3][re,im]
//   gauge         144  (uncompressed) or 96 (two-row): 8 links, link
//                 l = 2*mu + (0: U_mu(x - mu), 1: U_mu(x)), each [row][col][re,im]
//                 row-major U (rows 0,1 only when two-row compressed)
//   clover        72   per block b: diag[6], then offd[15][re,im] (lower
//                 triangle, offd[i(i-1)/2 + j] = M(i,j), i > j)
//   full clover   144  per block b: M[6][6][re,im]
// The checkerboard site index is s = ((t*Nz + z)*Ny + y)*Nxh + x/2 with
// x = 2*(x/2) + ((y + z + t + cb) & 1), which is the canonical order of the
// QDP++ packers (include/qphix/qdp_packer_parscalar.h:75).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qphix_amd {

// P16: IEEE half, PB16: bfloat16 (the FP32 exponent with an 8-bit
// significand); both store a per-site FP32 scale and compute in FP32
enum Prec { P64 = 0, P32 = 1, P16 = 2, PB16 = 3 };

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));

template <int P> struct PT;
template <> struct PT<P64> {
  typedef double calc;
  static constexpr int per_word = 2;
  static constexpr bool scaled = false;
};
template <> struct PT<P32> {
  typedef float calc;
  static constexpr int per_word = 4;
  static constexpr bool scaled = false;
};
template <> struct PT<P16> {
  typedef float calc;
  static constexpr int per_word = 8;
  static constexpr bool scaled = true;
};
template <> struct PT<PB16> {
  typedef float calc;
  static constexpr int per_word = 8;
  static constexpr bool scaled = true;
};

constexpr int kMaxRhs = 32;  // right-hand sides per multi-RHS launch
constexpr int kSpinorReals = 24;
constexpr int kCloverReals = 72;
constexpr int kFullCloverReals = 144;
__host__ __device__ constexpr int gauge_reals(bool c12) { return c12 ? 96 : 144; }

template <int P> __host__ __device__ constexpr int words_of(int nreal) { return nreal / PT<P>::per_word; }

// bytes of one field with nreal reals per site and plane stride vs
template <int P> __host__ __device__ inline size_t field_bytes(int nreal, size_t vs) {
  return (size_t)words_of<P>(nreal) * vs * 16 + (PT<P>::scaled ? vs * 4 : 0);
}

// ---------------------------------------------------------------------------
// word loads / stores
// ---------------------------------------------------------------------------
typedef double d2_t __attribute__((ext_vector_type(2)));
typedef float f4_t __attribute__((ext_vector_type(4)));

#ifndef QPHIX_SITE_BLOCK_LOG
#define QPHIX_SITE_BLOCK_LOG 12
#endif
constexpr int kSiteBlockLog = QPHIX_SITE_BLOCK_LOG;
constexpr long kSiteBlock = 1l << kSiteBlockLog;  // sites per layout block (a multiple of 64)
static_assert(kSiteBlockLog >= 6, "a layout block holds whole wavefronts");

// 16-byte index of word w of site s in a field of nw words per site
__host__ __device__ inline size_t widx(long s, int w, int nw) {
  return ((size_t)(s >> kSiteBlockLog) * nw + w) * kSiteBlock + (size_t)(s & (kSiteBlock - 1));
}

// NT selects non-temporal (streaming) loads / stores for data touched once
template <int P> struct Word;

template <> struct Word<P64> {
  template <bool NT> __device__ static inline void load(const void *base, size_t i, double *o) {
    const d2_t *p = reinterpret_cast<const d2_t *>(base) + i;
    const d2_t v = NT ? __builtin_nontemporal_load(p) : *p;
    o[0] = v.x;
    o[1] = v.y;
  }
  template <bool NT> __device__ static inline void store(void *base, size_t i, const double *in) {
    d2_t v = {in[0], in[1]};
    d2_t *p = reinterpret_cast<d2_t *>(base) + i;
    if (NT) __builtin_nontemporal_store(v, p); else *p = v;
  }
};

template <> struct Word<P32> {
  template <bool NT> __device__ static inline void load(const void *base, size_t i, float *o) {
    const f4_t *p = reinterpret_cast<const f4_t *>(base) + i;
    const f4_t v = NT ? __builtin_nontemporal_load(p) : *p;
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
    o[3] = v.w;
  }
  template <bool NT> __device__ static inline void store(void *base, size_t i, const float *in) {
    f4_t v = {in[0], in[1], in[2], in[3]};
    f4_t *p = reinterpret_cast<f4_t *>(base) + i;
    if (NT) __builtin_nontemporal_store(v, p); else *p = v;
  }
};

template <> struct Word<P16> {
  // loads raw (unscaled) values; callers multiply by the site scale
  template <bool NT> __device__ static inline void load(const void *base, size_t i, float *o) {
    const half8_t *p = reinterpret_cast<const half8_t *>(base) + i;
    const half8_t v = NT ? __builtin_nontemporal_load(p) : *p;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = (float)v[k];
  }
  template <bool NT> __device__ static inline void store(void *base, size_t i, const float *in) {
    half8_t v;
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = (_Float16)in[k];
    half8_t *p = reinterpret_cast<half8_t *>(base) + i;
    if (NT) __builtin_nontemporal_store(v, p); else *p = v;
  }
};

// bfloat16 words: a value widens to FP32 by a shift; FP32 narrows with
// round-to-nearest-even (v_cvt_pk_bf16_f32)
template <> struct Word<PB16> {
  template <bool NT> __device__ static inline void load(const void *base, size_t i, float *o) {
    const bf8_t *p = reinterpret_cast<const bf8_t *>(base) + i;
    const bf8_t v = NT ? __builtin_nontemporal_load(p) : *p;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = (float)v[k];
  }
  template <bool NT> __device__ static inline void store(void *base, size_t i, const float *in) {
    bf8_t v;
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = (__bf16)in[k];
    bf8_t *p = reinterpret_cast<bf8_t *>(base) + i;
    if (NT) __builtin_nontemporal_store(v, p); else *p = v;
  }
};

template <int P> __device__ inline float site_scale(const void *base, size_t vs, long s, int nwords) {
  const float *sc = reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + (size_t)nwords * vs * 16);
  return sc[s];
}

// An empty asm with a read-write register operand: the value must exist in a
// register as computed, so the instruction that produced it cannot be folded
// into its users. Emits no instruction.
__device__ inline void keep_reg(float &x) { asm("" : "+v"(x)); }

// Load reals [W0*per_word, (W0+NW)*per_word) of a field with NWT words per site.
// SCALED = false leaves FP16 values unscaled (the caller applies site_scale).
// Unscaled half-to-float conversions stay instructions of their own: folded
// into mixed-precision FMAs, which the compiler does or not depending on the
// kernel's shape, they changed the last bits of some results between the
// single-source and register-tiled kernels.
template <int P, int NWT, int W0, int NW, bool NT = false, bool SCALED = true>
__device__ inline void load_range(const void *base, size_t vs, long s, typename PT<P>::calc *o) {
  typedef typename PT<P>::calc T;
#pragma unroll
  for (int w = 0; w < NW; w++) Word<P>::template load<NT>(base, widx(s, W0 + w, NWT), o + w * PT<P>::per_word);
  if constexpr (PT<P>::scaled && !SCALED) {
#pragma unroll
    for (int k = 0; k < NW * PT<P>::per_word; k++) keep_reg(o[k]);
  }
  if constexpr (PT<P>::scaled && SCALED) {
    T sc = site_scale<P>(base, vs, s, NWT);
#pragma unroll
    for (int k = 0; k < NW * PT<P>::per_word; k++) o[k] *= sc;
  }
}

// Store all NWT words of a site (computes the F16 per-site scale).
template <int P, int NWT, bool NT = false>
__device__ inline void store_all(void *base, size_t vs, long s, const typename PT<P>::calc *v) {
  constexpr int N = NWT * PT<P>::per_word;
  if constexpr (PT<P>::scaled) {
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < N; k++) m = fmaxf(m, fabsf(v[k]));
    float inv = m > 0.f ? 1.0f / m : 0.f;
    float tmp[N];
#pragma unroll
    for (int k = 0; k < N; k++) tmp[k] = v[k] * inv;
#pragma unroll
    for (int w = 0; w < NWT; w++) Word<P>::template store<NT>(base, widx(s, w, NWT), tmp + w * PT<P>::per_word);
    float *sc = reinterpret_cast<float *>(reinterpret_cast<char *>(base) + (size_t)NWT * vs * 16);
    sc[s] = m;
  } else {
#pragma unroll
    for (int w = 0; w < NWT; w++) Word<P>::template store<NT>(base, widx(s, w, NWT), v + w * PT<P>::per_word);
  }
}

template <int P, bool NT = false, bool SCALED = true>
__device__ inline void load_spinor(const void *base, size_t vs, long s, typename PT<P>::calc *o) {
  load_range<P, words_of<P>(kSpinorReals), 0, words_of<P>(kSpinorReals), NT, SCALED>(base, vs, s, o);
}
template <int P, bool NT = false>
__device__ inline void store_spinor(void *base, size_t vs, long s, const typename PT<P>::calc *v) {
  store_all<P, words_of<P>(kSpinorReals), NT>(base, vs, s, v);
}

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
template <typename T> struct cx {
  T r, i;
};
template <typename T> __device__ inline cx<T> mk(T r, T i) { cx<T> c; c.r = r; c.i = i; return c; }
template <typename T> __device__ inline cx<T> operator+(cx<T> a, cx<T> b) { return mk(a.r + b.r, a.i + b.i); }
template <typename T> __device__ inline cx<T> operator-(cx<T> a, cx<T> b) { return mk(a.r - b.r, a.i - b.i); }
template <typename T> __device__ inline cx<T> operator*(cx<T> a, cx<T> b) {
  return mk(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r);
}
template <typename T> __device__ inline cx<T> cmulc(cx<T> a, cx<T> b) {  // conj(a) * b
  return mk(a.r * b.r + a.i * b.i, a.r * b.i - a.i * b.r);
}
template <typename T> __device__ inline cx<T> scale(T s, cx<T> a) { return mk(s * a.r, s * a.i); }
template <typename T> __device__ inline cx<T> conj(cx<T> a) { return mk(a.r, -a.i); }

// Complex products and multiply-adds as explicit FMA chains (one rounding per
// FMA, product terms never rounded on their own): a b, conj(a) b, s + a b and
// s + conj(a) b.
__device__ inline float fmat(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ inline double fmat(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <typename T> __device__ inline cx<T> cmul_f(cx<T> a, cx<T> b) {
  return mk(fmat(-a.i, b.i, a.r * b.r), fmat(a.i, b.r, a.r * b.i));
}
template <typename T> __device__ inline cx<T> cmulc_f(cx<T> a, cx<T> b) {
  return mk(fmat(a.i, b.i, a.r * b.r), fmat(-a.i, b.r, a.r * b.i));
}
template <typename T> __device__ inline cx<T> cfma(cx<T> a, cx<T> b, cx<T> s) {
  return mk(fmat(-a.i, b.i, fmat(a.r, b.r, s.r)), fmat(a.i, b.r, fmat(a.r, b.i, s.i)));
}
template <typename T> __device__ inline cx<T> cfmac(cx<T> a, cx<T> b, cx<T> s) {
  return mk(fmat(a.i, b.i, fmat(a.r, b.r, s.r)), fmat(-a.i, b.r, fmat(a.r, b.i, s.i)));
}

// multiply by a compile-time phase in {1, -1, i, -i} given as (PR, PI)
template <int PR, int PI, typename T> __device__ inline cx<T> mulph(cx<T> a) {
  if constexpr (PI == 0) {
    if constexpr (PR == 1) return a;
    else return mk(-a.r, -a.i);
  } else {
    if constexpr (PI == 1) return mk(-a.i, a.r);
    else return mk(a.i, -a.r);
  }
}

// ---------------------------------------------------------------------------
// DeGrand-Rossi gamma matrices as (column, phase) per row
// (the basis of codegen/dslash_common.cc:848-990; gamma_5 = diag(1,1,-1,-1))
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int gcol(int mu, int row) {
  return (mu < 2) ? (3 - row) : ((row + 2) & 3);
}
// phase real / imag parts
__host__ __device__ constexpr int gpr(int mu, int row) {
  return mu == 0 ? 0 : mu == 1 ? ((row == 0 || row == 3) ? -1 : 1) : mu == 2 ? 0 : 1;
}
__host__ __device__ constexpr int gpi(int mu, int row) {
  return mu == 0 ? (row < 2 ? 1 : -1) : mu == 1 ? 0 : mu == 2 ? ((row == 0 || row == 3) ? 1 : -1) : 0;
}

// h = rows 0,1 of (1 + S*gamma_mu) psi, psi as cx[4][3]
template <int MU, int S, typename T>
__device__ inline void project(const cx<T> (&psi)[4][3], cx<T> (&h)[2][3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    h[0][c] = psi[0][c] + mulph<S * gpr(MU, 0), S * gpi(MU, 0)>(psi[gcol(MU, 0)][c]);
    h[1][c] = psi[1][c] + mulph<S * gpr(MU, 1), S * gpi(MU, 1)>(psi[gcol(MU, 1)][c]);
  }
}

// acc += coef * (1 + S*gamma_mu)-reconstruction of the colour-rotated h
template <int MU, int S, typename T>
__device__ inline void recon_acc(cx<T> (&acc)[4][3], const cx<T> (&h)[2][3], T coef) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    cx<T> a = scale(coef, h[0][c]);
    cx<T> b = scale(coef, h[1][c]);
    acc[0][c] = acc[0][c] + a;
    acc[1][c] = acc[1][c] + b;
    // row k (k = 2,3) = S * phase(k) * h[col(k)]
    acc[2][c] = acc[2][c] + mulph<S * gpr(MU, 2), S * gpi(MU, 2)>(gcol(MU, 2) == 0 ? a : b);
    acc[3][c] = acc[3][c] + mulph<S * gpr(MU, 3), S * gpi(MU, 3)>(gcol(MU, 3) == 0 ? a : b);
  }
}

}  // namespace qphix_amd
