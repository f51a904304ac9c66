// Block (multi-RHS) BLAS for block Krylov solvers: the Gram matrix of two sets
// of spinor fields and the combination of one set by a small dense matrix.
//
// Both kernels walk the fields 64 sites and one 16-byte word at a time: the
// word of every right-hand side is staged in LDS with coalesced 1 KiB loads, so
// each work-item then combines all right-hand sides of its site without
// re-reading HBM. Gram products are summed in the storage precision over one
// word of 64 sites and then in FP64; per-work-group partials are
// summed in a fixed order, so results do not depend on timing.
#include "blockblas.hpp"

namespace qphix_amd {

constexpr int kGramThreads = 256;
constexpr int kGramPad = 65;  // LDS row stride in words: 64 sites + 1 (no bank conflicts)

template <int P>
__global__ void __launch_bounds__(kGramThreads) k_block_gram(const BlockArgs ba, int nsb, double *partials) {
  typedef typename PT<P>::calc T;
  constexpr int W = words_of<P>(kSpinorReals);
  constexpr int PW = PT<P>::per_word;
  constexpr int CW = PW / 2;
  __shared__ T la[kMaxRhs * kGramPad * PW];
  __shared__ T lb[kMaxRhs * kGramPad * PW];
  // pointer tables read in place from the kernel-argument segment (BlockArgs
  // is the first parameter): a run-time index into the by-value parameter
  // would copy it to private memory
  const BlockArgs *ka = (const BlockArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  const int na = ba.na, nb = ba.nb;
  const int tid = threadIdx.x;
  const int ti = tid & 15, tj = tid >> 4;  // this work-item's pairs: rows 2ti..2ti+1, columns 2tj..2tj+1
  const bool mine = 2 * ti < na && 2 * tj < nb;
  double acc[2][2][2];
#pragma unroll
  for (int u = 0; u < 2; u++)
#pragma unroll
    for (int v = 0; v < 2; v++) acc[u][v][0] = acc[u][v][1] = 0.0;

  for (int sb = blockIdx.x; sb < nsb; sb += gridDim.x) {
    for (int w = 0; w < W; w++) {
      __syncthreads();
      for (int idx = tid; idx < (na + nb) * 64; idx += kGramThreads) {
        const int f = idx >> 6, l = idx & 63;
        T v[PW];
        if (f < na) {
          Word<P>::template load<true>(ka->a[f], widx((long)sb * 64 + l, w, W), v);
#pragma unroll
          for (int k = 0; k < PW; k++) la[(f * kGramPad + l) * PW + k] = v[k];
        } else {
          Word<P>::template load<true>(ka->b[f - na], widx((long)sb * 64 + l, w, W), v);
#pragma unroll
          for (int k = 0; k < PW; k++) lb[((f - na) * kGramPad + l) * PW + k] = v[k];
        }
      }
      __syncthreads();
      if (mine) {
        T s[2][2][2];
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
          for (int v = 0; v < 2; v++) s[u][v][0] = s[u][v][1] = 0;
        for (int l = 0; l < 64; l++) {
          T va[2][PW], vb[2][PW];
#pragma unroll
          for (int u = 0; u < 2; u++)
#pragma unroll
            for (int k = 0; k < PW; k++) {
              va[u][k] = la[((2 * ti + u) * kGramPad + l) * PW + k];
              vb[u][k] = lb[((2 * tj + u) * kGramPad + l) * PW + k];
            }
#pragma unroll
          for (int u = 0; u < 2; u++)
#pragma unroll
            for (int v = 0; v < 2; v++)
#pragma unroll
              for (int c = 0; c < CW; c++) {
                const T ar = va[u][2 * c], ai = va[u][2 * c + 1], br = vb[v][2 * c], bi = vb[v][2 * c + 1];
                s[u][v][0] += ar * br + ai * bi;  // conj(a) b
                s[u][v][1] += ar * bi - ai * br;
              }
        }
#pragma unroll
        for (int u = 0; u < 2; u++)
#pragma unroll
          for (int v = 0; v < 2; v++) {
            acc[u][v][0] += (double)s[u][v][0];
            acc[u][v][1] += (double)s[u][v][1];
          }
      }
    }
  }
  if (mine) {
    double *o = partials + (size_t)blockIdx.x * kMaxRhs * kMaxRhs * 2;
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int v = 0; v < 2; v++) {
        const int i = 2 * ti + u, j = 2 * tj + v;
        if (i < na && j < nb) {
          o[(i * kMaxRhs + j) * 2] = acc[u][v][0];
          o[(i * kMaxRhs + j) * 2 + 1] = acc[u][v][1];
        }
      }
  }
}

// out[(i * nb + j) * 2 + {re, im}] = sum over work-groups, in order
__global__ void __launch_bounds__(256) k_block_gram_finalize(const double *partials, int nblocks, int na, int nb,
                                                             double *out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= na * nb * 2) return;
  const int ij = e >> 1, ri = e & 1;
  const int i = ij / nb, j = ij - i * nb;
  double s = 0;
  for (int b = 0; b < nblocks; b++) s += partials[((size_t)b * kMaxRhs * kMaxRhs + i * kMaxRhs + j) * 2 + ri];
  out[e] = s;
}

template <int P>
__global__ void __launch_bounds__(kGramThreads) k_block_combine(const BlockArgs ba, int nsb, const double *mat) {
  typedef typename PT<P>::calc T;
  constexpr int W = words_of<P>(kSpinorReals);
  constexpr int PW = PT<P>::per_word;
  constexpr int CW = PW / 2;
  __shared__ T la[kMaxRhs * 64 * PW];
  __shared__ T lm[kMaxRhs * kMaxRhs * 2];
  const BlockArgs *ka = (const BlockArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  const int na = ba.na, nb = ba.nb;
  const int tid = threadIdx.x;
  for (int e = tid; e < na * nb * 2; e += kGramThreads) lm[e] = (T)mat[e];
  const int l = tid & 63, jb = tid >> 6;
  for (int sb = blockIdx.x; sb < nsb; sb += gridDim.x) {
    for (int w = 0; w < W; w++) {
      __syncthreads();  // also orders the matrix staging before first use
      for (int idx = tid; idx < na * 64; idx += kGramThreads) {
        const int f = idx >> 6, ll = idx & 63;
        T v[PW];
        Word<P>::template load<false>(ka->a[f], widx((long)sb * 64 + ll, w, W), v);
#pragma unroll
        for (int k = 0; k < PW; k++) la[(f * 64 + ll) * PW + k] = v[k];
      }
      __syncthreads();
      const size_t wi = widx((long)sb * 64 + l, w, W);
      for (int j = jb; j < nb; j += kGramThreads / 64) {
        T o[PW];
        const void *y = ka->y[j];
        if (y) {
          Word<P>::template load<false>(y, wi, o);
        } else {
#pragma unroll
          for (int k = 0; k < PW; k++) o[k] = 0;
        }
        for (int i = 0; i < na; i++) {
          const T mr = lm[(i * nb + j) * 2], mi = lm[(i * nb + j) * 2 + 1];
#pragma unroll
          for (int c = 0; c < CW; c++) {
            const T ar = la[(i * 64 + l) * PW + 2 * c], ai = la[(i * 64 + l) * PW + 2 * c + 1];
            o[2 * c] += ar * mr - ai * mi;
            o[2 * c + 1] += ar * mi + ai * mr;
          }
        }
        Word<P>::template store<false>(ka->o[j], wi, o);
      }
    }
  }
}

static int gram_grid(int nsb) { return nsb < kMaxGramBlocks ? nsb : kMaxGramBlocks; }

hipError_t block_gram_launch(const GeomDesc &g, const BlockArgs &ba, double *partials, double *out, hipStream_t st) {
  if (ba.na < 1 || ba.nb < 1 || ba.na > kMaxRhs || ba.nb > kMaxRhs) return hipErrorInvalidValue;
  const int nsb = (int)((g.vh + 63) / 64);
  const int grid = gram_grid(nsb);
  switch (g.prec) {
    case P64: k_block_gram<P64><<<grid, kGramThreads, 0, st>>>(ba, nsb, partials); break;
    case P32: k_block_gram<P32><<<grid, kGramThreads, 0, st>>>(ba, nsb, partials); break;
    default: return hipErrorInvalidValue;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int ne = ba.na * ba.nb * 2;
  k_block_gram_finalize<<<(ne + 255) / 256, 256, 0, st>>>(partials, grid, ba.na, ba.nb, out);
  return hipGetLastError();
}

hipError_t block_combine_launch(const GeomDesc &g, const BlockArgs &ba, const double *mat, hipStream_t st) {
  if (ba.na < 1 || ba.nb < 1 || ba.na > kMaxRhs || ba.nb > kMaxRhs) return hipErrorInvalidValue;
  const int nsb = (int)((g.vh + 63) / 64);
  const int grid = nsb < 2048 ? nsb : 2048;
  switch (g.prec) {
    case P64: k_block_combine<P64><<<grid, kGramThreads, 0, st>>>(ba, nsb, mat); break;
    case P32: k_block_combine<P32><<<grid, kGramThreads, 0, st>>>(ba, nsb, mat); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace qphix_amd
