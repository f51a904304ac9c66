// Multi-RHS hop with the link x RHS-block contraction on the matrix cores
// (FP32 storage and arithmetic).
//
// A work-group takes 64 target sites and 8 right-hand sides. For each of the
// eight hops it stages in LDS the projected neighbour half spinors of all
// (site, RHS) pairs and the 64 links, then every wave contracts, two sites at
// a time, the real 12 x 12 block-diagonal form of the two links with the
// 12 x 16 block of half spinors (columns: 8 RHS x 2 spin components) by three
// v_mfma_f32_16x16x4f32 steps, and the products go back through LDS to the
// work-items that own the (site, RHS) accumulators. The epilogues of the
// kernel family are the ones of the vector kernels.
//
// Operand layouts of v_mfma_f32_16x16x4f32 (checked on the device by a
// lane-layout probe): lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16];
// D[4 (l / 16) + r][l % 16] comes back in register r of lane l.
#pragma once

#include "hop_impl.hpp"

namespace qphix_amd {

typedef float mf4_t __attribute__((ext_vector_type(4)));

constexpr int kMfmaRhs = 8;     // right-hand sides per work-group (16 MFMA columns)
constexpr int kMfmaSites = 64;  // target sites per work-group

// entry (ri, ck) of the real 6 x 6 form of U (FWD) or U^+ acting on
// (Re h_0..2, Im h_0..2); u is the link row-major, complex interleaved
template <bool FWD, typename TU>
__device__ inline float uhat(const TU *u, int ri, int ck) {
  const int c = ri < 3 ? ri : ri - 3, d = ck < 3 ? ck : ck - 3;
  const bool im_row = ri >= 3, im_col = ck >= 3;
  // FWD: W = U h;  otherwise W = U^+ h with (U^+)[c][d] = conj(U[d][c])
  const int e = FWD ? (3 * c + d) : (3 * d + c);
  const float ur = (float)u[2 * e], ui = FWD ? (float)u[2 * e + 1] : -(float)u[2 * e + 1];
  if (!im_row) return im_col ? -ui : ur;
  return im_col ? ur : ui;
}

// one hop (direction MU, forward or backward) for the work-group's 64 x 8
// (site, RHS) pairs; each work-item owns sites sl, RHS j0, j0 + 1
template <bool C12, int ISIGN, int MU, bool FWD>
__device__ inline void mfma_hop(const HopArgs &a, const void *const (&psi)[2], bool active, unsigned s, unsigned sn,
                                float coef, float tr, float ti, int sl, int j0, float (*lh)[kMfmaRhs][12],
                                float (*lu)[18], float (*lw)[kMfmaRhs][12], cx<float> (&acc)[2][4][3]) {
  typedef GaugeLayout<P32, C12> GL;
  constexpr int S = FWD ? -ISIGN : ISIGN;
  const int tid = threadIdx.x;
  // stage: projected neighbours of this work-item's two right-hand sides, and
  // (first wave) the links of the 64 sites
#pragma unroll
  for (int jj = 0; jj < 2; jj++) {
    float hv[12];
    if (active) {
      float sp[kSpinorReals];
      cx<float> ps[4][3], h[2][3];
      load_spinor<P32>(psi[jj], a.vs, sn, sp);
      to_spinor(sp, ps);
      project<MU, S>(ps, h);
#pragma unroll
      for (int sh = 0; sh < 2; sh++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          hv[(sh * 3 + c) * 2] = h[sh][c].r;
          hv[(sh * 3 + c) * 2 + 1] = h[sh][c].i;
        }
    } else {
#pragma unroll
      for (int e = 0; e < 12; e++) hv[e] = 0.f;
    }
#pragma unroll
    for (int e = 0; e < 12; e++) lh[sl][j0 + jj][e] = hv[e];
  }
  if (tid < kMfmaSites) {
    cx<float> U[3][3];
    if (active) {
      if constexpr (GL::per_link) {
        float lv[GL::link_reals];
        load_range<P32, GL::nw, (2 * MU + (FWD ? 1 : 0)) * GL::link_words, GL::link_words, true>(a.u, a.vs, s, lv);
        make_link<C12>(lv, U);
      } else {
        float lv[2 * GL::link_reals];
        load_range<P32, GL::nw, MU * GL::pair_words, GL::pair_words, true>(a.u, a.vs, s, lv);
        make_link<C12>(lv + (FWD ? GL::link_reals : 0), U);
      }
    } else {
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) U[r][c] = mk(0.f, 0.f);
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        lu[sl][(3 * r + c) * 2] = U[r][c].r;
        lu[sl][(3 * r + c) * 2 + 1] = U[r][c].i;
      }
  }
  __syncthreads();
  // contraction: wave w takes the site pairs 8w .. 8w + 7
  {
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, j = lane & 15;
    const int rhs = j >> 1, sh = j & 1;
    for (int q = 0; q < 8; q++) {
      const int sp = wave * 8 + q;
      mf4_t d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int step = 0; step < 3; step++) {
        const int kk = (lane >> 4) + 4 * step;  // 0..11: (site in pair, Re/Im, colour)
        const int sk = kk / 6, ck = kk - 6 * sk;
        float av = 0.f;
        if (i < 12 && i / 6 == sk) av = uhat<FWD>(lu[2 * sp + sk], i - 6 * sk, ck);
        const float bv = lh[2 * sp + sk][rhs][(sh * 3 + (ck % 3)) * 2 + ck / 3];
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, d, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = 4 * (lane >> 4) + r;
        if (row < 12) {
          const int si = row / 6, rr = row - 6 * si;
          lw[2 * sp + si][rhs][(sh * 3 + (rr % 3)) * 2 + rr / 3] = d[r];
        }
      }
    }
  }
  __syncthreads();
  // reconstruct into the owners' accumulators
  if (active) {
    const cx<float> ph = mk(tr, FWD ? ti : -ti);
#pragma unroll
    for (int jj = 0; jj < 2; jj++) {
      cx<float> uh[2][3];
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int c = 0; c < 3; c++)
          uh[h][c] = ph * mk(lw[sl][j0 + jj][(h * 3 + c) * 2], lw[sl][j0 + jj][(h * 3 + c) * 2 + 1]);
      recon_acc<MU, S>(acc[jj], uh, coef);
    }
  }
  __syncthreads();
}

template <bool C12, int ISIGN, int MU>
__device__ inline void mfma_dim(const HopArgs &a, const void *const (&psi)[2], bool active, unsigned s, unsigned sb,
                                unsigned sf, float cb_coef, float cf_coef, int sl, int j0, float (*lh)[kMfmaRhs][12],
                                float (*lu)[18], float (*lw)[kMfmaRhs][12], cx<float> (&acc)[2][4][3]) {
  const bool tbc = (a.tbc_mask >> MU) & 1;
  const float tr = tbc ? (float)a.tbc_r[MU] : 1.f, ti = tbc ? (float)a.tbc_i[MU] : 0.f;
  mfma_hop<C12, ISIGN, MU, false>(a, psi, active, s, sb, cb_coef, tr, ti, sl, j0, lh, lu, lw, acc);
  mfma_hop<C12, ISIGN, MU, true>(a, psi, active, s, sf, cf_coef, tr, ti, sl, j0, lh, lu, lw, acc);
}

// a.nrhs (a multiple of 8) right-hand sides in a.psis; grid = site blocks x RHS groups
template <bool C12, int ISIGN, int KIND>
__global__ void __launch_bounds__(256) hop_mfma_kernel(const HopArgs a) {
  __shared__ float lh[kMfmaSites][kMfmaRhs][12];
  __shared__ float lu[kMfmaSites][18];
  __shared__ float lw[kMfmaSites][kMfmaRhs][12];
  if (a.skip && *a.skip) return;
  int lb = xcd_block();
  const int ng = a.nrhs / kMfmaRhs;
  const int g = lb % ng;
  lb /= ng;
  const int tid = threadIdx.x;
  const int sl = tid & 63, j0 = 2 * (tid >> 6);
  const unsigned k = (unsigned)lb * kMfmaSites + sl;
  const bool active = k < (unsigned)a.vh;
  const HopArgs *ka = (const HopArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  const void *psi[2], *chi[2];
  void *out[2];
#pragma unroll
  for (int jj = 0; jj < 2; jj++) {
    const int r = g * kMfmaRhs + j0 + jj;
    psi[jj] = ka->psis[r];
    out[jj] = ka->outs[r];
    chi[jj] = ka->chis[r];
  }
  SiteNbrs n;
  n.s = n.xb = n.xf = n.yb = n.yf = n.zb = n.zf = n.tb = n.tf = 0;
  n.t_lo = n.t_hi = false;
  if (active) n = site_nbrs(a, k);
  const unsigned s = n.s;
  const float cs = (float)a.cs;
  const float ctb = (float)(n.t_lo ? a.ctb : a.ct);
  const float ctf = (float)(n.t_hi ? a.ctf : a.ct);
  cx<float> acc[2][4][3];
#pragma unroll
  for (int jj = 0; jj < 2; jj++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[jj][sp][c] = mk(0.f, 0.f);
  if (active) hop_init_r<P32, ISIGN, KIND, 2>(a, s, chi, acc);
  mfma_dim<C12, ISIGN, 0>(a, psi, active, s, n.xb, n.xf, cs, cs, sl, j0, lh, lu, lw, acc);
  mfma_dim<C12, ISIGN, 1>(a, psi, active, s, n.yb, n.yf, cs, cs, sl, j0, lh, lu, lw, acc);
  mfma_dim<C12, ISIGN, 2>(a, psi, active, s, n.zb, n.zf, cs, cs, sl, j0, lh, lu, lw, acc);
  mfma_dim<C12, ISIGN, 3>(a, psi, active, s, n.tb, n.tf, ctb, ctf, sl, j0, lh, lu, lw, acc);
  if (active) hop_finish_r<P32, ISIGN, KIND, 2>(a, s, acc, out);
}

}  // namespace qphix_amd

namespace qphix_amd {

template <bool C12, int ISIGN>
static hipError_t launch_mfma2(int kind, const HopArgs &a, hipStream_t st) {
  const dim3 grid((unsigned)(((a.vh + kMfmaSites - 1) / kMfmaSites) * (a.nrhs / kMfmaRhs)));
  switch (kind) {
#define MK(K) \
  case K: hop_mfma_kernel<C12, ISIGN, K><<<grid, 256, 0, st>>>(a); break;
    MK(KD_WILSON) MK(KA_WILSON) MK(KD_CLOVER) MK(KA_CLOVER) MK(KD_TM) MK(KA_TM) MK(KD_FULL)
#undef MK
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// FP32 storage only; a.nrhs a positive multiple of 8
inline hipError_t launch_hop_mfma(int kind, bool c12, int isign, const HopArgs &a, hipStream_t st) {
  if (a.nrhs < kMfmaRhs || a.nrhs % kMfmaRhs) return hipErrorInvalidValue;
  if (c12) return isign > 0 ? launch_mfma2<true, 1>(kind, a, st) : launch_mfma2<true, -1>(kind, a, st);
  return isign > 0 ? launch_mfma2<false, 1>(kind, a, st) : launch_mfma2<false, -1>(kind, a, st);
}

}  // namespace qphix_amd

// The same contraction for FP16 storage on v_mfma_f32_16x16x16f16: one MFMA
// per site pair and hop covers K = 12 (two sites x 6 reals) in one step. The
// operands are the stored halves without their per-site scales: the projected
// neighbour half spinors of the unscaled source halves (rounded to half once
// more) and the links with their third row rebuilt times the link scale, so
// that each product carries one factor of the link scale and one of the
// source's, which multiply the term through its boundary phase as in the
// vector kernels. Operand layout of v_mfma_f32_16x16x16f16 (checked on the
// device by a lane-layout probe): lane l supplies A[l % 16][4 (l / 16) + v]
// and B[4 (l / 16) + v][l % 16] for v = 0..3; D as for the FP32 form.
// Results agree with the vector kernels to the half rounding of the
// operands (about 1e-3), not bit for bit.
namespace qphix_amd {

typedef _Float16 mh4_t __attribute__((ext_vector_type(4)));

template <int ISIGN, int MU, bool FWD>
__device__ inline void mfma16_hop(const HopArgs &a, const void *const (&psi)[2], bool active, unsigned s, unsigned sn,
                                  float coef, float tr, float ti, int sl, int j0, _Float16 (*lh)[kMfmaRhs][12],
                                  float (*ls)[kMfmaRhs], _Float16 (*lu)[18], float *lus,
                                  float (*lw)[kMfmaRhs][12], cx<float> (&acc)[2][4][3]) {
  typedef GaugeLayout<P16, true> GL;
  constexpr int S = FWD ? -ISIGN : ISIGN;
  const int tid = threadIdx.x;
#pragma unroll
  for (int jj = 0; jj < 2; jj++) {
    float hv[12];
    float sc = 0.f;
    if (active) {
      float sp[kSpinorReals];
      cx<float> ps[4][3], h[2][3];
      load_spinor<P16, false, false>(psi[jj], a.vs, sn, sp);
      sc = site_scale<P16>(psi[jj], a.vs, sn, words_of<P16>(kSpinorReals));
      to_spinor(sp, ps);
      project<MU, S>(ps, h);
#pragma unroll
      for (int sh = 0; sh < 2; sh++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          hv[(sh * 3 + c) * 2] = h[sh][c].r;
          hv[(sh * 3 + c) * 2 + 1] = h[sh][c].i;
        }
    } else {
#pragma unroll
      for (int e = 0; e < 12; e++) hv[e] = 0.f;
    }
#pragma unroll
    for (int e = 0; e < 12; e++) lh[sl][j0 + jj][e] = (_Float16)hv[e];
    ls[sl][j0 + jj] = sc;
  }
  if (tid < kMfmaSites) {
    cx<float> U[3][3];
    float su = 0.f;
    if (active) {
      float lv[2 * GL::link_reals];
      load_range<P16, GL::nw, MU * GL::pair_words, GL::pair_words, true, false>(a.u, a.vs, s, lv);
      su = site_scale<P16>(a.u, a.vs, s, GL::nw);
      make_link<true>(lv + (FWD ? GL::link_reals : 0), U, su);
    } else {
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) U[r][c] = mk(0.f, 0.f);
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        lu[sl][(3 * r + c) * 2] = (_Float16)U[r][c].r;
        lu[sl][(3 * r + c) * 2 + 1] = (_Float16)U[r][c].i;
      }
    lus[sl] = su;
  }
  __syncthreads();
  // contraction: wave w takes the site pairs 8w .. 8w + 7
  {
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, j = lane & 15;
    const int rhs = j >> 1, sh = j & 1;
    for (int q = 0; q < 8; q++) {
      const int sp = wave * 8 + q;
      mh4_t av, bv;
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int kk = 4 * (lane >> 4) + v;  // 0..15: (site in pair, Re/Im, colour), 12..15 padding
        float x = 0.f, y = 0.f;
        if (kk < 12) {
          const int sk = kk / 6, ck = kk - 6 * sk;
          if (i < 12 && i / 6 == sk) x = uhat<FWD>(lu[2 * sp + sk], i - 6 * sk, ck);
          y = (float)lh[2 * sp + sk][rhs][(sh * 3 + (ck % 3)) * 2 + ck / 3];
        }
        av[v] = (_Float16)x;
        bv[v] = (_Float16)y;
      }
      mf4_t d = {0.f, 0.f, 0.f, 0.f};
      d = __builtin_amdgcn_mfma_f32_16x16x16f16(av, bv, d, 0, 0, 0);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = 4 * (lane >> 4) + r;
        if (row < 12) {
          const int si = row / 6, rr = row - 6 * si;
          lw[2 * sp + si][rhs][(sh * 3 + (rr % 3)) * 2 + rr / 3] = d[r];
        }
      }
    }
  }
  __syncthreads();
  // reconstruct into the owners' accumulators; the link and source scales
  // multiply the term through its boundary phase
  if (active) {
#pragma unroll
    for (int jj = 0; jj < 2; jj++) {
      const float f = lus[sl] * ls[sl][j0 + jj];
      const cx<float> ph = mk(tr * f, FWD ? ti * f : -(ti * f));
      cx<float> uh[2][3];
#pragma unroll
      for (int h = 0; h < 2; h++)
#pragma unroll
        for (int c = 0; c < 3; c++)
          uh[h][c] = ph * mk(lw[sl][j0 + jj][(h * 3 + c) * 2], lw[sl][j0 + jj][(h * 3 + c) * 2 + 1]);
      recon_acc<MU, S>(acc[jj], uh, coef);
    }
  }
  __syncthreads();
}

template <int ISIGN, int MU>
__device__ inline void mfma16_dim(const HopArgs &a, const void *const (&psi)[2], bool active, unsigned s, unsigned sb,
                                  unsigned sf, float cb_coef, float cf_coef, int sl, int j0,
                                  _Float16 (*lh)[kMfmaRhs][12], float (*ls)[kMfmaRhs], _Float16 (*lu)[18], float *lus,
                                  float (*lw)[kMfmaRhs][12], cx<float> (&acc)[2][4][3]) {
  const bool tbc = (a.tbc_mask >> MU) & 1;
  const float tr = tbc ? (float)a.tbc_r[MU] : 1.f, ti = tbc ? (float)a.tbc_i[MU] : 0.f;
  mfma16_hop<ISIGN, MU, false>(a, psi, active, s, sb, cb_coef, tr, ti, sl, j0, lh, ls, lu, lus, lw, acc);
  mfma16_hop<ISIGN, MU, true>(a, psi, active, s, sf, cf_coef, tr, ti, sl, j0, lh, ls, lu, lus, lw, acc);
}

// FP16 storage (two-row links), a.nrhs (a multiple of 8) right-hand sides in
// a.psis; grid = site blocks x RHS groups
template <int ISIGN, int KIND>
__global__ void __launch_bounds__(256) hop_mfma16_kernel(const HopArgs a) {
  __shared__ _Float16 lh[kMfmaSites][kMfmaRhs][12];
  __shared__ float ls[kMfmaSites][kMfmaRhs];
  __shared__ _Float16 lu[kMfmaSites][18];
  __shared__ float lus[kMfmaSites];
  __shared__ float lw[kMfmaSites][kMfmaRhs][12];
  if (a.skip && *a.skip) return;
  int lb = xcd_block();
  const int ng = a.nrhs / kMfmaRhs;
  const int g = lb % ng;
  lb /= ng;
  const int tid = threadIdx.x;
  const int sl = tid & 63, j0 = 2 * (tid >> 6);
  const unsigned k = (unsigned)lb * kMfmaSites + sl;
  const bool active = k < (unsigned)a.vh;
  const HopArgs *ka = (const HopArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  const void *psi[2], *chi[2];
  void *out[2];
#pragma unroll
  for (int jj = 0; jj < 2; jj++) {
    const int r = g * kMfmaRhs + j0 + jj;
    psi[jj] = ka->psis[r];
    out[jj] = ka->outs[r];
    chi[jj] = ka->chis[r];
  }
  SiteNbrs n;
  n.s = n.xb = n.xf = n.yb = n.yf = n.zb = n.zf = n.tb = n.tf = 0;
  n.t_lo = n.t_hi = false;
  if (active) n = site_nbrs(a, k);
  const unsigned s = n.s;
  const float cs = (float)a.cs;
  const float ctb = (float)(n.t_lo ? a.ctb : a.ct);
  const float ctf = (float)(n.t_hi ? a.ctf : a.ct);
  cx<float> acc[2][4][3];
#pragma unroll
  for (int jj = 0; jj < 2; jj++)
#pragma unroll
    for (int sp = 0; sp < 4; sp++)
#pragma unroll
      for (int c = 0; c < 3; c++) acc[jj][sp][c] = mk(0.f, 0.f);
  if (active) hop_init_r<P16, ISIGN, KIND, 2>(a, s, chi, acc);
  mfma16_dim<ISIGN, 0>(a, psi, active, s, n.xb, n.xf, cs, cs, sl, j0, lh, ls, lu, lus, lw, acc);
  mfma16_dim<ISIGN, 1>(a, psi, active, s, n.yb, n.yf, cs, cs, sl, j0, lh, ls, lu, lus, lw, acc);
  mfma16_dim<ISIGN, 2>(a, psi, active, s, n.zb, n.zf, cs, cs, sl, j0, lh, ls, lu, lus, lw, acc);
  mfma16_dim<ISIGN, 3>(a, psi, active, s, n.tb, n.tf, ctb, ctf, sl, j0, lh, ls, lu, lus, lw, acc);
  if (active) hop_finish_r<P16, ISIGN, KIND, 2>(a, s, acc, out);
}

template <int ISIGN>
static hipError_t launch_mfma16_2(int kind, const HopArgs &a, hipStream_t st) {
  const dim3 grid((unsigned)(((a.vh + kMfmaSites - 1) / kMfmaSites) * (a.nrhs / kMfmaRhs)));
  switch (kind) {
#define MK(K) \
  case K: hop_mfma16_kernel<ISIGN, K><<<grid, 256, 0, st>>>(a); break;
    MK(KD_WILSON) MK(KA_WILSON) MK(KD_CLOVER) MK(KA_CLOVER) MK(KD_TM) MK(KA_TM) MK(KD_FULL)
#undef MK
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// FP16 storage with two-row links only; a.nrhs a positive multiple of 8
inline hipError_t launch_hop_mfma16(int kind, bool c12, int isign, const HopArgs &a, hipStream_t st) {
  if (!c12 || a.nrhs < kMfmaRhs || a.nrhs % kMfmaRhs) return hipErrorInvalidValue;
  return isign > 0 ? launch_mfma16_2<1>(kind, a, st) : launch_mfma16_2<-1>(kind, a, st);
}

}  // namespace qphix_amd
