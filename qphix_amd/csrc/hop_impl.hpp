// Even-odd hopping-term kernels for gfx950.
//
// One work-item per target-checkerboard site. For every dimension mu the
// kernel gathers the two neighbour spinors of the source checkerboard, projects
// them to half spinors with (1 -/+ isign gamma_mu), multiplies by U_mu(x) or
// U_mu(x-mu)^+, and accumulates the reconstructed spinor with the hop
// coefficient. The fused epilogues cover the whole kernel family of the
// reference (codegen dslash_common.cc:1967-2183):
//   KD_WILSON   res = hop                                      (dslash_body.h:338)
//   KA_WILSON   res = alpha chi - beta hop                     (dslash_body.h:361)
//   KD_CLOVER   res = A^-1 hop                                  (clover_dslash_body.h:339)
//   KA_CLOVER   res = A chi + i isign rho g5 chi - beta hop     (clover_dslash_body.h:351)
//   KD_TM       res = muinv (1 - i isign mu' g5) hop            (tm_dslash_body.h:344)
//   KA_TM       res = alpha (1 + i isign mu' g5) chi - beta hop (tm_dslash_body.h:356)
//   KD_FULL     res = F hop, F a full 6x6 block per chirality   (tm_clov_dslash_body.h:339)
// The -beta factor is folded into the hop coefficients by the host.
#pragma once

#include "common.hpp"

namespace qphix_amd {

// KD_DIR: one term of the hop (Dslash::dslashDir, dslash_body.h:1074-1137),
// a.dir = 2 mu takes the forward neighbour term, 2 mu + 1 the backward one
enum HopKind { KD_WILSON = 0, KA_WILSON, KD_CLOVER, KA_CLOVER, KD_TM, KA_TM, KD_FULL, KD_DIR, NUM_HOP_KINDS };


struct HopArgs {
  const void *psi;
  void *out;
  const void *u;     // gauge array of the target checkerboard
  const void *clov;  // clover / full clover of the target checkerboard
  const void *chi;
  size_t vs;
  int vh, nxh, ny, nz, nt, cb;
  double cs;         // spatial hop coefficient
  double ct;         // temporal coefficient away from the boundary
  double ctf, ctb;   // temporal coefficient of the forward hop at t = Nt-1 / backward hop at t = 0
  double tbc_r[4], tbc_i[4];
  int tbc_mask;
  double alpha, mu, muinv, rho;
  // traversal: bz = 0 is the natural site order; otherwise the (y, z) cross
  // section is cut into tiles of by rows (by = 0: all of y) by bz planes and
  // each of `slabs` t-slabs is walked tile by tile, t innermost, so consecutive
  // work-groups of one XCD reuse the source spinor slices they share
  int bz, slabs, by;
  int dir;  // KD_DIR only
  // multi-RHS: nrhs > 1 applies the operator to psis[r] -> outs[r] (chis[r]);
  // the right-hand side is the fastest index of the work-group order, so the
  // work-groups sharing a site range (and its links and clover) run together
  int nrhs;
  // dynamic LDS bytes requested per work-group (unused by the kernel): caps
  // the work-groups resident per CU, so the sites in flight on one XCD, whose
  // neighbour spinors must stay in its L2, fit it (0: no cap)
  int lds_cap;
  // non-null: the launch does nothing when *skip != 0 (a solver iteration
  // enqueued after its device state says the solve is done)
  const int *skip;
  const void *psis[kMaxRhs];
  void *outs[kMaxRhs];
  const void *chis[kMaxRhs];
};

// block -> lattice-tile remap so that each XCD (8 per device, round-robin
// dispatch) walks one contiguous range of sites and keeps its L2 warm
__device__ inline int xcd_block() {
  const int nb = gridDim.x, bid = blockIdx.x;
  const int q = nb >> 3, r = nb & 7, xcd = bid & 7, loc = bid >> 3;
  return (xcd < r) ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

template <typename T> __device__ inline void to_spinor(const T *v, cx<T> (&p)[4][3]) {
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int c = 0; c < 3; c++) p[s][c] = mk(v[(3 * s + c) * 2], v[(3 * s + c) * 2 + 1]);
}
template <typename T> __device__ inline void from_spinor(const cx<T> (&p)[4][3], T *v) {
#pragma unroll
  for (int s = 0; s < 4; s++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      v[(3 * s + c) * 2] = p[s][c].r;
      v[(3 * s + c) * 2 + 1] = p[s][c].i;
    }
}

// 3x3 link from 18 reals, or from 12 (rows 0,1) with row 2 = conj(row0 x row1)
template <bool C12, typename T> __device__ inline void make_link(const T *v, cx<T> (&U)[3][3]) {
#pragma unroll
  for (int r = 0; r < (C12 ? 2 : 3); r++)
#pragma unroll
    for (int c = 0; c < 3; c++) U[r][c] = mk(v[(3 * r + c) * 2], v[(3 * r + c) * 2 + 1]);
  if constexpr (C12) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int a = (c + 1) % 3, b = (c + 2) % 3;
      U[2][c] = conj(U[0][a] * U[1][b] - U[0][b] * U[1][a]);
    }
  }
}

// The link from unscaled stored rows v = U / su (FP16 storage, su the site
// scale): returns U / su. A reconstructed third row is quadratic in the stored
// rows, so it takes one factor su to come out on the same footing.
template <bool C12, typename T> __device__ inline void make_link(const T *v, cx<T> (&U)[3][3], T su) {
  make_link<C12>(v, U);
  if constexpr (C12) {
#pragma unroll
    for (int c = 0; c < 3; c++) U[2][c] = scale(su, U[2][c]);
  }
}

// link products and clover blocks as FMA chains (cmul_f / cfma): FP16
// storage, where the kernels are close to issue-bound; FP32 and FP64 keep
// separate products and sums (FMA chains cut the FP32 VALU count and left the
// memory-bound FP32 single- and multi-RHS times unchanged)
template <int P> constexpr bool fma_of() { return PT<P>::scaled; }

template <int P, bool C12> struct GaugeLayout {
  static constexpr int link_reals = C12 ? 12 : 18;
  static constexpr int nw = words_of<P>(gauge_reals(C12));
  static constexpr bool per_link = (link_reals % PT<P>::per_word) == 0;
  static constexpr int link_words = link_reals / PT<P>::per_word;  // valid if per_link
  static constexpr int pair_words = 2 * link_reals / PT<P>::per_word;
};

// The twisted-boundary phase (tr, ti) is applied unconditionally, (1, 0) when
// the dimension has none: a branch here would split the kernel into basic
// blocks and stop the scheduler from issuing the next hop's loads early.
// apply_hop_h takes the projected half spinor h = (1 -/+ isign g_mu) psi.
// FMA: the link products as FMA chains (cmul_f / cfma; FP16 storage)
template <int MU, int S, bool FWD, bool C12, bool FMA = false, typename T>
__device__ inline void apply_hop_h(const cx<T> (&U)[3][3], const cx<T> (&h)[2][3], T coef, T tr, T ti,
                                   cx<T> (&acc)[4][3]) {
  cx<T> uh[2][3];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int i = 0; i < 3; i++) {
      cx<T> sum;
      if constexpr (FMA && FWD) {
        sum = cmul_f(U[i][0], h[r][0]);
        sum = cfma(U[i][1], h[r][1], sum);
        sum = cfma(U[i][2], h[r][2], sum);
      } else if constexpr (FMA) {
        sum = cmulc_f(U[0][i], h[r][0]);
        sum = cfmac(U[1][i], h[r][1], sum);
        sum = cfmac(U[2][i], h[r][2], sum);
      } else if constexpr (FWD) {
        sum = U[i][0] * h[r][0];
        sum = sum + U[i][1] * h[r][1];
        sum = sum + U[i][2] * h[r][2];
      } else {
        sum = cmulc(U[0][i], h[r][0]);
        sum = sum + cmulc(U[1][i], h[r][1]);
        sum = sum + cmulc(U[2][i], h[r][2]);
      }
      uh[r][i] = sum;
    }
  const cx<T> ph = mk(tr, FWD ? ti : -ti);
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int i = 0; i < 3; i++) uh[r][i] = ph * uh[r][i];
  recon_acc<MU, S>(acc, uh, coef);
}

template <int MU, int S, bool FWD, bool C12, bool FMA = false, typename T>
__device__ inline void apply_hop(const cx<T> (&U)[3][3], const cx<T> (&psi)[4][3], T coef, T tr, T ti,
                                 cx<T> (&acc)[4][3]) {
  cx<T> h[2][3];
  project<MU, S>(psi, h);
  apply_hop_h<MU, S, FWD, C12, FMA>(U, h, coef, tr, ti, acc);
}

template <int P, bool C12, int ISIGN, int MU>
__device__ inline void hop_dim(const HopArgs &a, const void *psi_p, long s, long sb, long sf, typename PT<P>::calc cb_coef,
                               typename PT<P>::calc cf_coef, cx<typename PT<P>::calc> (&acc)[4][3]) {
  typedef typename PT<P>::calc T;
  typedef GaugeLayout<P, C12> GL;
  const bool tbc = (a.tbc_mask >> MU) & 1;
  const T tr = tbc ? (T)a.tbc_r[MU] : (T)1, ti = tbc ? (T)a.tbc_i[MU] : (T)0;
  T sp[kSpinorReals];
  cx<T> psi[4][3], U[3][3];
  if constexpr (PT<P>::scaled) {
    // FP16 storage: links and spinors stay unscaled; the product of the link
    // and neighbour-spinor site scales multiplies the term once, through its
    // boundary phase (every term is linear in both); the link products are
    // FMA chains
    static_assert(!GL::per_link, "FP16 links are read in pairs");
    constexpr int nws = words_of<P>(kSpinorReals);
    T lv[2 * GL::link_reals];
    load_range<P, GL::nw, MU * GL::pair_words, GL::pair_words, true, false>(a.u, a.vs, s, lv);
    const T su = site_scale<P>(a.u, a.vs, s, GL::nw);
    load_spinor<P, false, false>(psi_p, a.vs, sb, sp);
    const T fb = su * site_scale<P>(psi_p, a.vs, sb, nws);
    to_spinor(sp, psi);
    make_link<C12>(lv, U, su);
    apply_hop<MU, ISIGN, false, C12, true>(U, psi, cb_coef, tr * fb, ti * fb, acc);
    load_spinor<P, false, false>(psi_p, a.vs, sf, sp);
    const T ff = su * site_scale<P>(psi_p, a.vs, sf, nws);
    to_spinor(sp, psi);
    make_link<C12>(lv + GL::link_reals, U, su);
    apply_hop<MU, -ISIGN, true, C12, true>(U, psi, cf_coef, tr * ff, ti * ff, acc);
  } else if constexpr (GL::per_link) {
    T lv[GL::link_reals];
    // backward: U_mu(x-mu)^+ (1 + isign gamma_mu) psi(x-mu)
    load_spinor<P>(psi_p, a.vs, sb, sp);
    load_range<P, GL::nw, (2 * MU) * GL::link_words, GL::link_words, true>(a.u, a.vs, s, lv);
    to_spinor(sp, psi);
    make_link<C12>(lv, U);
    apply_hop<MU, ISIGN, false, C12, fma_of<P>()>(U, psi, cb_coef, tr, ti, acc);
    // forward: U_mu(x) (1 - isign gamma_mu) psi(x+mu)
    load_spinor<P>(psi_p, a.vs, sf, sp);
    load_range<P, GL::nw, (2 * MU + 1) * GL::link_words, GL::link_words, true>(a.u, a.vs, s, lv);
    to_spinor(sp, psi);
    make_link<C12>(lv, U);
    apply_hop<MU, -ISIGN, true, C12, fma_of<P>()>(U, psi, cf_coef, tr, ti, acc);
  } else {
    T lv[2 * GL::link_reals];
    load_range<P, GL::nw, MU * GL::pair_words, GL::pair_words, true>(a.u, a.vs, s, lv);
    load_spinor<P>(psi_p, a.vs, sb, sp);
    to_spinor(sp, psi);
    make_link<C12>(lv, U);
    apply_hop<MU, ISIGN, false, C12, fma_of<P>()>(U, psi, cb_coef, tr, ti, acc);
    load_spinor<P>(psi_p, a.vs, sf, sp);
    to_spinor(sp, psi);
    make_link<C12>(lv + GL::link_reals, U);
    apply_hop<MU, -ISIGN, true, C12, fma_of<P>()>(U, psi, cf_coef, tr, ti, acc);
  }
}

// Direction MU of the hop with the neighbour half spinors taken from the
// lanes of the same wavefront where possible. When the wave's 64 target sites
// are consecutive (s0 .. s0 + 63), the source spinor at index n is the
// wave's own-index spinor of lane n - s0, so each lane projects its own-index
// spinor with the direction's projector and the consumers read the 12-real
// half spinor across lanes; lanes whose neighbour lies outside the wave load
// it. The arithmetic is that of hop_dim for FP32 storage (same projection,
// same products); the kernels use this form for FP32 only (hop_shfl).
template <typename T> __device__ inline void shfl_half(const cx<T> (&ho)[2][3], int src, cx<T> (&h)[2][3]) {
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) h[r][c] = mk(__shfl(ho[r][c].r, src, 64), __shfl(ho[r][c].i, src, 64));
}

template <int P, bool C12, int ISIGN, int MU, bool FWD>
__device__ inline void hop_half_shfl(const HopArgs &a, const void *psi_p, const cx<typename PT<P>::calc> (&own)[4][3],
                                     unsigned s, unsigned sn, unsigned s0, bool contig, typename PT<P>::calc coef,
                                     typename PT<P>::calc tr, typename PT<P>::calc ti,
                                     cx<typename PT<P>::calc> (&acc)[4][3]) {
  typedef typename PT<P>::calc T;
  typedef GaugeLayout<P, C12> GL;
  constexpr int S = FWD ? -ISIGN : ISIGN;
  cx<T> ho[2][3], h[2][3];
  project<MU, S>(own, ho);
  const unsigned d = sn - s0;
  const bool in = contig && d < 64u;
  shfl_half(ho, in ? (int)d : 0, h);
  if (!in) {
    T sp[kSpinorReals];
    cx<T> ps[4][3];
    load_spinor<P>(psi_p, a.vs, sn, sp);
    to_spinor(sp, ps);
    project<MU, S>(ps, h);
  }
  cx<T> U[3][3];
  if constexpr (GL::per_link) {
    T lv[GL::link_reals];
    load_range<P, GL::nw, (2 * MU + (FWD ? 1 : 0)) * GL::link_words, GL::link_words, true>(a.u, a.vs, s, lv);
    make_link<C12>(lv, U);
  } else {
    T lv[2 * GL::link_reals];
    load_range<P, GL::nw, MU * GL::pair_words, GL::pair_words, true>(a.u, a.vs, s, lv);
    make_link<C12>(lv + (FWD ? GL::link_reals : 0), U);
  }
  apply_hop_h<MU, S, FWD, C12, fma_of<P>()>(U, h, coef, tr, ti, acc);
}

template <int P, bool C12, int ISIGN, int MU>
__device__ inline void hop_dim_shfl(const HopArgs &a, const void *psi_p, const cx<typename PT<P>::calc> (&own)[4][3],
                                    unsigned s, unsigned sb, unsigned sf, unsigned s0, bool contig,
                                    typename PT<P>::calc coef, cx<typename PT<P>::calc> (&acc)[4][3]) {
  typedef typename PT<P>::calc T;
  const bool tbc = (a.tbc_mask >> MU) & 1;
  const T tr = tbc ? (T)a.tbc_r[MU] : (T)1, ti = tbc ? (T)a.tbc_i[MU] : (T)0;
  hop_half_shfl<P, C12, ISIGN, MU, false>(a, psi_p, own, s, sb, s0, contig, coef, tr, ti, acc);
  hop_half_shfl<P, C12, ISIGN, MU, true>(a, psi_p, own, s, sf, s0, contig, coef, tr, ti, acc);
}

// --- clover blocks --------------------------------------------------------
// packed Hermitian block b (36 reals) of a 72-real clover site (SCALED =
// false: FP16 values unscaled)
template <int P, int B, bool SCALED = true, bool NT = true>
__device__ inline void load_clover_block(const void *c, size_t vs, long s, typename PT<P>::calc *o) {
  constexpr int nw = words_of<P>(kCloverReals);
  if constexpr ((36 % PT<P>::per_word) == 0) {
    load_range<P, nw, B * (36 / PT<P>::per_word), 36 / PT<P>::per_word, NT, SCALED>(c, vs, s, o);
  } else {
    typename PT<P>::calc all[kCloverReals];
    load_range<P, nw, 0, nw, NT, SCALED>(c, vs, s, all);
#pragma unroll
    for (int k = 0; k < 36; k++) o[k] = all[36 * B + k];
  }
}

// out = M in for a packed Hermitian 6x6 block m = {diag[6], offd[15][2]}
// (FMA: the multiply-adds as FMA chains, used for FP16 storage)
template <bool FMA = false, typename T>
__device__ inline void herm6(const T *m, const cx<T> (&in)[6], cx<T> (&out)[6]) {
#pragma unroll
  for (int i = 0; i < 6; i++) {
    cx<T> s = scale(m[i], in[i]);
#pragma unroll
    for (int j = 0; j < 6; j++) {
      if (j < i) {
        const int e = i * (i - 1) / 2 + j;
        if constexpr (FMA) s = cfma(mk(m[6 + 2 * e], m[6 + 2 * e + 1]), in[j], s);
        else s = s + mk(m[6 + 2 * e], m[6 + 2 * e + 1]) * in[j];
      } else if (j > i) {
        const int e = j * (j - 1) / 2 + i;
        if constexpr (FMA) s = cfmac(mk(m[6 + 2 * e], m[6 + 2 * e + 1]), in[j], s);
        else s = s + cmulc(mk(m[6 + 2 * e], m[6 + 2 * e + 1]), in[j]);
      }
    }
    out[i] = s;
  }
}

template <bool FMA = false, typename T>
__device__ inline void full6(const T *m, const cx<T> (&in)[6], cx<T> (&out)[6]) {
#pragma unroll
  for (int i = 0; i < 6; i++) {
    cx<T> s;
    if constexpr (FMA) s = cmul_f(mk(m[12 * i], m[12 * i + 1]), in[0]);
    else s = mk(m[12 * i], m[12 * i + 1]) * in[0];
#pragma unroll
    for (int j = 1; j < 6; j++) {
      if constexpr (FMA) s = cfma(mk(m[12 * i + 2 * j], m[12 * i + 2 * j + 1]), in[j], s);
      else s = s + mk(m[12 * i + 2 * j], m[12 * i + 2 * j + 1]) * in[j];
    }
    out[i] = s;
  }
}

template <int P, int B, typename T>
__device__ inline void clover_block_apply(const void *clov, size_t vs, long s, const cx<T> (&in)[4][3],
                                          cx<T> (&out)[4][3]) {
  T m[36];
  load_clover_block<P, B>(clov, vs, s, m);
  cx<T> v[6], w[6];
#pragma unroll
  for (int i = 0; i < 6; i++) v[i] = in[2 * B + i / 3][i % 3];
  herm6(m, v, w);
#pragma unroll
  for (int i = 0; i < 6; i++) out[2 * B + i / 3][i % 3] = w[i];
}

template <int P, int B, typename T>
__device__ inline void full_block_apply(const void *fcl, size_t vs, long s, const cx<T> (&in)[4][3],
                                        cx<T> (&out)[4][3]) {
  constexpr int nw = words_of<P>(kFullCloverReals);
  constexpr int bw = nw / 2;
  T m[72];
  load_range<P, nw, B * bw, bw, true>(fcl, vs, s, m);
  cx<T> v[6], w[6];
#pragma unroll
  for (int i = 0; i < 6; i++) v[i] = in[2 * B + i / 3][i % 3];
  full6(m, v, w);
#pragma unroll
  for (int i = 0; i < 6; i++) out[2 * B + i / 3][i % 3] = w[i];
}

// out += i * g * gamma_5 in  (g5 = +1 on spins 0,1 and -1 on spins 2,3)
template <typename T> __device__ inline void add_ig5(T g, const cx<T> (&in)[4][3], cx<T> (&out)[4][3]) {
#pragma unroll
  for (int sp = 0; sp < 4; sp++) {
    const T gg = sp < 2 ? g : -g;
#pragma unroll
    for (int c = 0; c < 3; c++) out[sp][c] = out[sp][c] + mk(-gg * in[sp][c].i, gg * in[sp][c].r);
  }
}

// neighbour indices and temporal coefficients of one target site
struct SiteNbrs {
  unsigned s, xb, xf, yb, yf, zb, zf, tb, tf;
  bool t_lo, t_hi;  // t == 0, t == Nt-1
};

// target site of traversal index k and its eight neighbours. 32-bit index
// arithmetic throughout: a parity of one GPU's lattice has fewer than 2^31
// sites, and 64-bit division would cost a branchy library sequence.
__device__ inline SiteNbrs site_nbrs(const HopArgs &a, unsigned k) {
  SiteNbrs n;
  const unsigned nxh = (unsigned)a.nxh, ny = (unsigned)a.ny, nz = (unsigned)a.nz, nt = (unsigned)a.nt;
  unsigned s, xh, y, z, t;
  if (a.bz > 0) {
    // tile slice: all x, by rows (all of y for by = 0), bz planes (x fastest,
    // then y, then z); then t within a slab, then the tiles, then the slabs
    const unsigned by = a.by > 0 ? (unsigned)a.by : ny, bz = (unsigned)a.bz;
    const unsigned nts = nt / (unsigned)a.slabs, nyb = ny / by, nzb = nz / bz;
    const unsigned row = k / nxh;
    xh = k - row * nxh;
    unsigned q = row / by;
    const unsigned yi = row - q * by;
    const unsigned zi = q % bz;
    q /= bz;
    const unsigned ti = q % nts;
    q /= nts;
    const unsigned yb = q % nyb;
    q /= nyb;
    const unsigned zb = q % nzb;
    const unsigned slab = q / nzb;
    y = yb * by + yi;
    z = zb * bz + zi;
    t = slab * nts + ti;
    s = ((t * nz + z) * ny + y) * nxh + xh;
  } else {
    s = k;
    unsigned r = s;
    xh = r % nxh;
    r /= nxh;
    y = r % ny;
    r /= ny;
    z = r % nz;
    t = r / nz;
  }
  const unsigned nx = 2 * nxh;
  const unsigned x = 2 * xh + ((y + z + t + (unsigned)a.cb) & 1);
  const unsigned row = s - xh;
  const unsigned sy = nxh, sz = nxh * ny, st = sz * nz;
  n.s = s;
  n.xb = row + ((x == 0 ? nx - 1 : x - 1) >> 1);
  n.xf = row + ((x == nx - 1 ? 0 : x + 1) >> 1);
  n.yb = y == 0 ? s + (ny - 1) * sy : s - sy;
  n.yf = y == ny - 1 ? s - (ny - 1) * sy : s + sy;
  n.zb = z == 0 ? s + (nz - 1) * sz : s - sz;
  n.zf = z == nz - 1 ? s - (nz - 1) * sz : s + sz;
  n.tb = t == 0 ? s + (nt - 1) * st : s - st;
  n.tf = t == nt - 1 ? s - (nt - 1) * st : s + st;
  n.t_lo = t == 0;
  n.t_hi = t == nt - 1;
  return n;
}

// One direction mu of the hop for R right-hand sides: each link is loaded and
// rebuilt once and multiplies the projected neighbours of all R sources (the
// 3 x (2 R) block contraction of a multi-RHS Dslash).
// FP16 storage: U and the spinors are unscaled, su is the link site scale and
// each term takes su times its spinor's site scale through the phase, as in
// hop_dim (same expressions, same rounding)
template <int P, bool C12, int ISIGN, int MU, bool FWD, int R>
__device__ inline void hop_half_r(const cx<typename PT<P>::calc> (&U)[3][3], const HopArgs &a,
                                  const void *const (&psi)[R], unsigned sn, typename PT<P>::calc coef,
                                  typename PT<P>::calc tr, typename PT<P>::calc ti, typename PT<P>::calc su,
                                  cx<typename PT<P>::calc> (&acc)[R][4][3]) {
  typedef typename PT<P>::calc T;
#pragma unroll
  for (int j = 0; j < R; j++) {
    T sp[kSpinorReals];
    cx<T> ps[4][3];
    if constexpr (PT<P>::scaled) {
      load_spinor<P, false, false>(psi[j], a.vs, sn, sp);
      const T f = su * site_scale<P>(psi[j], a.vs, sn, words_of<P>(kSpinorReals));
      to_spinor(sp, ps);
      apply_hop<MU, FWD ? -ISIGN : ISIGN, FWD, C12, true>(U, ps, coef, tr * f, ti * f, acc[j]);
    } else {
      load_spinor<P>(psi[j], a.vs, sn, sp);
      to_spinor(sp, ps);
      apply_hop<MU, FWD ? -ISIGN : ISIGN, FWD, C12, fma_of<P>()>(U, ps, coef, tr, ti, acc[j]);
    }
  }
}


// NT: links loaded nontemporally (false where several work-groups of one
// site range read them, the register-tiled multi-RHS launch, so that the L2
// keeps them for the others)
template <int P, bool C12, int ISIGN, int MU, int R, bool NT = true>
__device__ inline void hop_dim_r(const HopArgs &a, const void *const (&psi)[R], unsigned s, unsigned sb, unsigned sf,
                                 typename PT<P>::calc cb_coef, typename PT<P>::calc cf_coef,
                                 cx<typename PT<P>::calc> (&acc)[R][4][3]) {
  typedef typename PT<P>::calc T;
  typedef GaugeLayout<P, C12> GL;
  const bool tbc = (a.tbc_mask >> MU) & 1;
  const T tr = tbc ? (T)a.tbc_r[MU] : (T)1, ti = tbc ? (T)a.tbc_i[MU] : (T)0;
  cx<T> U[3][3];
  if constexpr (GL::per_link) {
    static_assert(!PT<P>::scaled, "FP16 links are read in pairs");
    T lv[GL::link_reals];
    // backward: U_mu(x-mu)^+ (1 + isign gamma_mu) psi(x-mu)
    load_range<P, GL::nw, (2 * MU) * GL::link_words, GL::link_words, NT>(a.u, a.vs, s, lv);
    make_link<C12>(lv, U);
    hop_half_r<P, C12, ISIGN, MU, false, R>(U, a, psi, sb, cb_coef, tr, ti, (T)1, acc);
    // forward: U_mu(x) (1 - isign gamma_mu) psi(x+mu)
    load_range<P, GL::nw, (2 * MU + 1) * GL::link_words, GL::link_words, NT>(a.u, a.vs, s, lv);
    make_link<C12>(lv, U);
    hop_half_r<P, C12, ISIGN, MU, true, R>(U, a, psi, sf, cf_coef, tr, ti, (T)1, acc);
  } else {
    T lv[2 * GL::link_reals];
    load_range<P, GL::nw, MU * GL::pair_words, GL::pair_words, NT, !PT<P>::scaled>(a.u, a.vs, s, lv);
    if constexpr (PT<P>::scaled) {
      const T su = site_scale<P>(a.u, a.vs, s, GL::nw);
      make_link<C12>(lv, U, su);
      hop_half_r<P, C12, ISIGN, MU, false, R>(U, a, psi, sb, cb_coef, tr, ti, su, acc);
      make_link<C12>(lv + GL::link_reals, U, su);
      hop_half_r<P, C12, ISIGN, MU, true, R>(U, a, psi, sf, cf_coef, tr, ti, su, acc);
    } else {
      make_link<C12>(lv, U);
      hop_half_r<P, C12, ISIGN, MU, false, R>(U, a, psi, sb, cb_coef, tr, ti, (T)1, acc);
      make_link<C12>(lv + GL::link_reals, U);
      hop_half_r<P, C12, ISIGN, MU, true, R>(U, a, psi, sf, cf_coef, tr, ti, (T)1, acc);
    }
  }
}

// a single hop term (direction MU, forward or backward) for R right-hand sides
template <int P, bool C12, int ISIGN, int MU, bool FWD, int R>
__device__ inline void hop_term_r(const HopArgs &a, const void *const (&psi)[R], unsigned s, unsigned sn,
                                  typename PT<P>::calc coef, cx<typename PT<P>::calc> (&acc)[R][4][3]) {
  typedef typename PT<P>::calc T;
  typedef GaugeLayout<P, C12> GL;
  const bool tbc = (a.tbc_mask >> MU) & 1;
  const T tr = tbc ? (T)a.tbc_r[MU] : (T)1, ti = tbc ? (T)a.tbc_i[MU] : (T)0;
  cx<T> U[3][3];
  T su = (T)1;
  if constexpr (GL::per_link) {
    static_assert(!PT<P>::scaled, "FP16 links are read in pairs");
    T lv[GL::link_reals];
    load_range<P, GL::nw, (2 * MU + (FWD ? 1 : 0)) * GL::link_words, GL::link_words, true>(a.u, a.vs, s, lv);
    make_link<C12>(lv, U);
  } else {
    T lv[2 * GL::link_reals];
    load_range<P, GL::nw, MU * GL::pair_words, GL::pair_words, true, !PT<P>::scaled>(a.u, a.vs, s, lv);
    if constexpr (PT<P>::scaled) {
      su = site_scale<P>(a.u, a.vs, s, GL::nw);
      make_link<C12>(lv + (FWD ? GL::link_reals : 0), U, su);
    } else {
      make_link<C12>(lv + (FWD ? GL::link_reals : 0), U);
    }
  }
  hop_half_r<P, C12, ISIGN, MU, FWD, R>(U, a, psi, sn, coef, tr, ti, su, acc);
}

// packed Hermitian clover block B applied to R spinors (block loaded once).
// FP16 storage: the unscaled block is applied and its site scale multiplies
// the six outputs instead of the 36 stored reals.
template <int P, int B, int R, bool FMA = fma_of<P>(), bool NT = true, typename T>
__device__ inline void clover_block_apply_r(const void *clov, size_t vs, long s, const cx<T> (&in)[R][4][3],
                                            cx<T> (&out)[R][4][3]) {
  constexpr bool raw = PT<P>::scaled;
  T m[36];
  load_clover_block<P, B, !raw, NT>(clov, vs, s, m);
  const T sc = raw ? site_scale<P>(clov, vs, s, words_of<P>(kCloverReals)) : (T)1;
#pragma unroll
  for (int j = 0; j < R; j++) {
    cx<T> v[6], w[6];
#pragma unroll
    for (int i = 0; i < 6; i++) v[i] = in[j][2 * B + i / 3][i % 3];
    herm6<FMA>(m, v, w);
#pragma unroll
    for (int i = 0; i < 6; i++) out[j][2 * B + i / 3][i % 3] = raw ? scale(sc, w[i]) : w[i];
  }
}

template <int P, int B, int R, bool NT = true, typename T>
__device__ inline void full_block_apply_r(const void *fcl, size_t vs, long s, const cx<T> (&in)[R][4][3],
                                          cx<T> (&out)[R][4][3]) {
  constexpr bool raw = PT<P>::scaled;
  constexpr int nw = words_of<P>(kFullCloverReals);
  constexpr int bw = nw / 2;
  T m[72];
  load_range<P, nw, B * bw, bw, NT, !raw>(fcl, vs, s, m);
  const T sc = raw ? site_scale<P>(fcl, vs, s, nw) : (T)1;
#pragma unroll
  for (int j = 0; j < R; j++) {
    cx<T> v[6], w[6];
#pragma unroll
    for (int i = 0; i < 6; i++) v[i] = in[j][2 * B + i / 3][i % 3];
    full6<fma_of<P>()>(m, v, w);
#pragma unroll
    for (int i = 0; i < 6; i++) out[j][2 * B + i / 3][i % 3] = raw ? scale(sc, w[i]) : w[i];
  }
}

// accumulator start of the kernel family: 0, or the A chi part of the
// achimbdpsi forms, for R right-hand sides at site s
template <int P, int ISIGN, int KIND, int R, bool NT = true>
__device__ inline void hop_init_r(const HopArgs &a, unsigned s, const void *const (&chi)[R],
                                  cx<typename PT<P>::calc> (&acc)[R][4][3]) {
  typedef typename PT<P>::calc T;
  if constexpr (KIND == KA_WILSON || KIND == KA_CLOVER || KIND == KA_TM) {
    cx<T> ch[R][4][3];
#pragma unroll
    for (int j = 0; j < R; j++) {
      T cv[kSpinorReals];
      load_spinor<P, true>(chi[j], a.vs, s, cv);
      to_spinor(cv, ch[j]);
    }
    if constexpr (KIND == KA_WILSON) {
      const T al = (T)a.alpha;
#pragma unroll
      for (int j = 0; j < R; j++)
#pragma unroll
        for (int sp = 0; sp < 4; sp++)
#pragma unroll
          for (int c = 0; c < 3; c++) acc[j][sp][c] = scale(al, ch[j][sp][c]);
    } else if constexpr (KIND == KA_CLOVER) {
      clover_block_apply_r<P, 0, R, false, NT>(a.clov, a.vs, s, ch, acc);
      clover_block_apply_r<P, 1, R, false, NT>(a.clov, a.vs, s, ch, acc);
#pragma unroll
      for (int j = 0; j < R; j++) add_ig5((T)(ISIGN * a.rho), ch[j], acc[j]);
    } else {
      const T al = (T)a.alpha;
#pragma unroll
      for (int j = 0; j < R; j++) {
#pragma unroll
        for (int sp = 0; sp < 4; sp++)
#pragma unroll
          for (int c = 0; c < 3; c++) acc[j][sp][c] = ch[j][sp][c];
        add_ig5((T)(ISIGN * a.mu), ch[j], acc[j]);
#pragma unroll
        for (int sp = 0; sp < 4; sp++)
#pragma unroll
          for (int c = 0; c < 3; c++) acc[j][sp][c] = scale(al, acc[j][sp][c]);
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < R; j++)
#pragma unroll
      for (int sp = 0; sp < 4; sp++)
#pragma unroll
        for (int c = 0; c < 3; c++) acc[j][sp][c] = mk((T)0, (T)0);
  }
}

// epilogue of the kernel family (A^-1, twisted-mass factors, full block) and
// the store, for R right-hand sides at site s
template <int P, int ISIGN, int KIND, int R, bool NT = true>
__device__ inline void hop_finish_r(const HopArgs &a, unsigned s, const cx<typename PT<P>::calc> (&acc)[R][4][3],
                                    void *const (&out)[R]) {
  typedef typename PT<P>::calc T;
  if constexpr (KIND == KD_CLOVER || KIND == KD_FULL) {
    cx<T> o[R][4][3];
    if constexpr (KIND == KD_CLOVER) {
      clover_block_apply_r<P, 0, R, fma_of<P>(), NT>(a.clov, a.vs, s, acc, o);
      clover_block_apply_r<P, 1, R, fma_of<P>(), NT>(a.clov, a.vs, s, acc, o);
    } else {
      full_block_apply_r<P, 0, R, NT>(a.clov, a.vs, s, acc, o);
      full_block_apply_r<P, 1, R, NT>(a.clov, a.vs, s, acc, o);
    }
#pragma unroll
    for (int j = 0; j < R; j++) {
      T ov[kSpinorReals];
      from_spinor(o[j], ov);
      store_spinor<P, true>(out[j], a.vs, s, ov);
    }
  } else {
#pragma unroll
    for (int j = 0; j < R; j++) {
      T ov[kSpinorReals];
      if constexpr (KIND == KD_TM) {
        cx<T> o[4][3];
#pragma unroll
        for (int sp = 0; sp < 4; sp++)
#pragma unroll
          for (int c = 0; c < 3; c++) o[sp][c] = acc[j][sp][c];
        add_ig5((T)(-ISIGN * a.mu), acc[j], o);
        const T mi = (T)a.muinv;
#pragma unroll
        for (int sp = 0; sp < 4; sp++)
#pragma unroll
          for (int c = 0; c < 3; c++) o[sp][c] = scale(mi, o[sp][c]);
        from_spinor(o, ov);
      } else {
        from_spinor(acc[j], ov);
      }
      store_spinor<P, true>(out[j], a.vs, s, ov);
    }
  }
}

// the hop for traversal index k (< a.vh) of R right-hand sides psi[j] -> out[j]
// Kernels that take their x / y neighbours across lanes (hop_dim_shfl). The
// exchange keeps the own-index spinor live through two dimensions, which costs
// registers, and the 24 lane exchanges per site cost the same whatever the
// storage: measured at 48^3x96 it makes the FP32 forms ~5% faster, the FP16
// forms (52-byte spinors, cheap to load) ~8% slower, and the FP64 forms would
// drop to one or two waves per SIMD. FP32 only.
constexpr bool hop_shfl(int p, int kind) { return p == P32 && kind != KD_DIR; }

// wave_full: all 64 lanes of the wavefront run this site loop (the x / y
// neighbours may then come from other lanes, hop_dim_shfl)
template <int P, bool C12, int ISIGN, int KIND, int R, bool NT = true>
__device__ inline void hop_body_r(const HopArgs &a, unsigned k, const void *const (&psi)[R], void *const (&out)[R],
                                  const void *const (&chi)[R], bool wave_full = false) {
  typedef typename PT<P>::calc T;
  const SiteNbrs n = site_nbrs(a, k);
  const unsigned s = n.s;
  const T cs = (T)a.cs;
  const T ctb = (T)(n.t_lo ? a.ctb : a.ct);
  const T ctf = (T)(n.t_hi ? a.ctf : a.ct);

  cx<T> acc[R][4][3];
  hop_init_r<P, ISIGN, KIND, R, NT>(a, s, chi, acc);
  if constexpr (KIND == KD_DIR) {
    switch (a.dir) {
      case 0: hop_term_r<P, C12, ISIGN, 0, true, R>(a, psi, s, n.xf, cs, acc); break;
      case 1: hop_term_r<P, C12, ISIGN, 0, false, R>(a, psi, s, n.xb, cs, acc); break;
      case 2: hop_term_r<P, C12, ISIGN, 1, true, R>(a, psi, s, n.yf, cs, acc); break;
      case 3: hop_term_r<P, C12, ISIGN, 1, false, R>(a, psi, s, n.yb, cs, acc); break;
      case 4: hop_term_r<P, C12, ISIGN, 2, true, R>(a, psi, s, n.zf, cs, acc); break;
      case 5: hop_term_r<P, C12, ISIGN, 2, false, R>(a, psi, s, n.zb, cs, acc); break;
      case 6: hop_term_r<P, C12, ISIGN, 3, true, R>(a, psi, s, n.tf, ctf, acc); break;
      default: hop_term_r<P, C12, ISIGN, 3, false, R>(a, psi, s, n.tb, ctb, acc); break;
    }
  } else if constexpr (R == 1) {
    if (hop_shfl(P, KIND) && wave_full) {
      T ov[kSpinorReals];
      cx<T> own[4][3];
      load_spinor<P>(psi[0], a.vs, s, ov);
      to_spinor(ov, own);
      const unsigned s0 = __shfl(s, 0, 64);
      const bool contig = __shfl(s, 63, 64) == s0 + 63u;
      hop_dim_shfl<P, C12, ISIGN, 0>(a, psi[0], own, s, n.xb, n.xf, s0, contig, cs, acc[0]);
      hop_dim_shfl<P, C12, ISIGN, 1>(a, psi[0], own, s, n.yb, n.yf, s0, contig, cs, acc[0]);
    } else {
      hop_dim<P, C12, ISIGN, 0>(a, psi[0], s, n.xb, n.xf, cs, cs, acc[0]);
      hop_dim<P, C12, ISIGN, 1>(a, psi[0], s, n.yb, n.yf, cs, cs, acc[0]);
    }
    hop_dim<P, C12, ISIGN, 2>(a, psi[0], s, n.zb, n.zf, cs, cs, acc[0]);
    hop_dim<P, C12, ISIGN, 3>(a, psi[0], s, n.tb, n.tf, ctb, ctf, acc[0]);
  } else {
    hop_dim_r<P, C12, ISIGN, 0, R, NT>(a, psi, s, n.xb, n.xf, cs, cs, acc);
    hop_dim_r<P, C12, ISIGN, 1, R, NT>(a, psi, s, n.yb, n.yf, cs, cs, acc);
    hop_dim_r<P, C12, ISIGN, 2, R, NT>(a, psi, s, n.zb, n.zf, cs, cs, acc);
    hop_dim_r<P, C12, ISIGN, 3, R, NT>(a, psi, s, n.tb, n.tf, ctb, ctf, acc);
  }
  hop_finish_r<P, ISIGN, KIND, R, NT>(a, s, acc, out);
}

// the hop for traversal index k (< a.vh) of one right-hand side
template <int P, bool C12, int ISIGN, int KIND>
__device__ inline void hop_body(const HopArgs &a, long k, const void *psi_p, void *out_p, const void *chi_p,
                                bool wave_full = false) {
  const void *const psi[1] = {psi_p};
  void *const out[1] = {out_p};
  const void *const chi[1] = {chi_p};
  hop_body_r<P, C12, ISIGN, KIND, 1>(a, (unsigned)k, psi, out, chi, wave_full);
}

// minimum waves per SIMD the register allocation must allow for the hop
// kernels. None is imposed: with the 4096-site field blocks the compiler's own
// allocation gives every FP16 / BF16 form four waves (at most 126 VGPRs) and
// the FP32 forms three or four, while an explicit floor of four made its scheduler
// spill the FP16 A chi - b D psi forms (9 VGPRs at isign +1). The issue-bound
// FP16 kernels need those four waves to hide their load latency; a change
// that takes them past 128 VGPRs costs a wave.
#ifndef QPHIX_HOP_WAVES16
#define QPHIX_HOP_WAVES16 1
#endif
constexpr int hop_min_waves(int p) { return (p == P16 || p == PB16) ? QPHIX_HOP_WAVES16 : 1; }

template <int P, bool C12, int ISIGN, int KIND>
__global__ void __launch_bounds__(256, hop_min_waves(P)) hop_kernel(const HopArgs a) {
  if (a.skip && *a.skip) return;
  int lb = xcd_block();
  const void *psi_p = a.psi, *chi_p = a.chi;
  void *out_p = a.out;
  if (a.nrhs > 1) {
    const int r = lb % a.nrhs;
    lb /= a.nrhs;
    // The per-RHS pointers are read in place from the kernel-argument segment,
    // where this struct (the kernel's only parameter) starts at offset 0:
    // indexing the by-value parameter with a run-time index would make the
    // compiler copy the whole struct to private memory for every lane.
    const HopArgs *ka = (const HopArgs *)__builtin_amdgcn_kernarg_segment_ptr();
    psi_p = ka->psis[r];
    out_p = ka->outs[r];
    chi_p = ka->chis[r];
  }
  const long k = (long)lb * blockDim.x + threadIdx.x;
  // the cross-lane neighbour exchange needs every lane of the wave (one
  // RHS per launch: the per-RHS pointers are uniform over the wave)
  const bool wave_full = __ballot(k < a.vh) == ~0ull;
  if (k >= a.vh) return;
  hop_body<P, C12, ISIGN, KIND>(a, k, psi_p, out_p, chi_p, wave_full);
}

// Multi-RHS hop with a register tile of R right-hand sides per work-item:
// a.nrhs (a multiple of R) sources in a.psis; the RHS group is the fastest
// index of the work-group order, so the groups sharing a site range run
// together and also share links and clover through the L2.
template <int P, bool C12, int ISIGN, int KIND, int R>
__global__ void __launch_bounds__(256) hop_tile_kernel(const HopArgs a) {
  if (a.skip && *a.skip) return;
  int lb = xcd_block();
  const int ng = a.nrhs / R;
  const int g = lb % ng;
  lb /= ng;
  const unsigned k = (unsigned)lb * blockDim.x + threadIdx.x;
  if (k >= (unsigned)a.vh) return;
  // per-RHS pointers read in place from the kernel-argument segment (see hop_kernel)
  const HopArgs *ka = (const HopArgs *)__builtin_amdgcn_kernarg_segment_ptr();
  const void *psi[R];
  void *out[R];
  const void *chi[R];
#pragma unroll
  for (int j = 0; j < R; j++) {
    psi[j] = ka->psis[g * R + j];
    out[j] = ka->outs[g * R + j];
    chi[j] = ka->chis[g * R + j];
  }
  hop_body_r<P, C12, ISIGN, KIND, R, false>(a, k, psi, out, chi);
}

// Batched hop: one launch covers several systems (possibly on different
// lattices). Segment g owns the work-groups [blk_off[g], blk_off[g+1]) of a
// flat index space that a fixed grid walks with a block stride (persistent
// form), so the launch shape does not depend on the batch and can be captured
// once in a graph. Segments whose system has converged (active[sys] == 0) are
// skipped.
constexpr int kBatchTile = 2;  // systems per segment of the tiled batched hop
struct HopSeg {
  HopArgs a;
  int sys;   // first system of the segment
  int nsys;  // tiled form: systems sys .. sys + nsys - 1 share the operator (a.psis / outs / chis)
};
// Occupancy ceiling of the batched hops. The FP16 A chi - beta D psi forms
// (146-169 VGPRs) sit near the three-waves-per-SIMD line, where the scheduler
// trades load overlap for registers: their flat loads then end up waited on
// one by one (the register-tiled form ran 63% slower). Two waves leave them
// the registers to keep the loads in flight; the other kinds keep theirs
// (the full-block A^-1 tile runs 15% faster at three waves).
constexpr int batch_max_waves(int p, int kind) { return (p == P16 || p == PB16) && kind == KA_CLOVER ? 2 : 8; }

template <int P, bool C12, int ISIGN, int KIND>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, batch_max_waves(P, KIND))))
hop_batch_kernel(const HopSeg *segs, const int *blk_off, int nseg,
                                                         const int *active) {
  const int total = blk_off[nseg];
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    int g = 0;
    while (g + 1 < nseg && blk_off[g + 1] <= b) g++;
    const HopSeg &sg = segs[g];
    if (active && !active[sg.sys]) continue;
    const long k = (long)(b - blk_off[g]) * blockDim.x + threadIdx.x;
    if (k < sg.a.vh) hop_body<P, C12, ISIGN, KIND>(sg.a, k, sg.a.psi, sg.a.out, sg.a.chi);
  }
}

// Batched hop with a register tile: each segment holds R systems that share
// one operator (the same lattice, links and clover, R right-hand sides), so
// every work-item applies the links and clover of its site to all R sources
// (hop_body_r). A segment is skipped once all its systems have converged.
template <int P, bool C12, int ISIGN, int KIND, int R>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, batch_max_waves(P, KIND))))
hop_batch_tile_kernel(const HopSeg *segs, const int *blk_off, int nseg,
                                                              const int *active) {
  const int total = blk_off[nseg];
  for (int b = blockIdx.x; b < total; b += gridDim.x) {
    int g = 0;
    while (g + 1 < nseg && blk_off[g + 1] <= b) g++;
    const HopSeg &sg = segs[g];
    if (active) {
      bool any = false;
#pragma unroll
      for (int j = 0; j < R; j++) any = any || active[sg.sys + j] != 0;
      if (!any) continue;
    }
    const unsigned k = (unsigned)(b - blk_off[g]) * blockDim.x + threadIdx.x;
    if (k < (unsigned)sg.a.vh) {
      const void *psi[R], *chi[R];
      void *out[R];
#pragma unroll
      for (int j = 0; j < R; j++) {
        psi[j] = sg.a.psis[j];
        out[j] = sg.a.outs[j];
        chi[j] = sg.a.chis[j];
      }
      hop_body_r<P, C12, ISIGN, KIND, R>(sg.a, k, psi, out, chi);
    }
  }
}

template <int P>
hipError_t launch_hop_prec(int kind, bool c12, int isign, const HopArgs &a, int block, hipStream_t st);
template <int P>
hipError_t launch_hop_tile_prec(int kind, bool c12, int isign, int rt, const HopArgs &a, int block, hipStream_t st);
template <int P>
hipError_t launch_hop_batch_tile_prec(int kind, bool c12, int isign, const HopSeg *segs, const int *off, int nseg,
                                      const int *active, int grid, hipStream_t st);
template <int P>
hipError_t launch_hop_batch_prec(int kind, bool c12, int isign, const HopSeg *segs, const int *off, int nseg,
                                 const int *active, int grid, hipStream_t st);
// matrix-core multi-RHS launch (FP32 storage, a.nrhs a multiple of 8)
hipError_t launch_hop_mfma_f32(int kind, bool c12, int isign, const HopArgs &a, hipStream_t st);
// matrix-core multi-RHS launch (FP16 storage, two-row links, a.nrhs a multiple of 8)
hipError_t launch_hop_mfma_f16(int kind, bool c12, int isign, const HopArgs &a, hipStream_t st);

}  // namespace qphix_amd
