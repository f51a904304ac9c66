// Field preparation kernels: canonical <-> device layout packers, seeded gauge
// and spinor generation, clover term construction / inversion and precision
// conversion. These run once per configuration; they favour clarity over speed.
#include "fields.hpp"
#include "hop_impl.hpp"

namespace qphix_amd {

// ---------------------------------------------------------------------------
// canonical index helpers (QDP++ checkerboarded lexicographic order)
// ---------------------------------------------------------------------------
struct Lat {
  int L[4];
  int nxh;
  long vh;
};

__device__ inline void cb_coords(const Lat &l, int cb, long s, int c[4]) {
  long r = s;
  const int xh = (int)(r % l.nxh);
  r /= l.nxh;
  c[1] = (int)(r % l.L[1]);
  r /= l.L[1];
  c[2] = (int)(r % l.L[2]);
  c[3] = (int)(r / l.L[2]);
  c[0] = 2 * xh + ((c[1] + c[2] + c[3] + cb) & 1);
}
__device__ inline long cb_site(const Lat &l, const int c[4]) {
  return (((long)c[3] * l.L[2] + c[2]) * l.L[1] + c[1]) * l.nxh + c[0] / 2;
}
__device__ inline int cb_of(const int c[4]) { return (c[0] + c[1] + c[2] + c[3]) & 1; }

// ---------------------------------------------------------------------------
// spinor pack / unpack / generate / convert
// ---------------------------------------------------------------------------
template <int P>
__global__ void k_pack_spinor(void *dst, size_t vs, long vh, const double *src) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= vh) return;
  typename PT<P>::calc v[kSpinorReals];
  for (int k = 0; k < kSpinorReals; k++) v[k] = (typename PT<P>::calc)src[s * kSpinorReals + k];
  store_spinor<P>(dst, vs, s, v);
}

template <int P>
__global__ void k_unpack_spinor(double *dst, size_t vs, long vh, const void *src) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= vh) return;
  typename PT<P>::calc v[kSpinorReals];
  load_spinor<P>(src, vs, s, v);
  for (int k = 0; k < kSpinorReals; k++) dst[s * kSpinorReals + k] = (double)v[k];
}

__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ inline double uniform01(uint64_t seed, uint64_t counter) {
  const uint64_t h = splitmix64(splitmix64(seed) ^ counter);
  return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
__device__ inline void gauss_pair(uint64_t seed, uint64_t counter, double &g0, double &g1) {
  const double u1 = uniform01(seed, 2 * counter), u2 = uniform01(seed, 2 * counter + 1);
  const double r = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586476925286766559 * u2;
  g0 = r * cos(th);
  g1 = r * sin(th);
}
constexpr uint64_t kTagGauge = 0x6761756765000000ULL;
constexpr uint64_t kTagSpinor = 0x7370696e6f720000ULL;

template <int P>
__global__ void k_gen_spinor(void *dst, size_t vs, Lat l, int cb, uint64_t seed) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= l.vh) return;
  int c[4];
  cb_coords(l, cb, s, c);
  const long lf = c[0] + (long)l.L[0] * (c[1] + (long)l.L[1] * (c[2] + (long)l.L[2] * c[3]));
  typename PT<P>::calc v[kSpinorReals];
  for (int k = 0; k < 12; k++) {
    double g0, g1;
    gauss_pair(seed ^ kTagSpinor, (uint64_t)lf * 12 + k, g0, g1);
    v[2 * k] = (typename PT<P>::calc)g0;
    v[2 * k + 1] = (typename PT<P>::calc)g1;
  }
  store_spinor<P>(dst, vs, s, v);
}

template <int PO, int PI>
__global__ void k_convert_spinor(void *out, const void *in, size_t vs, long vh, double scale) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= vh) return;
  typename PT<PI>::calc a[kSpinorReals];
  typename PT<PO>::calc b[kSpinorReals];
  load_spinor<PI>(in, vs, s, a);
  typedef typename PT<PO>::calc TO;
  const TO sc = (TO)scale;
  for (int k = 0; k < kSpinorReals; k++) b[k] = sc * (TO)a[k];
  store_spinor<PO>(out, vs, s, b);
}

// ---------------------------------------------------------------------------
// gauge
// ---------------------------------------------------------------------------
// canonical full gauge (double [V][4][18]) -> one checkerboard array
template <int P, bool C12>
__global__ void k_pack_gauge(void *dst, size_t vs, Lat l, int cb, const double *src) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= l.vh) return;
  constexpr int LR = C12 ? 12 : 18;
  typename PT<P>::calc v[8 * LR];
  int c[4];
  cb_coords(l, cb, s, c);
  const long self = cb * l.vh + s;
  for (int mu = 0; mu < 4; mu++) {
    int d[4] = {c[0], c[1], c[2], c[3]};
    d[mu] = (d[mu] - 1 + l.L[mu]) % l.L[mu];
    const long nb = (long)cb_of(d) * l.vh + cb_site(l, d);
    for (int k = 0; k < LR; k++) {
      v[(2 * mu) * LR + k] = (typename PT<P>::calc)src[(nb * 4 + mu) * 18 + k];
      v[(2 * mu + 1) * LR + k] = (typename PT<P>::calc)src[(self * 4 + mu) * 18 + k];
    }
  }
  store_all<P, words_of<P>(8 * LR)>(dst, vs, s, v);
}

// U_mu(x) = reunit(exp(i eps A)) into a canonical double gauge array
__global__ void k_gen_gauge(double *u, Lat l, uint64_t seed, double eps) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 2 * l.vh) return;
  const int cb = (int)(i / l.vh);
  int c[4];
  cb_coords(l, cb, i - cb * l.vh, c);
  const long lf = c[0] + (long)l.L[0] * (c[1] + (long)l.L[1] * (c[2] + (long)l.L[2] * c[3]));
  for (int mu = 0; mu < 4; mu++) {
    cx<double> G[9], M[9], R[9], T[9], tmp[9];
    for (int k = 0; k < 9; k++) {
      double g0, g1;
      gauss_pair(seed ^ kTagGauge, (uint64_t)(lf * 4 + mu) * 9 + k, g0, g1);
      G[k] = mk(g0, g1);
    }
    double tr = G[0].r + G[4].r + G[8].r;
    double nrm = 0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        cx<double> h = scale(0.5, G[3 * a + b] + conj(G[3 * b + a]));
        if (a == b) h.r -= tr / 3.0;
        M[3 * a + b] = mk(-eps * h.i, eps * h.r);  // i * eps * A
        nrm += M[3 * a + b].r * M[3 * a + b].r + M[3 * a + b].i * M[3 * a + b].i;
      }
    nrm = sqrt(nrm);
    int sq = 0;
    while (nrm > 0.25) {
      nrm *= 0.5;
      sq++;
    }
    const double sc = ldexp(1.0, -sq);
    for (int k = 0; k < 9; k++) {
      M[k] = scale(sc, M[k]);
      R[k] = mk(k % 4 == 0 ? 1.0 : 0.0, 0.0);
      T[k] = R[k];
    }
    for (int n = 1; n <= 18; n++) {
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) tmp[3 * a + b] = T[3 * a] * M[b] + T[3 * a + 1] * M[3 + b] + T[3 * a + 2] * M[6 + b];
      for (int k = 0; k < 9; k++) {
        T[k] = scale(1.0 / n, tmp[k]);
        R[k] = R[k] + T[k];
      }
    }
    for (int q = 0; q < sq; q++) {
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) tmp[3 * a + b] = R[3 * a] * R[b] + R[3 * a + 1] * R[3 + b] + R[3 * a + 2] * R[6 + b];
      for (int k = 0; k < 9; k++) R[k] = tmp[k];
    }
    // reunitarise: row 0 normalised, row 1 orthonormalised, row 2 = conj(r0 x r1)
    double n0 = 0;
    for (int k = 0; k < 3; k++) n0 += R[k].r * R[k].r + R[k].i * R[k].i;
    n0 = 1.0 / sqrt(n0);
    for (int k = 0; k < 3; k++) R[k] = scale(n0, R[k]);
    cx<double> dd = mk(0.0, 0.0);
    for (int k = 0; k < 3; k++) dd = dd + cmulc(R[k], R[3 + k]);
    for (int k = 0; k < 3; k++) R[3 + k] = R[3 + k] - dd * R[k];
    double n1 = 0;
    for (int k = 0; k < 3; k++) n1 += R[3 + k].r * R[3 + k].r + R[3 + k].i * R[3 + k].i;
    n1 = 1.0 / sqrt(n1);
    for (int k = 0; k < 3; k++) R[3 + k] = scale(n1, R[3 + k]);
    for (int k = 0; k < 3; k++) {
      const int a = (k + 1) % 3, b = (k + 2) % 3;
      R[6 + k] = conj(R[a] * R[3 + b] - R[b] * R[3 + a]);
    }
    double *o = u + (i * 4 + mu) * 18;
    for (int k = 0; k < 9; k++) {
      o[2 * k] = R[k].r;
      o[2 * k + 1] = R[k].i;
    }
  }
}

template <int PO, int PI, bool CO, bool CI>
__global__ void k_convert_gauge(void *out, const void *in, size_t vs, long vh) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= vh) return;
  constexpr int LI = CI ? 12 : 18, LO = CO ? 12 : 18;
  typename PT<PI>::calc a[8 * LI];
  typename PT<PO>::calc b[8 * LO];
  load_range<PI, words_of<PI>(8 * LI), 0, words_of<PI>(8 * LI)>(in, vs, s, a);
  for (int l = 0; l < 8; l++) {
    typedef typename PT<PI>::calc TI;
    cx<TI> U[3][3];
    make_link<CI>(a + l * LI, U);
    for (int r = 0; r < (CO ? 2 : 3); r++)
      for (int c = 0; c < 3; c++) {
        b[l * LO + (3 * r + c) * 2] = (typename PT<PO>::calc)U[r][c].r;
        b[l * LO + (3 * r + c) * 2 + 1] = (typename PT<PO>::calc)U[r][c].i;
      }
  }
  store_all<PO, words_of<PO>(8 * LO)>(out, vs, s, b);
}

// ---------------------------------------------------------------------------
// clover
// ---------------------------------------------------------------------------
// Chroma site record (diag[2][6], offd[2][15][2]) <-> per-block (diag[6], offd[15][2])
template <int P>
__global__ void k_pack_clover(void *dst, size_t vs, long vh, const double *src) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= vh) return;
  typename PT<P>::calc v[kCloverReals];
  const double *c = src + s * kCloverReals;
  for (int b = 0; b < 2; b++) {
    for (int k = 0; k < 6; k++) v[36 * b + k] = (typename PT<P>::calc)c[6 * b + k];
    for (int k = 0; k < 30; k++) v[36 * b + 6 + k] = (typename PT<P>::calc)c[12 + 30 * b + k];
  }
  store_all<P, words_of<P>(kCloverReals)>(dst, vs, s, v);
}
template <int P>
__global__ void k_unpack_clover(double *dst, size_t vs, long vh, const void *src) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= vh) return;
  typename PT<P>::calc v[kCloverReals];
  load_range<P, words_of<P>(kCloverReals), 0, words_of<P>(kCloverReals)>(src, vs, s, v);
  double *c = dst + s * kCloverReals;
  for (int b = 0; b < 2; b++) {
    for (int k = 0; k < 6; k++) c[6 * b + k] = (double)v[36 * b + k];
    for (int k = 0; k < 30; k++) c[12 + 30 * b + k] = (double)v[36 * b + 6 + k];
  }
}
template <int P>
__global__ void k_pack_full_clover(void *dst, size_t vs, long vh, const double *src) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= vh) return;
  typename PT<P>::calc v[kFullCloverReals];
  for (int k = 0; k < kFullCloverReals; k++) v[k] = (typename PT<P>::calc)src[s * kFullCloverReals + k];
  store_all<P, words_of<P>(kFullCloverReals)>(dst, vs, s, v);
}
template <int PO, int PI, int N>
__global__ void k_convert_words(void *out, const void *in, size_t vs, long vh) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= vh) return;
  typename PT<PI>::calc a[N];
  typename PT<PO>::calc b[N];
  load_range<PI, words_of<PI>(N), 0, words_of<PI>(N)>(in, vs, s, a);
  for (int k = 0; k < N; k++) b[k] = (typename PT<PO>::calc)a[k];
  store_all<PO, words_of<PO>(N)>(out, vs, s, b);
}

typedef cx<double> zd;

// ---------------------------------------------------------------------------
// clover term: field strength per plane, then per chiral block the Hermitian
// 6x6 matrix and what is asked of it, then the store in the field precision.
// Three launches over chunks of sites keep every work-item's arrays in
// registers (no private memory); the FP64 intermediates live in a chunk-sized
// scratch allocated by the host launcher.
// ---------------------------------------------------------------------------

// U_mu at site c (either checkerboard): link 2mu+1 of the array of c's own
// checkerboard, reading only the words that hold it
template <int P, bool C12, int MU>
__device__ inline void site_link(const void *const u[2], size_t vs, const Lat &l, const int c[4], zd (&U)[3][3]) {
  typedef GaugeLayout<P, C12> GL;
  typedef typename PT<P>::calc T;
  constexpr int pw = PT<P>::per_word;
  constexpr int first = (2 * MU + 1) * GL::link_reals;
  constexpr int w0 = first / pw;
  constexpr int w1 = (first + GL::link_reals + pw - 1) / pw;
  T buf[(w1 - w0) * pw];
  load_range<P, GL::nw, w0, w1 - w0>(u[cb_of(c)], vs, cb_site(l, c), buf);
  cx<T> Ut[3][3];
  make_link<C12>(buf + (first - w0 * pw), Ut);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) U[a][b] = mk((double)Ut[a][b].r, (double)Ut[a][b].i);
}

// X = U (FIRST) or X = X U, with U = U_mu(c) or its adjoint (DAG)
template <int P, bool C12, int MU, bool DAG, bool FIRST>
__device__ inline void leaf_mul(const void *const u[2], size_t vs, const Lat &l, const int c[4], zd (&X)[3][3]) {
  zd U[3][3];
  site_link<P, C12, MU>(u, vs, l, c, U);
  if constexpr (FIRST) {
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) X[a][b] = DAG ? conj(U[b][a]) : U[a][b];
  } else {
    zd Y[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) {
        zd t = mk(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 3; k++) t = t + X[a][k] * (DAG ? conj(U[b][k]) : U[k][b]);
        Y[a][b] = t;
      }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int b = 0; b < 3; b++) X[a][b] = Y[a][b];
  }
}

__device__ inline void shifted(const Lat &l, const int c[4], int mu, int dm, int nu, int dn, int o[4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) o[k] = c[k];
  o[mu] = (o[mu] + dm + l.L[mu]) % l.L[mu];
  o[nu] = (o[nu] + dn + l.L[nu]) % l.L[nu];
}

// coef * F_{MU NU}(x) = coef * (Q - Q^+) / 8, Q the sum of the four clover
// leaves at x (tests/mesfield.cc:24-64), for the chunk sites s0 + j; written
// as 18 doubles, element e at F[e * n + j]
template <int P, bool C12, int MU, int NU>
__global__ void __launch_bounds__(128) k_field_strength(const void *u0, const void *u1, size_t vs, Lat l, int cb,
                                                        long s0, long n, double coef, double *F) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const void *const u[2] = {u0, u1};
  int x[4], xpm[4], xpn[4], xmm[4], xmmpn[4], xmmmn[4], xmn[4], xpmmn[4];
  cb_coords(l, cb, s0 + j, x);
  shifted(l, x, MU, 1, NU, 0, xpm);     // x + mu
  shifted(l, x, MU, 0, NU, 1, xpn);     // x + nu
  shifted(l, x, MU, -1, NU, 0, xmm);    // x - mu
  shifted(l, x, MU, -1, NU, 1, xmmpn);  // x - mu + nu
  shifted(l, x, MU, -1, NU, -1, xmmmn); // x - mu - nu
  shifted(l, x, MU, 0, NU, -1, xmn);    // x - nu
  shifted(l, x, MU, 1, NU, -1, xpmmn);  // x + mu - nu
  zd Q[3][3], X[3][3];
  // U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
  leaf_mul<P, C12, MU, false, true>(u, vs, l, x, X);
  leaf_mul<P, C12, NU, false, false>(u, vs, l, xpm, X);
  leaf_mul<P, C12, MU, true, false>(u, vs, l, xpn, X);
  leaf_mul<P, C12, NU, true, false>(u, vs, l, x, X);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) Q[a][b] = X[a][b];
  // U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu)
  leaf_mul<P, C12, NU, false, true>(u, vs, l, x, X);
  leaf_mul<P, C12, MU, true, false>(u, vs, l, xmmpn, X);
  leaf_mul<P, C12, NU, true, false>(u, vs, l, xmm, X);
  leaf_mul<P, C12, MU, false, false>(u, vs, l, xmm, X);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) Q[a][b] = Q[a][b] + X[a][b];
  // U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu)
  leaf_mul<P, C12, MU, true, true>(u, vs, l, xmm, X);
  leaf_mul<P, C12, NU, true, false>(u, vs, l, xmmmn, X);
  leaf_mul<P, C12, MU, false, false>(u, vs, l, xmmmn, X);
  leaf_mul<P, C12, NU, false, false>(u, vs, l, xmn, X);
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) Q[a][b] = Q[a][b] + X[a][b];
  // U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) U_mu(x)^+
  leaf_mul<P, C12, NU, true, true>(u, vs, l, xmn, X);
  leaf_mul<P, C12, MU, false, false>(u, vs, l, xmn, X);
  leaf_mul<P, C12, NU, false, false>(u, vs, l, xpmmn, X);
  leaf_mul<P, C12, MU, true, false>(u, vs, l, x, X);
  const double f = 0.125 * coef;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const zd v = scale(f, (Q[a][b] + X[a][b]) - conj(Q[b][a] + X[b][a]));
      F[(2 * (3 * a + b)) * n + j] = v.r;
      F[(2 * (3 * a + b) + 1) * n + j] = v.i;
    }
}

// lower-triangle storage of a Hermitian 6x6 matrix: (i, j), i >= j
__host__ __device__ constexpr int tri(int i, int j) { return i * (i + 1) / 2 + j; }
__device__ inline zd hel(const zd (&M)[21], int i, int j) { return i >= j ? M[tri(i, j)] : conj(M[tri(j, i)]); }

__device__ inline zd fplane(const double *F, long n, long j, int k, int a, int b) {
  const double *p = F + (size_t)k * 18 * n;
  return mk(p[(2 * (3 * a + b)) * n + j], p[(2 * (3 * a + b) + 1) * n + j]);
}

// chiral block b of the clover term in Chroma's packing
// (tests/clover_term_qdp_w.h:325-442), index i = 3 * (spin - 2b) + colour:
// with s = +1 (b = 0) or -1 (b = 1), G = i (F01 - s F23), H = (i F12 - F02) - s (i F03 + F13)
//   A_b = [[diag_mass + G, H^+], [H, diag_mass - G]]
// (F already carries the plane's clover coefficient)
__device__ inline void clover_block(const double *F, long n, long j, int blk, double dm, zd (&A)[21]) {
  const double s = blk == 0 ? 1.0 : -1.0;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b <= a; b++) {
      const zd d = fplane(F, n, j, 0, a, b) - scale(s, fplane(F, n, j, 5, a, b));
      const zd g = mk(-d.i, d.r);
      A[tri(a, b)] = a == b ? mk(dm + g.r, 0.0) : g;
      A[tri(a + 3, b + 3)] = a == b ? mk(dm - g.r, 0.0) : scale(-1.0, g);
    }
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      const zd f2 = fplane(F, n, j, 2, a, b), f3 = fplane(F, n, j, 3, a, b);
      const zd em = mk(-f2.i, f2.r) + fplane(F, n, j, 4, a, b);
      const zd bm = mk(-f3.i, f3.r) - fplane(F, n, j, 1, a, b);
      A[tri(a + 3, b)] = bm - scale(s, em);
    }
}

// M <- M^-1 for Hermitian positive definite M: Cholesky M = L L^+ in place,
// Li = L^-1 by forward substitution, M^-1 = Li^+ Li
__device__ inline void hpd_inverse(zd (&M)[21]) {
  double dinv[6];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    double d = M[tri(c, c)].r;
#pragma unroll
    for (int k = 0; k < c; k++) d -= M[tri(c, k)].r * M[tri(c, k)].r + M[tri(c, k)].i * M[tri(c, k)].i;
    d = sqrt(d);
    dinv[c] = 1.0 / d;
    M[tri(c, c)] = mk(d, 0.0);
#pragma unroll
    for (int i = c + 1; i < 6; i++) {
      zd t = M[tri(i, c)];
#pragma unroll
      for (int k = 0; k < c; k++) t = t - M[tri(i, k)] * conj(M[tri(c, k)]);
      M[tri(i, c)] = scale(dinv[c], t);
    }
  }
  zd Li[21];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    Li[tri(c, c)] = mk(dinv[c], 0.0);
#pragma unroll
    for (int i = c + 1; i < 6; i++) {
      zd t = mk(0.0, 0.0);
#pragma unroll
      for (int k = c; k < i; k++) t = t + M[tri(i, k)] * Li[tri(k, c)];
      Li[tri(i, c)] = scale(-dinv[i], t);
    }
  }
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      zd t = mk(0.0, 0.0);
#pragma unroll
      for (int k = i; k < 6; k++) t = t + cmulc(Li[tri(k, i)], Li[tri(k, j)]);
      M[tri(i, j)] = t;
    }
}

// A <- A A + shift
__device__ inline void square_shift(zd (&A)[21], double shift) {
  zd S[21];
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int j = 0; j <= i; j++) {
      zd t = mk(i == j ? shift : 0.0, 0.0);
#pragma unroll
      for (int q = 0; q < 6; q++) t = t + hel(A, i, q) * hel(A, q, j);
      S[tri(i, j)] = t;
    }
#pragma unroll
  for (int k = 0; k < 21; k++) A[k] = S[k];
}

// what: 0 A, 1 A^-1, 5 eps (A^2 + mu^2 - eps^2)^-1 (packed Hermitian, 36
// reals per block); 2 (A + i mu g5)^-1 = (A^2 + mu^2)^-1 (A - i mu g5),
// 3 (A^2 + mu^2 - eps^2)^-1 (A - i mu g5), 4 the same with + i mu g5 (full
// complex blocks, 72 reals per block); results as FP64 into R[e * n + j]
template <int WHAT>
__global__ void __launch_bounds__(64) k_clover_block(const double *F, long n, double dm, double mu, double eps,
                                                     double *R) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int blk = blockIdx.y;
  if (j >= n) return;
  zd A[21];
  clover_block(F, n, j, blk, dm, A);
  if constexpr (WHAT == 0 || WHAT == 1 || WHAT == 5) {
    if constexpr (WHAT == 1) hpd_inverse(A);
    if constexpr (WHAT == 5) {
      square_shift(A, mu * mu - eps * eps);
      hpd_inverse(A);
#pragma unroll
      for (int k = 0; k < 21; k++) A[k] = scale(eps, A[k]);
    }
    double *o = R + (size_t)36 * blk * n;
#pragma unroll
    for (int i = 0; i < 6; i++) o[i * n + j] = A[tri(i, i)].r;
#pragma unroll
    for (int i = 1; i < 6; i++)
#pragma unroll
      for (int k = 0; k < i; k++) {
        const int e = i * (i - 1) / 2 + k;
        o[(6 + 2 * e) * n + j] = A[tri(i, k)].r;
        o[(6 + 2 * e + 1) * n + j] = A[tri(i, k)].i;
      }
  } else {
    square_shift(A, WHAT == 2 ? mu * mu : mu * mu - eps * eps);
    hpd_inverse(A);  // A now holds N = (A^2 + shift)^-1
    zd B[21];
    clover_block(F, n, j, blk, dm, B);  // the block again, for the factor on the right
    const double g5 = blk == 0 ? 1.0 : -1.0;
    const double tw = (WHAT == 4 ? 1.0 : -1.0) * g5 * mu;
    double *o = R + (size_t)72 * blk * n;
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
      for (int k = 0; k < 6; k++) {
        zd t = mk(0.0, 0.0);
#pragma unroll
        for (int q = 0; q < 6; q++) t = t + hel(A, i, q) * hel(B, q, k);
        const zd nik = hel(A, i, k);
        t = t + mk(-tw * nik.i, tw * nik.r);
        o[(12 * i + 2 * k) * n + j] = t.r;
        o[(12 * i + 2 * k + 1) * n + j] = t.i;
      }
  }
}

// FP64 site values R[e * n + j] -> field of precision PO (sites s0 + j), one
// word at a time (as store_all, including the FP16 per-site scale)
template <int PO, int N>
__global__ void k_clover_store(const double *R, long n, long s0, size_t vs, void *out) {
  const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  typedef typename PT<PO>::calc T;
  constexpr int pw = PT<PO>::per_word, nw = N / pw;
  T inv = 1;
  if constexpr (PT<PO>::scaled) {
    float m = 0.f;
    for (int e = 0; e < N; e++) m = fmaxf(m, fabsf((float)R[(size_t)e * n + j]));
    inv = m > 0.f ? 1.0f / m : 0.f;
    float *sc = reinterpret_cast<float *>(reinterpret_cast<char *>(out) + (size_t)nw * vs * 16);
    sc[s0 + j] = m;
  }
  for (int w = 0; w < nw; w++) {
    T v[pw];
#pragma unroll
    for (int k = 0; k < pw; k++) {
      v[k] = (T)R[(size_t)(w * pw + k) * n + j];
      if constexpr (PT<PO>::scaled) v[k] *= inv;
    }
    Word<PO>::template store<false>(out, widx(s0 + j, w, nw), v);
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static inline dim3 grid_for(long n, int b = 256) { return dim3((unsigned)((n + b - 1) / b)); }
static inline Lat make_lat(const GeomDesc &g) {
  Lat l;
  for (int k = 0; k < 4; k++) l.L[k] = g.L[k];
  l.nxh = g.L[0] / 2;
  l.vh = g.vh;
  return l;
}

#define PREC_SWITCH(p, F) \
  switch (p) {            \
    case P64: { constexpr int PP = P64; F; } break; \
    case P32: { constexpr int PP = P32; F; } break; \
    case P16: { constexpr int PP = P16; F; } break; \
    case PB16: { constexpr int PP = PB16; F; } break; \
    default: return hipErrorInvalidValue; \
  }

hipError_t fld_pack_spinor(const GeomDesc &g, void *dst, const double *dsrc, hipStream_t st) {
  PREC_SWITCH(g.prec, (k_pack_spinor<PP><<<grid_for(g.vh), 256, 0, st>>>(dst, g.vs, g.vh, dsrc)));
  return hipGetLastError();
}
hipError_t fld_unpack_spinor(const GeomDesc &g, double *ddst, const void *src, hipStream_t st) {
  PREC_SWITCH(g.prec, (k_unpack_spinor<PP><<<grid_for(g.vh), 256, 0, st>>>(ddst, g.vs, g.vh, src)));
  return hipGetLastError();
}
hipError_t fld_gen_spinor(const GeomDesc &g, void *dst, uint64_t seed, int cb, hipStream_t st) {
  Lat l = make_lat(g);
  PREC_SWITCH(g.prec, (k_gen_spinor<PP><<<grid_for(g.vh), 256, 0, st>>>(dst, g.vs, l, cb, seed)));
  return hipGetLastError();
}

template <int PO>
static hipError_t conv_spinor_in(int pin, void *out, const void *in, size_t vs, long vh, double scale, hipStream_t st) {
  switch (pin) {
    case P64: k_convert_spinor<PO, P64><<<grid_for(vh), 256, 0, st>>>(out, in, vs, vh, scale); break;
    case P32: k_convert_spinor<PO, P32><<<grid_for(vh), 256, 0, st>>>(out, in, vs, vh, scale); break;
    case P16: k_convert_spinor<PO, P16><<<grid_for(vh), 256, 0, st>>>(out, in, vs, vh, scale); break;
    case PB16: k_convert_spinor<PO, PB16><<<grid_for(vh), 256, 0, st>>>(out, in, vs, vh, scale); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t fld_convert_spinor(const GeomDesc &go, void *out, double scale, const GeomDesc &gi, const void *in,
                              hipStream_t st) {
  PREC_SWITCH(go.prec, return conv_spinor_in<PP>(gi.prec, out, in, go.vs, go.vh, scale, st));
  return hipSuccess;
}

hipError_t fld_pack_gauge_cb(const GeomDesc &g, void *dst, int cb, const double *dsrc, hipStream_t st) {
  Lat l = make_lat(g);
  if (g.c12) {
    PREC_SWITCH(g.prec, (k_pack_gauge<PP, true><<<grid_for(g.vh), 256, 0, st>>>(dst, g.vs, l, cb, dsrc)));
  } else {
    switch (g.prec) {
      case P64: k_pack_gauge<P64, false><<<grid_for(g.vh), 256, 0, st>>>(dst, g.vs, l, cb, dsrc); break;
      case P32: k_pack_gauge<P32, false><<<grid_for(g.vh), 256, 0, st>>>(dst, g.vs, l, cb, dsrc); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}
hipError_t fld_gen_gauge_canonical(const GeomDesc &g, double *du, uint64_t seed, double eps, hipStream_t st) {
  Lat l = make_lat(g);
  k_gen_gauge<<<grid_for(2 * g.vh, 128), 128, 0, st>>>(du, l, seed, eps);
  return hipGetLastError();
}

template <int PO, bool CO>
static hipError_t conv_gauge_in(const GeomDesc &gi, void *out, const void *in, size_t vs, long vh, hipStream_t st) {
  const dim3 gr = grid_for(vh, 128);
  if (gi.c12) {
    switch (gi.prec) {
      case P64: k_convert_gauge<PO, P64, CO, true><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
      case P32: k_convert_gauge<PO, P32, CO, true><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
      case P16: k_convert_gauge<PO, P16, CO, true><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
      case PB16: k_convert_gauge<PO, PB16, CO, true><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
      default: return hipErrorInvalidValue;
    }
  } else {
    switch (gi.prec) {
      case P64: k_convert_gauge<PO, P64, CO, false><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
      case P32: k_convert_gauge<PO, P32, CO, false><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
      default: return hipErrorInvalidValue;
    }
  }
  return hipGetLastError();
}
hipError_t fld_convert_gauge(const GeomDesc &go, void *out, const GeomDesc &gi, const void *in, hipStream_t st) {
  if (go.c12) {
    PREC_SWITCH(go.prec, return (conv_gauge_in<PP, true>(gi, out, in, go.vs, go.vh, st)));
  } else {
    switch (go.prec) {
      case P64: return conv_gauge_in<P64, false>(gi, out, in, go.vs, go.vh, st);
      case P32: return conv_gauge_in<P32, false>(gi, out, in, go.vs, go.vh, st);
      default: return hipErrorInvalidValue;
    }
  }
  return hipSuccess;
}

hipError_t fld_pack_clover(const GeomDesc &g, void *dst, const double *dsrc, hipStream_t st) {
  PREC_SWITCH(g.prec, (k_pack_clover<PP><<<grid_for(g.vh), 256, 0, st>>>(dst, g.vs, g.vh, dsrc)));
  return hipGetLastError();
}
hipError_t fld_unpack_clover(const GeomDesc &g, double *ddst, const void *src, hipStream_t st) {
  PREC_SWITCH(g.prec, (k_unpack_clover<PP><<<grid_for(g.vh), 256, 0, st>>>(ddst, g.vs, g.vh, src)));
  return hipGetLastError();
}
hipError_t fld_pack_full_clover(const GeomDesc &g, void *dst, const double *dsrc, hipStream_t st) {
  PREC_SWITCH(g.prec, (k_pack_full_clover<PP><<<grid_for(g.vh, 128), 128, 0, st>>>(dst, g.vs, g.vh, dsrc)));
  return hipGetLastError();
}

template <int PO, int N>
static hipError_t conv_words_in(int pin, void *out, const void *in, size_t vs, long vh, hipStream_t st) {
  const dim3 gr = grid_for(vh, 128);
  switch (pin) {
    case P64: k_convert_words<PO, P64, N><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
    case P32: k_convert_words<PO, P32, N><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
    case P16: k_convert_words<PO, P16, N><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
    case PB16: k_convert_words<PO, PB16, N><<<gr, 128, 0, st>>>(out, in, vs, vh); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
hipError_t fld_convert_clover(const GeomDesc &go, void *out, const GeomDesc &gi, const void *in, int full,
                              hipStream_t st) {
  if (full) {
    PREC_SWITCH(go.prec, return (conv_words_in<PP, kFullCloverReals>(gi.prec, out, in, go.vs, go.vh, st)));
  } else {
    PREC_SWITCH(go.prec, return (conv_words_in<PP, kCloverReals>(gi.prec, out, in, go.vs, go.vh, st)));
  }
  return hipSuccess;
}

template <int P, bool C12, int MU, int NU>
static void launch_fs(const GeomDesc &g, const Lat &l, const void *u0, const void *u1, int cb, long s0, long n,
                      double coef, double *F, hipStream_t st) {
  k_field_strength<P, C12, MU, NU><<<grid_for(n, 128), 128, 0, st>>>(u0, u1, g.vs, l, cb, s0, n, coef, F);
}

// all launches of one clover build, chunk by chunk, into the FP64 scratch
// F (108 doubles per chunk site) and R (up to 144); stream-ordered, no
// synchronisation
template <int P, bool C12>
static hipError_t clover_build_chunks(const GeomDesc &g, const void *u0, const void *u1, double cr, double ct,
                                      double dm, int cb, int what, double mu, double eps, void *out, double *F,
                                      double *R, long chunk, hipStream_t st) {
  Lat l = make_lat(g);
  const bool full = what == 2 || what == 3 || what == 4;
  for (long s0 = 0; s0 < g.vh; s0 += chunk) {
    const long n = (g.vh - s0) < chunk ? (g.vh - s0) : chunk;
    // planes 01 02 03 12 13 23; the planes with t take the temporal coefficient
    launch_fs<P, C12, 0, 1>(g, l, u0, u1, cb, s0, n, cr, F + 0 * 18 * n, st);
    launch_fs<P, C12, 0, 2>(g, l, u0, u1, cb, s0, n, cr, F + 1 * 18 * n, st);
    launch_fs<P, C12, 0, 3>(g, l, u0, u1, cb, s0, n, ct, F + 2 * 18 * n, st);
    launch_fs<P, C12, 1, 2>(g, l, u0, u1, cb, s0, n, cr, F + 3 * 18 * n, st);
    launch_fs<P, C12, 1, 3>(g, l, u0, u1, cb, s0, n, ct, F + 4 * 18 * n, st);
    launch_fs<P, C12, 2, 3>(g, l, u0, u1, cb, s0, n, ct, F + 5 * 18 * n, st);
    const dim3 gb((unsigned)((n + 63) / 64), 2);
    switch (what) {
      case 0: k_clover_block<0><<<gb, 64, 0, st>>>(F, n, dm, mu, eps, R); break;
      case 1: k_clover_block<1><<<gb, 64, 0, st>>>(F, n, dm, mu, eps, R); break;
      case 2: k_clover_block<2><<<gb, 64, 0, st>>>(F, n, dm, mu, eps, R); break;
      case 3: k_clover_block<3><<<gb, 64, 0, st>>>(F, n, dm, mu, eps, R); break;
      case 4: k_clover_block<4><<<gb, 64, 0, st>>>(F, n, dm, mu, eps, R); break;
      default: k_clover_block<5><<<gb, 64, 0, st>>>(F, n, dm, mu, eps, R); break;
    }
    // output in the same precision as the gauge field
    if (full)
      k_clover_store<P, kFullCloverReals><<<grid_for(n), 256, 0, st>>>(R, n, s0, g.vs, out);
    else
      k_clover_store<P, kCloverReals><<<grid_for(n), 256, 0, st>>>(R, n, s0, g.vs, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <int P, bool C12>
static hipError_t clover_build_po(const GeomDesc &g, const void *u0, const void *u1, double cr, double ct, double dm,
                                  int cb, int what, double mu, double eps, void *out, hipStream_t st) {
  // chunks of at most 2^20 sites bound the FP64 scratch (field strength 108,
  // result up to 144 doubles per site) to about 2 GiB
  const long chunk = g.vh < (1l << 20) ? g.vh : (1l << 20);
  double *F = nullptr, *R = nullptr;
  hipError_t e = hipMalloc(&F, sizeof(double) * 108 * (size_t)chunk);
  if (e == hipSuccess) e = hipMalloc(&R, sizeof(double) * kFullCloverReals * (size_t)chunk);
  if (e == hipSuccess) e = clover_build_chunks<P, C12>(g, u0, u1, cr, ct, dm, cb, what, mu, eps, out, F, R, chunk, st);
  // the scratch is freed only after the stream has drained
  const hipError_t es = hipStreamSynchronize(st);
  if (F) (void)hipFree(F);
  if (R) (void)hipFree(R);
  return e != hipSuccess ? e : es;
}
// ---------------------------------------------------------------------------
// allocation-free clover build
//
// `scratch` holds the FP64 field strength (108 doubles per site) and the
// assembled blocks (up to 144 doubles per site) of kNoAllocSites sites; a
// build of any lattice walks it in chunks of that many sites. Each device
// holds its own copy of the array at its own address, looked up once per
// device by fld_noalloc_init; host code passes only that address to kernels.
// Every scratch entry a kernel reads is written earlier in the same call, so
// nothing is cleared. Calls that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kNoAllocSites * (108 + kFullCloverReals)];
static double *g_scratch_dev[kNoAllocMaxDevices] = {};

hipError_t fld_noalloc_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  if (g_scratch_dev[device]) return hipSuccess;
  double *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_dev[device] = p;
  return hipSuccess;
}

template <int P, bool C12>
static hipError_t clover_build_noalloc_po(const GeomDesc &g, const void *u0, const void *u1, double cr, double ct,
                                          double dm, int cb, int what, double mu, double eps, void *out,
                                          double *scr, hipStream_t st) {
  return clover_build_chunks<P, C12>(g, u0, u1, cr, ct, dm, cb, what, mu, eps, out, scr,
                                     scr + 108 * kNoAllocSites, kNoAllocSites, st);
}
hipError_t fld_clover_build_noalloc(const GeomDesc &g, const void *u0, const void *u1, double cr, double ct,
                                    double dm, int cb, int what, double mu, double eps, void *out, int device,
                                    hipStream_t st) {
  if (device < 0 || device >= kNoAllocMaxDevices || !g_scratch_dev[device]) return hipErrorNotInitialized;
  double *scr = g_scratch_dev[device];
  if (g.c12) {
    PREC_SWITCH(g.prec, return (clover_build_noalloc_po<PP, true>(g, u0, u1, cr, ct, dm, cb, what, mu, eps, out,
                                                                   scr, st)));
  } else {
    switch (g.prec) {
      case P64: return clover_build_noalloc_po<P64, false>(g, u0, u1, cr, ct, dm, cb, what, mu, eps, out, scr, st);
      case P32: return clover_build_noalloc_po<P32, false>(g, u0, u1, cr, ct, dm, cb, what, mu, eps, out, scr, st);
      default: return hipErrorInvalidValue;
    }
  }
  return hipSuccess;
}

hipError_t fld_clover_build(const GeomDesc &g, const void *u0, const void *u1, double cr, double ct, double dm,
                            int cb, int what, double mu, double eps, void *out, hipStream_t st) {
  if (g.c12) {
    PREC_SWITCH(g.prec, return (clover_build_po<PP, true>(g, u0, u1, cr, ct, dm, cb, what, mu, eps, out, st)));
  } else {
    switch (g.prec) {
      case P64: return clover_build_po<P64, false>(g, u0, u1, cr, ct, dm, cb, what, mu, eps, out, st);
      case P32: return clover_build_po<P32, false>(g, u0, u1, cr, ct, dm, cb, what, mu, eps, out, st);
      default: return hipErrorInvalidValue;
    }
  }
  return hipSuccess;
}

}  // namespace qphix_amd
