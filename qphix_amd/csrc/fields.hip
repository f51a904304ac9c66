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

// forward link U_mu at site c from the packed arrays of its checkerboard
template <int P, bool C12>
__device__ inline void fwd_link(const void *const u[2], size_t vs, const Lat &l, const int c[4], int mu,
                                zd (&U)[3][3]) {
  typedef GaugeLayout<P, C12> GL;
  typedef typename PT<P>::calc T;
  const int cb = cb_of(c);
  const long s = cb_site(l, c);
  T all[8 * GL::link_reals];
  load_range<P, GL::nw, 0, GL::nw>(u[cb], vs, s, all);
  cx<T> Ut[3][3];
  make_link<C12>(all + (2 * mu + 1) * GL::link_reals, Ut);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) U[a][b] = mk((double)Ut[a][b].r, (double)Ut[a][b].i);
}

__device__ inline void mm3(const zd (&A)[3][3], const zd (&B)[3][3], zd (&C)[3][3]) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) C[a][b] = A[a][0] * B[0][b] + A[a][1] * B[1][b] + A[a][2] * B[2][b];
}
__device__ inline void adj3(zd (&A)[3][3]) {
  zd T[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) T[a][b] = conj(A[b][a]);
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) A[a][b] = T[a][b];
}
__device__ inline void shift2(const Lat &l, const int c[4], int mu, int dm, int nu, int dn, int o[4]) {
  for (int k = 0; k < 4; k++) o[k] = c[k];
  o[mu] = (o[mu] + dm + l.L[mu]) % l.L[mu];
  o[nu] = (o[nu] + dn + l.L[nu]) % l.L[nu];
}

// leaf product A B C D of four links (each optionally daggered) added to Q
template <int P, bool C12>
__device__ inline void add_leaf(const void *const u[2], size_t vs, const Lat &l, const int (&s)[4][4],
                                const int (&m)[4], const bool (&dg)[4], zd (&Q)[3][3]) {
  zd X[3][3], Y[3][3], Z[3][3];
  fwd_link<P, C12>(u, vs, l, s[0], m[0], X);
  if (dg[0]) adj3(X);
  for (int k = 1; k < 4; k++) {
    fwd_link<P, C12>(u, vs, l, s[k], m[k], Y);
    if (dg[k]) adj3(Y);
    mm3(X, Y, Z);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) X[a][b] = Z[a][b];
  }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) Q[a][b] = Q[a][b] + X[a][b];
}

// Hermitian positive definite 6x6 inverse by Cholesky (M = L L^+)
__device__ inline void hpd_inverse6(const zd (&M)[6][6], zd (&Mi)[6][6]) {
  zd Lm[6][6];
  double dinv[6];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) Lm[i][j] = mk(0.0, 0.0);
  for (int j = 0; j < 6; j++) {
    double d = M[j][j].r;
    for (int k = 0; k < j; k++) d -= Lm[j][k].r * Lm[j][k].r + Lm[j][k].i * Lm[j][k].i;
    d = sqrt(d);
    Lm[j][j] = mk(d, 0.0);
    dinv[j] = 1.0 / d;
    for (int i = j + 1; i < 6; i++) {
      zd s = M[i][j];
      for (int k = 0; k < j; k++) s = s - Lm[i][k] * conj(Lm[j][k]);
      Lm[i][j] = scale(dinv[j], s);
    }
  }
  // columns of L^-1 by forward substitution, then Mi = L^-+ L^-1
  zd Li[6][6];
  for (int c = 0; c < 6; c++) {
    for (int i = 0; i < 6; i++) {
      zd s = mk(i == c ? 1.0 : 0.0, 0.0);
      for (int k = 0; k < i; k++) s = s - Lm[i][k] * Li[k][c];
      Li[i][c] = scale(dinv[i], s);
    }
  }
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      zd s = mk(0.0, 0.0);
      for (int k = 0; k < 6; k++) s = s + cmulc(Li[k][i], Li[k][j]);
      Mi[i][j] = s;
    }
}

template <int P, bool C12, int PO>
__global__ void k_clover_build(const void *u0, const void *u1, size_t vs, Lat l, int cb, double cr, double ct,
                               double diag_mass, int what, double mu, double eps, void *out) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= l.vh) return;
  const void *const u[2] = {u0, u1};
  int c[4];
  cb_coords(l, cb, s, c);
  // F_{mu nu} = (Q - Q^+)/8 times the clover coefficient (tests/mesfield.cc:24-64)
  zd F[6][3][3];
  int k = 0;
  for (int m = 0; m < 3; m++)
    for (int n = m + 1; n < 4; n++, k++) {
      zd Q[3][3];
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) Q[a][b] = mk(0.0, 0.0);
      int x0[4], xm[4], xn[4], xmn[4], xpm_mn[4], xpn_mm[4];
      for (int q = 0; q < 4; q++) x0[q] = c[q];
      shift2(l, c, m, 1, n, 0, xm);       // x+mu
      shift2(l, c, m, 0, n, 1, xn);       // x+nu
      shift2(l, c, m, -1, n, 0, xmn);     // x-mu
      shift2(l, c, m, -1, n, 1, xpn_mm);  // x-mu+nu
      shift2(l, c, m, 1, n, -1, xpm_mn);  // x+mu-nu
      int xmm[4], xmnn[4], xnn[4];
      shift2(l, c, m, -1, n, -1, xmnn);   // x-mu-nu
      shift2(l, c, m, 0, n, -1, xnn);     // x-nu
      for (int q = 0; q < 4; q++) xmm[q] = xmn[q];
      {  // U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+
        const int S[4][4] = {{x0[0], x0[1], x0[2], x0[3]}, {xm[0], xm[1], xm[2], xm[3]},
                             {xn[0], xn[1], xn[2], xn[3]}, {x0[0], x0[1], x0[2], x0[3]}};
        const int M[4] = {m, n, m, n};
        const bool D[4] = {false, false, true, true};
        add_leaf<P, C12>(u, vs, l, S, M, D, Q);
      }
      {  // U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu)
        const int S[4][4] = {{x0[0], x0[1], x0[2], x0[3]}, {xpn_mm[0], xpn_mm[1], xpn_mm[2], xpn_mm[3]},
                             {xmm[0], xmm[1], xmm[2], xmm[3]}, {xmm[0], xmm[1], xmm[2], xmm[3]}};
        const int M[4] = {n, m, n, m};
        const bool D[4] = {false, true, true, false};
        add_leaf<P, C12>(u, vs, l, S, M, D, Q);
      }
      {  // U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu)
        const int S[4][4] = {{xmm[0], xmm[1], xmm[2], xmm[3]}, {xmnn[0], xmnn[1], xmnn[2], xmnn[3]},
                             {xmnn[0], xmnn[1], xmnn[2], xmnn[3]}, {xnn[0], xnn[1], xnn[2], xnn[3]}};
        const int M[4] = {m, n, m, n};
        const bool D[4] = {true, true, false, false};
        add_leaf<P, C12>(u, vs, l, S, M, D, Q);
      }
      {  // U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) U_mu(x)^+
        const int S[4][4] = {{xnn[0], xnn[1], xnn[2], xnn[3]}, {xnn[0], xnn[1], xnn[2], xnn[3]},
                             {xpm_mn[0], xpm_mn[1], xpm_mn[2], xpm_mn[3]}, {x0[0], x0[1], x0[2], x0[3]}};
        const int M[4] = {n, m, n, m};
        const bool D[4] = {true, false, false, true};
        add_leaf<P, C12>(u, vs, l, S, M, D, Q);
      }
      const double cf = (n == 3) ? ct : cr;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) F[k][a][b] = scale(0.125 * cf, Q[a][b] - conj(Q[b][a]));
    }
  // Chroma clover blocks (tests/clover_term_qdp_w.h:325-442), dense 6x6 per block
  zd A[2][6][6];
  for (int b = 0; b < 2; b++)
    for (int i = 0; i < 6; i++)
      for (int j = 0; j < 6; j++) A[b][i][j] = mk(i == j ? diag_mass : 0.0, 0.0);
  for (int a = 0; a < 3; a++) {
    const double r0 = F[5][a][a].i - F[0][a][a].i;
    A[0][a][a].r += r0;
    A[0][a + 3][a + 3].r -= r0;
    const double r1 = F[5][a][a].i + F[0][a][a].i;
    A[1][a][a].r -= r1;
    A[1][a + 3][a + 3].r += r1;
  }
  for (int a = 1; a < 3; a++)
    for (int b = 0; b < a; b++) {
      const zd d0 = F[0][a][b] - F[5][a][b];
      const zd c0 = mk(-d0.i, d0.r);
      const zd d1 = F[5][a][b] + F[0][a][b];
      const zd c1 = mk(-d1.i, d1.r);
      A[0][a][b] = c0;
      A[0][a + 3][b + 3] = scale(-1.0, c0);
      A[1][a][b] = c1;
      A[1][a + 3][b + 3] = scale(-1.0, c1);
    }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      const zd Em = mk(-F[2][a][b].i, F[2][a][b].r) + F[4][a][b];
      const zd Bm = mk(-F[3][a][b].i, F[3][a][b].r) - F[1][a][b];
      A[0][a + 3][b] = Bm - Em;
      A[1][a + 3][b] = Em + Bm;
    }
  // complete the Hermitian upper triangles
  for (int b = 0; b < 2; b++)
    for (int i = 0; i < 6; i++)
      for (int j = i + 1; j < 6; j++) A[b][i][j] = conj(A[b][j][i]);

  typedef typename PT<PO>::calc TO;
  if (what == 0 || what == 1 || what == 5) {  // packed Hermitian output
    TO v[kCloverReals];
    for (int b = 0; b < 2; b++) {
      zd R[6][6];
      if (what == 0) {
        for (int i = 0; i < 6; i++)
          for (int j = 0; j < 6; j++) R[i][j] = A[b][i][j];
      } else if (what == 1) {
        hpd_inverse6(A[b], R);
      } else {
        zd A2[6][6], Ni[6][6];
        for (int i = 0; i < 6; i++)
          for (int j = 0; j < 6; j++) {
            zd sum = mk(0.0, 0.0);
            for (int q = 0; q < 6; q++) sum = sum + A[b][i][q] * A[b][q][j];
            A2[i][j] = sum;
          }
        for (int i = 0; i < 6; i++) A2[i][i].r += mu * mu - eps * eps;
        hpd_inverse6(A2, Ni);
        for (int i = 0; i < 6; i++)
          for (int j = 0; j < 6; j++) R[i][j] = scale(eps, Ni[i][j]);
      }
      for (int i = 0; i < 6; i++) v[36 * b + i] = (TO)R[i][i].r;
      for (int i = 1; i < 6; i++)
        for (int j = 0; j < i; j++) {
          const int e = i * (i - 1) / 2 + j;
          v[36 * b + 6 + 2 * e] = (TO)R[i][j].r;
          v[36 * b + 6 + 2 * e + 1] = (TO)R[i][j].i;
        }
    }
    store_all<PO, words_of<PO>(kCloverReals)>(out, vs, s, v);
  } else {  // full blocks
    TO v[kFullCloverReals];
    for (int b = 0; b < 2; b++) {
      const double g5 = b == 0 ? 1.0 : -1.0;
      zd A2[6][6], Ni[6][6], Tm[6][6];
      const double shift = (what == 2) ? mu * mu : mu * mu - eps * eps;
      for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
          zd sum = mk(0.0, 0.0);
          for (int q = 0; q < 6; q++) sum = sum + A[b][i][q] * A[b][q][j];
          A2[i][j] = sum;
        }
      for (int i = 0; i < 6; i++) A2[i][i].r += shift;
      hpd_inverse6(A2, Ni);
      // (A + i mu g5)^-1 = (A^2 + mu^2)^-1 (A - i mu g5); ND up: (..)^-1 (A - i mu g5); ND down: + i mu g5
      const double sgn = (what == 4) ? 1.0 : -1.0;
      for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) Tm[i][j] = A[b][i][j];
      for (int i = 0; i < 6; i++) Tm[i][i].i += sgn * g5 * mu;
      for (int i = 0; i < 6; i++)
        for (int j = 0; j < 6; j++) {
          zd sum = mk(0.0, 0.0);
          for (int q = 0; q < 6; q++) sum = sum + Ni[i][q] * Tm[q][j];
          v[72 * b + 12 * i + 2 * j] = (TO)sum.r;
          v[72 * b + 12 * i + 2 * j + 1] = (TO)sum.i;
        }
    }
    store_all<PO, words_of<PO>(kFullCloverReals)>(out, vs, s, v);
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

template <int P, bool C12>
static hipError_t clover_build_po(const GeomDesc &g, const void *u0, const void *u1, double cr, double ct, double dm,
                                  int cb, int what, double mu, double eps, void *out, hipStream_t st) {
  Lat l = make_lat(g);
  const dim3 gr = grid_for(g.vh, 64);
  // output in the same precision as the gauge field
  k_clover_build<P, C12, P><<<gr, 64, 0, st>>>(u0, u1, g.vs, l, cb, cr, ct, dm, what, mu, eps, out);
  return hipGetLastError();
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
