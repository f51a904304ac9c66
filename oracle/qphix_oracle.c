/*
 * qphix_oracle.c -- FP64 CPU restatement of the QPhiX Dslash / clover /
 * twisted-mass / BLAS / solver algorithms.  TEST INFRASTRUCTURE ONLY: the
 * reference results of tests/, of the smoke check and of bench.py's CPU
 * baseline; never part of the product path.  Layout and
 * conventions are described in qphix_oracle.h.
 *
 * The hopping term is written with explicit dense 4x4 gamma matrices (no
 * half-spinor projection), which is a different formulation from the GPU
 * kernels (projector / reconstruct); both restate tests/dslashm_w.cc:45-144.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "qphix_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double re, im; } cpx;
static inline cpx cmul(cpx a, cpx b) { cpx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }
static inline cpx cmulc(cpx a, cpx b) { /* conj(a)*b */ cpx r = {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re}; return r; }
static inline cpx cadd(cpx a, cpx b) { cpx r = {a.re + b.re, a.im + b.im}; return r; }
static inline cpx csub(cpx a, cpx b) { cpx r = {a.re - b.re, a.im - b.im}; return r; }
static inline cpx cscale(double s, cpx a) { cpx r = {s * a.re, s * a.im}; return r; }
static inline cpx cconj(cpx a) { cpx r = {a.re, -a.im}; return r; }
static inline cpx ctimesI(cpx a) { cpx r = {-a.im, a.re}; return r; }
static inline cpx cmk(double re, double im) { cpx r = {re, im}; return r; }

/* ------------------------------------------------------------------------ */
/* geometry                                                                  */
/* ------------------------------------------------------------------------ */
long orc_vh(const int L[4]) { return (long)L[0] * L[1] * L[2] * L[3] / 2; }

/* QDP++ checkerboarded lexicographic index (qdp_packer_parscalar.h:75). */
long orc_site_index(const int L[4], int x, int y, int z, int t)
{
  int cb = (x + y + z + t) & 1;
  long nxh = L[0] / 2;
  return cb * orc_vh(L) + (((long)t * L[2] + z) * L[1] + y) * nxh + x / 2;
}

void orc_site_coords(const int L[4], long i, int c[4])
{
  long vh = orc_vh(L);
  int cb = (int)(i / vh);
  long r = i - cb * vh;
  int nxh = L[0] / 2;
  int xh = (int)(r % nxh); r /= nxh;
  c[1] = (int)(r % L[1]); r /= L[1];
  c[2] = (int)(r % L[2]); r /= L[2];
  c[3] = (int)r;
  c[0] = 2 * xh + ((c[1] + c[2] + c[3] + cb) & 1);
}

static long lex_full(const int L[4], const int c[4])
{
  return c[0] + (long)L[0] * (c[1] + (long)L[1] * (c[2] + (long)L[2] * c[3]));
}

static long neighbour(const int L[4], const int c[4], int mu, int dir)
{
  int d[4] = {c[0], c[1], c[2], c[3]};
  d[mu] = (d[mu] + dir + L[mu]) % L[mu];
  return orc_site_index(L, d[0], d[1], d[2], d[3]);
}

/* ------------------------------------------------------------------------ */
/* counter-based RNG: SplitMix64 finaliser + Box-Muller                      */
/* ------------------------------------------------------------------------ */
static uint64_t splitmix64(uint64_t z)
{
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

/* uniform in (0,1): 53 random bits, offset by half an ulp */
double orc_uniform(uint64_t seed, uint64_t counter)
{
  uint64_t h = splitmix64(splitmix64(seed) ^ counter);
  return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

void orc_gauss_pair(uint64_t seed, uint64_t counter, double g[2])
{
  double u1 = orc_uniform(seed, 2 * counter);
  double u2 = orc_uniform(seed, 2 * counter + 1);
  double r = sqrt(-2.0 * log(u1));
  double th = 6.283185307179586476925286766559 * u2;
  g[0] = r * cos(th);
  g[1] = r * sin(th);
}

/* domain tags keep gauge and spinor streams apart for equal seeds */
#define ORC_TAG_GAUGE 0x6761756765000000ULL
#define ORC_TAG_SPINOR 0x7370696e6f720000ULL

static void mat_mul3(const cpx *a, const cpx *b, cpx *c)
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cpx s = {0, 0};
      for (int k = 0; k < 3; k++) s = cadd(s, cmul(a[3 * i + k], b[3 * k + j]));
      c[3 * i + j] = s;
    }
}

/* exp(i*eps*A) for Hermitian A by scaling-and-squaring of a Taylor series */
void orc_su3_exp_i(const double A[18], double eps, double U[18])
{
  cpx M[9];
  double nrm = 0;
  for (int k = 0; k < 9; k++) {
    M[k] = cmk(-eps * A[2 * k + 1], eps * A[2 * k]); /* i*eps*A */
    nrm += M[k].re * M[k].re + M[k].im * M[k].im;
  }
  nrm = sqrt(nrm);
  int sq = 0;
  while (nrm > 0.25) { nrm *= 0.5; sq++; }
  double sc = ldexp(1.0, -sq);
  for (int k = 0; k < 9; k++) M[k] = cscale(sc, M[k]);
  cpx R[9], T[9], tmp[9];
  for (int k = 0; k < 9; k++) { R[k] = cmk(k % 4 == 0 ? 1 : 0, 0); T[k] = R[k]; }
  for (int n = 1; n <= 18; n++) {
    mat_mul3(T, M, tmp);
    for (int k = 0; k < 9; k++) { T[k] = cscale(1.0 / n, tmp[k]); R[k] = cadd(R[k], T[k]); }
  }
  for (int s = 0; s < sq; s++) {
    mat_mul3(R, R, tmp);
    memcpy(R, tmp, sizeof(R));
  }
  for (int k = 0; k < 9; k++) { U[2 * k] = R[k].re; U[2 * k + 1] = R[k].im; }
}

/* row 0 normalised, row 1 orthonormalised, row 2 = conj(row0 x row1) */
void orc_reunit(double U[18])
{
  cpx r[9];
  for (int k = 0; k < 9; k++) r[k] = cmk(U[2 * k], U[2 * k + 1]);
  double n0 = 0;
  for (int c = 0; c < 3; c++) n0 += r[c].re * r[c].re + r[c].im * r[c].im;
  n0 = 1.0 / sqrt(n0);
  for (int c = 0; c < 3; c++) r[c] = cscale(n0, r[c]);
  cpx d = {0, 0};
  for (int c = 0; c < 3; c++) d = cadd(d, cmulc(r[c], r[3 + c]));
  for (int c = 0; c < 3; c++) r[3 + c] = csub(r[3 + c], cmul(d, r[c]));
  double n1 = 0;
  for (int c = 0; c < 3; c++) n1 += r[3 + c].re * r[3 + c].re + r[3 + c].im * r[3 + c].im;
  n1 = 1.0 / sqrt(n1);
  for (int c = 0; c < 3; c++) r[3 + c] = cscale(n1, r[3 + c]);
  for (int c = 0; c < 3; c++) {
    int a = (c + 1) % 3, b = (c + 2) % 3;
    r[6 + c] = cconj(csub(cmul(r[a], r[3 + b]), cmul(r[b], r[3 + a])));
  }
  for (int k = 0; k < 9; k++) { U[2 * k] = r[k].re; U[2 * k + 1] = r[k].im; }
}

void orc_gen_gauge(const int L[4], uint64_t seed, double eps, double *u)
{
  long V = 2 * orc_vh(L);
  uint64_t s = seed ^ ORC_TAG_GAUGE;
  for (long i = 0; i < V; i++) {
    int c[4];
    orc_site_coords(L, i, c);
    long lf = lex_full(L, c);
    for (int mu = 0; mu < 4; mu++) {
      cpx G[9], A[9];
      for (int k = 0; k < 9; k++) {
        double g[2];
        orc_gauss_pair(s, (uint64_t)(lf * 4 + mu) * 9 + k, g);
        G[k] = cmk(g[0], g[1]);
      }
      /* A = (G + G^+)/2 - tr(G + G^+)/6 */
      double tr = 0;
      for (int a = 0; a < 3; a++) tr += G[4 * a].re;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          cpx h = cscale(0.5, cadd(G[3 * a + b], cconj(G[3 * b + a])));
          if (a == b) h.re -= tr / 3.0;
          A[3 * a + b] = h;
        }
      double Ad[18], Ud[18];
      for (int k = 0; k < 9; k++) { Ad[2 * k] = A[k].re; Ad[2 * k + 1] = A[k].im; }
      orc_su3_exp_i(Ad, eps, Ud);
      orc_reunit(Ud);
      memcpy(u + (i * 4 + mu) * 18, Ud, sizeof(Ud));
    }
  }
}

void orc_gen_spinor(const int L[4], uint64_t seed, int cb, double *psi)
{
  long vh = orc_vh(L);
  uint64_t s = seed ^ ORC_TAG_SPINOR;
  for (long j = 0; j < vh; j++) {
    int c[4];
    orc_site_coords(L, cb * vh + j, c);
    long lf = lex_full(L, c);
    for (int k = 0; k < 12; k++) {
      double g[2];
      orc_gauss_pair(s, (uint64_t)lf * 12 + k, g);
      psi[j * 24 + 2 * k] = g[0];
      psi[j * 24 + 2 * k + 1] = g[1];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* gamma matrices (DeGrand-Rossi, the basis of codegen/dslash_common.cc:848)  */
/* ------------------------------------------------------------------------ */
static const double GAMMA[4][4][4][2] = {
    /* gamma_0 (x) */
    {{{0, 0}, {0, 0}, {0, 0}, {0, 1}},
     {{0, 0}, {0, 0}, {0, 1}, {0, 0}},
     {{0, 0}, {0, -1}, {0, 0}, {0, 0}},
     {{0, -1}, {0, 0}, {0, 0}, {0, 0}}},
    /* gamma_1 (y) */
    {{{0, 0}, {0, 0}, {0, 0}, {-1, 0}},
     {{0, 0}, {0, 0}, {1, 0}, {0, 0}},
     {{0, 0}, {1, 0}, {0, 0}, {0, 0}},
     {{-1, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* gamma_2 (z) */
    {{{0, 0}, {0, 0}, {0, 1}, {0, 0}},
     {{0, 0}, {0, 0}, {0, 0}, {0, -1}},
     {{0, -1}, {0, 0}, {0, 0}, {0, 0}},
     {{0, 0}, {0, 1}, {0, 0}, {0, 0}}},
    /* gamma_3 (t) */
    {{{0, 0}, {0, 0}, {1, 0}, {0, 0}},
     {{0, 0}, {0, 0}, {0, 0}, {1, 0}},
     {{1, 0}, {0, 0}, {0, 0}, {0, 0}},
     {{0, 0}, {1, 0}, {0, 0}, {0, 0}}}};

void orc_gamma(int mu, double g[32]) { memcpy(g, GAMMA[mu], 32 * sizeof(double)); }

/* v = (1 + s*gamma_mu) psi  on a 4x3 colour-spinor */
static void spin_proj_full(int mu, double s, const cpx *psi, cpx *v)
{
  for (int a = 0; a < 4; a++)
    for (int col = 0; col < 3; col++) {
      cpx acc = psi[3 * a + col];
      for (int b = 0; b < 4; b++) {
        cpx g = cmk(GAMMA[mu][a][b][0], GAMMA[mu][a][b][1]);
        if (g.re != 0 || g.im != 0) acc = cadd(acc, cscale(s, cmul(g, psi[3 * b + col])));
      }
      v[3 * a + col] = acc;
    }
}

static void load_cpx(const double *p, int n, cpx *c)
{
  for (int k = 0; k < n; k++) c[k] = cmk(p[2 * k], p[2 * k + 1]);
}

/* ------------------------------------------------------------------------ */
/* hopping term (dslash_body.h:491-512 coefficients; codegen dslash.cc:332   */
/* twisted-BC phases applied to every hop of that dimension)                 */
/* ------------------------------------------------------------------------ */
/* hop restricted to the terms whose bit is set in mask: bit 2mu = backward
 * neighbour term U_mu(x-mu)^+ (1 + isign g_mu) psi(x-mu), bit 2mu+1 = forward
 * term U_mu(x) (1 - isign g_mu) psi(x+mu) (the accumulate[] order of
 * dslash_body.h:410-642); mask 0xff is the full hop */
void orc_hop_masked(const int L[4], const double *u, const double *psi, double *out,
                    int isign, int cb, const orc_hop_params *p, unsigned mask)
{
  long vh = orc_vh(L);
  int ocb = 1 - cb;
  /* sites are independent: the threads share no writes */
#pragma omp parallel for schedule(static)
  for (long j = 0; j < vh; j++) {
    int c[4];
    orc_site_coords(L, cb * vh + j, c);
    long self = cb * vh + j;
    cpx acc[12];
    for (int k = 0; k < 12; k++) acc[k] = cmk(0, 0);
    for (int mu = 0; mu < 4; mu++) {
      double cs = (mu == 3) ? p->aniso_t : p->aniso_s;
      cpx ph = cmk(1, 0);
      if (p->use_tbc[mu]) ph = cmk(p->tbc[mu][0], p->tbc[mu][1]);
      /* forward */
      if (mask >> (2 * mu + 1) & 1u) {
        double cf = cs;
        if (mu == 3 && c[3] == L[3] - 1) cf *= p->t_boundary;
        long nb = neighbour(L, c, mu, +1);
        cpx ps[12], v[12], U[9];
        load_cpx(psi + (nb - ocb * vh) * 24, 12, ps);
        spin_proj_full(mu, -(double)isign, ps, v);
        load_cpx(u + (self * 4 + mu) * 18, 9, U);
        for (int a = 0; a < 4; a++)
          for (int r = 0; r < 3; r++) {
            cpx s = {0, 0};
            for (int k = 0; k < 3; k++) s = cadd(s, cmul(U[3 * r + k], v[3 * a + k]));
            s = cmul(ph, s);
            acc[3 * a + r] = cadd(acc[3 * a + r], cscale(cf, s));
          }
      }
      /* backward */
      if (mask >> (2 * mu) & 1u) {
        double cbk = cs;
        if (mu == 3 && c[3] == 0) cbk *= p->t_boundary;
        long nb = neighbour(L, c, mu, -1);
        cpx ps[12], v[12], U[9];
        load_cpx(psi + (nb - ocb * vh) * 24, 12, ps);
        spin_proj_full(mu, (double)isign, ps, v);
        load_cpx(u + (nb * 4 + mu) * 18, 9, U);
        for (int a = 0; a < 4; a++)
          for (int r = 0; r < 3; r++) {
            cpx s = {0, 0};
            for (int k = 0; k < 3; k++) s = cadd(s, cmulc(U[3 * k + r], v[3 * a + k]));
            s = cmul(cconj(ph), s);
            acc[3 * a + r] = cadd(acc[3 * a + r], cscale(cbk, s));
          }
      }
    }
    for (int k = 0; k < 12; k++) {
      out[j * 24 + 2 * k] = acc[k].re;
      out[j * 24 + 2 * k + 1] = acc[k].im;
    }
  }
}

void orc_hop(const int L[4], const double *u, const double *psi, double *out,
             int isign, int cb, const orc_hop_params *p)
{
  orc_hop_masked(L, u, psi, out, isign, cb, p, 0xffu);
}

/* Dslash::dslashDir (dslash_body.h:350-358, 1074-1137): one term of the hop
 * with isign = +1. DPsiDir keeps dir_mask[2mu + recv_from] with
 * recv_from = 1 - dir % 2, so dir = 2mu takes the forward term (psi(x+mu)) and
 * dir = 2mu+1 the backward one (psi(x-mu)). */
void orc_dslash_dir(const int L[4], const double *u, const double *psi, double *out,
                    int cb, int dir, const orc_hop_params *p)
{
  orc_hop_masked(L, u, psi, out, 1, cb, p, 1u << (dir ^ 1));
}

void orc_achimbdpsi(const int L[4], const double *u, const double *psi,
                    const double *chi, double *out, double alpha, double beta,
                    int isign, int cb, const orc_hop_params *p)
{
  long n = orc_vh(L) * 24;
  double *h = (double *)malloc(n * sizeof(double));
  orc_hop(L, u, psi, h, isign, cb, p);
  for (long k = 0; k < n; k++) out[k] = alpha * chi[k] - beta * h[k];
  free(h);
}

/* ------------------------------------------------------------------------ */
/* clover term                                                               */
/* ------------------------------------------------------------------------ */
static void mat_adj3(const cpx *a, cpx *b)
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) b[3 * i + j] = cconj(a[3 * j + i]);
}

static void link(const int L[4], const double *u, const int c[4], int mu, cpx *U)
{
  long i = orc_site_index(L, c[0], c[1], c[2], c[3]);
  load_cpx(u + (i * 4 + mu) * 18, 9, U);
}

static void shifted(const int L[4], const int c[4], int mu, int dmu, int nu, int dnu, int out[4])
{
  for (int k = 0; k < 4; k++) out[k] = c[k];
  out[mu] = (out[mu] + dmu + L[mu]) % L[mu];
  out[nu] = (out[nu] + dnu + L[nu]) % L[nu];
}

/* F_{mu nu}(x) = (Q - Q^+)/8, Q = sum of the four plaquette leaves at x
 * (tests/mesfield.cc:24-64) */
void orc_field_strength(const int L[4], const double *u, double *f)
{
  long V = 2 * orc_vh(L);
  for (long i = 0; i < V; i++) {
    int c[4];
    orc_site_coords(L, i, c);
    int k = 0;
    for (int mu = 0; mu < 3; mu++)
      for (int nu = mu + 1; nu < 4; nu++, k++) {
        int s1[4], s2[4], s3[4];
        cpx A[9], B[9], C[9], D[9], t1[9], t2[9], Q[9], P[9], Ad[9];
        for (int q = 0; q < 9; q++) Q[q] = cmk(0, 0);
        /* P1 = U_mu(x) U_nu(x+mu) U_mu(x+nu)^+ U_nu(x)^+ */
        link(L, u, c, mu, A);
        shifted(L, c, mu, 1, nu, 0, s1); link(L, u, s1, nu, B);
        shifted(L, c, mu, 0, nu, 1, s2); link(L, u, s2, mu, C); mat_adj3(C, Ad); memcpy(C, Ad, sizeof(C));
        link(L, u, c, nu, D); mat_adj3(D, Ad); memcpy(D, Ad, sizeof(D));
        mat_mul3(A, B, t1); mat_mul3(t1, C, t2); mat_mul3(t2, D, P);
        for (int q = 0; q < 9; q++) Q[q] = cadd(Q[q], P[q]);
        /* P2 = U_nu(x) U_mu(x-mu+nu)^+ U_nu(x-mu)^+ U_mu(x-mu) */
        link(L, u, c, nu, A);
        shifted(L, c, mu, -1, nu, 1, s1); link(L, u, s1, mu, B); mat_adj3(B, Ad); memcpy(B, Ad, sizeof(B));
        shifted(L, c, mu, -1, nu, 0, s2); link(L, u, s2, nu, C); mat_adj3(C, Ad); memcpy(C, Ad, sizeof(C));
        link(L, u, s2, mu, D);
        mat_mul3(A, B, t1); mat_mul3(t1, C, t2); mat_mul3(t2, D, P);
        for (int q = 0; q < 9; q++) Q[q] = cadd(Q[q], P[q]);
        /* P3 = U_mu(x-mu)^+ U_nu(x-mu-nu)^+ U_mu(x-mu-nu) U_nu(x-nu) */
        shifted(L, c, mu, -1, nu, 0, s1); link(L, u, s1, mu, A); mat_adj3(A, Ad); memcpy(A, Ad, sizeof(A));
        shifted(L, c, mu, -1, nu, -1, s2); link(L, u, s2, nu, B); mat_adj3(B, Ad); memcpy(B, Ad, sizeof(B));
        link(L, u, s2, mu, C);
        shifted(L, c, mu, 0, nu, -1, s3); link(L, u, s3, nu, D);
        mat_mul3(A, B, t1); mat_mul3(t1, C, t2); mat_mul3(t2, D, P);
        for (int q = 0; q < 9; q++) Q[q] = cadd(Q[q], P[q]);
        /* P4 = U_nu(x-nu)^+ U_mu(x-nu) U_nu(x-nu+mu) U_mu(x)^+ */
        shifted(L, c, mu, 0, nu, -1, s1); link(L, u, s1, nu, A); mat_adj3(A, Ad); memcpy(A, Ad, sizeof(A));
        link(L, u, s1, mu, B);
        shifted(L, c, mu, 1, nu, -1, s2); link(L, u, s2, nu, C);
        link(L, u, c, mu, D); mat_adj3(D, Ad); memcpy(D, Ad, sizeof(D));
        mat_mul3(A, B, t1); mat_mul3(t1, C, t2); mat_mul3(t2, D, P);
        for (int q = 0; q < 9; q++) Q[q] = cadd(Q[q], P[q]);
        double *fo = f + (i * 6 + k) * 18;
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            cpx v = cscale(0.125, csub(Q[3 * a + b], cconj(Q[3 * b + a])));
            fo[2 * (3 * a + b)] = v.re;
            fo[2 * (3 * a + b) + 1] = v.im;
          }
      }
  }
}

/* Chroma PrimitiveClovTriang packing inside the 72-double site record */
static inline double *tri_diag(double *s, int b) { return s + 6 * b; }
static inline double *tri_offd(double *s, int b) { return s + 12 + 30 * b; }
static inline const double *tri_diag_c(const double *s, int b) { return s + 6 * b; }
static inline const double *tri_offd_c(const double *s, int b) { return s + 12 + 30 * b; }

/* makeClovSiteLoop (tests/clover_term_qdp_w.h:325-442) with clover
 * coefficients 0.5*c_sw folded in as in create() (:166-181) */
void orc_clover_build(const int L[4], const double *u, double csw_r, double csw_t,
                      double diag_mass, double *clov)
{
  long V = 2 * orc_vh(L);
  double *f = (double *)malloc(V * 6 * 18 * sizeof(double));
  orc_field_strength(L, u, f);
  const double cr = 0.5 * csw_r, ct = 0.5 * csw_t;
  /* (mu,nu) order 01 02 03 12 13 23; planes with t (=3) take ct */
  const double coef[6] = {cr, cr, ct, cr, ct, ct};
  for (long i = 0; i < V; i++) {
    cpx F[6][9];
    for (int k = 0; k < 6; k++) {
      load_cpx(f + (i * 6 + k) * 18, 9, F[k]);
      for (int q = 0; q < 9; q++) F[k][q] = cscale(coef[k], F[k][q]);
    }
    double *s = clov + i * 72;
    double *d0 = tri_diag(s, 0), *d1 = tri_diag(s, 1);
    double *o0 = tri_offd(s, 0), *o1 = tri_offd(s, 1);
    for (int q = 0; q < 6; q++) { d0[q] = diag_mass; d1[q] = diag_mass; }
    for (int a = 0; a < 3; a++) {
      double r0 = F[5][4 * a].im - F[0][4 * a].im;
      d0[a] += r0;
      d0[a + 3] -= r0;
      double r1 = F[5][4 * a].im + F[0][4 * a].im;
      d1[a] -= r1;
      d1[a + 3] += r1;
    }
    for (int a = 1; a < 3; a++)
      for (int b = 0; b < a; b++) {
        int eij = a * (a - 1) / 2 + b;
        int etmp = (a + 3) * (a + 2) / 2 + b + 3;
        cpx c0 = ctimesI(csub(F[0][3 * a + b], F[5][3 * a + b]));
        o0[2 * eij] = c0.re; o0[2 * eij + 1] = c0.im;
        o0[2 * etmp] = -c0.re; o0[2 * etmp + 1] = -c0.im;
        cpx c1 = ctimesI(cadd(F[5][3 * a + b], F[0][3 * a + b]));
        o1[2 * eij] = c1.re; o1[2 * eij + 1] = c1.im;
        o1[2 * etmp] = -c1.re; o1[2 * etmp + 1] = -c1.im;
      }
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        int e = (a + 3) * (a + 2) / 2 + b;
        cpx Em = cadd(ctimesI(F[2][3 * a + b]), F[4][3 * a + b]);
        cpx Bm = csub(ctimesI(F[3][3 * a + b]), F[1][3 * a + b]);
        cpx v0 = csub(Bm, Em), v1 = cadd(Em, Bm);
        o0[2 * e] = v0.re; o0[2 * e + 1] = v0.im;
        o1[2 * e] = v1.re; o1[2 * e + 1] = v1.im;
      }
  }
  free(f);
}

/* unpack Hermitian block b of a packed site into a dense 6x6 complex matrix */
void orc_clover_unpack_block(const double *site72, int b, double M[72])
{
  const double *d = tri_diag_c(site72, b), *o = tri_offd_c(site72, b);
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      double re, im;
      if (i == j) { re = d[i]; im = 0; }
      else if (i > j) { int e = i * (i - 1) / 2 + j; re = o[2 * e]; im = o[2 * e + 1]; }
      else { int e = j * (j - 1) / 2 + i; re = o[2 * e]; im = -o[2 * e + 1]; }
      M[2 * (6 * i + j)] = re;
      M[2 * (6 * i + j) + 1] = im;
    }
}

void orc_clover_pack_block(const double M[72], int b, double *site72)
{
  double *d = tri_diag(site72, b), *o = tri_offd(site72, b);
  for (int i = 0; i < 6; i++) d[i] = M[2 * (7 * i)];
  for (int i = 1; i < 6; i++)
    for (int j = 0; j < i; j++) {
      int e = i * (i - 1) / 2 + j;
      o[2 * e] = M[2 * (6 * i + j)];
      o[2 * e + 1] = M[2 * (6 * i + j) + 1];
    }
}

/* complex Gauss-Jordan with partial pivoting; returns 0 on success */
int orc_cmat_inverse(int n, const double *A, double *Ainv)
{
  cpx a[12][24];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < 2 * n; j++)
      a[i][j] = (j < n) ? cmk(A[2 * (n * i + j)], A[2 * (n * i + j) + 1])
                        : cmk(j - n == i ? 1 : 0, 0);
  for (int c = 0; c < n; c++) {
    int piv = c;
    double best = -1;
    for (int r = c; r < n; r++) {
      double m = a[r][c].re * a[r][c].re + a[r][c].im * a[r][c].im;
      if (m > best) { best = m; piv = r; }
    }
    if (best <= 0) return 1;
    if (piv != c)
      for (int j = 0; j < 2 * n; j++) { cpx t = a[c][j]; a[c][j] = a[piv][j]; a[piv][j] = t; }
    cpx p = a[c][c];
    double den = p.re * p.re + p.im * p.im;
    cpx pinv = cmk(p.re / den, -p.im / den);
    for (int j = 0; j < 2 * n; j++) a[c][j] = cmul(pinv, a[c][j]);
    for (int r = 0; r < n; r++) {
      if (r == c) continue;
      cpx fct = a[r][c];
      if (fct.re == 0 && fct.im == 0) continue;
      for (int j = 0; j < 2 * n; j++) a[r][j] = csub(a[r][j], cmul(fct, a[c][j]));
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      Ainv[2 * (n * i + j)] = a[i][j + n].re;
      Ainv[2 * (n * i + j) + 1] = a[i][j + n].im;
    }
  return 0;
}

void orc_clover_invert(long nsites, const double *clov, double *invclov)
{
  for (long i = 0; i < nsites; i++)
    for (int b = 0; b < 2; b++) {
      double M[72], Mi[72];
      orc_clover_unpack_block(clov + i * 72, b, M);
      orc_cmat_inverse(6, M, Mi);
      orc_clover_pack_block(Mi, b, invclov + i * 72);
    }
}

/* block b carries gamma_5 = +1 (b = 0) or -1 (b = 1) */
void orc_full_clover_tm_inverse(long nsites, const double *clov, double mu, double *fcl)
{
  for (long i = 0; i < nsites; i++)
    for (int b = 0; b < 2; b++) {
      double M[72], Mi[72];
      orc_clover_unpack_block(clov + i * 72, b, M);
      double g5 = (b == 0) ? 1.0 : -1.0;
      for (int q = 0; q < 6; q++) M[2 * (7 * q) + 1] += g5 * mu;
      orc_cmat_inverse(6, M, Mi);
      memcpy(fcl + i * 144 + b * 72, Mi, sizeof(Mi));
    }
}

static void cmat_mul(int n, const double *A, const double *B, double *C)
{
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      cpx s = {0, 0};
      for (int k = 0; k < n; k++)
        s = cadd(s, cmul(cmk(A[2 * (n * i + k)], A[2 * (n * i + k) + 1]),
                         cmk(B[2 * (n * k + j)], B[2 * (n * k + j) + 1])));
      C[2 * (n * i + j)] = s.re;
      C[2 * (n * i + j) + 1] = s.im;
    }
}

void orc_ndclover_inverse(long nsites, const double *clov, double mu, double eps,
                          double *fcl_u, double *fcl_d, double *coff)
{
  for (long i = 0; i < nsites; i++)
    for (int b = 0; b < 2; b++) {
      double A[72], A2[72], Ni[72], T[72], C[72];
      orc_clover_unpack_block(clov + i * 72, b, A);
      cmat_mul(6, A, A, A2);
      for (int q = 0; q < 6; q++) A2[2 * (7 * q)] += mu * mu - eps * eps;
      orc_cmat_inverse(6, A2, Ni);
      double g5 = (b == 0) ? 1.0 : -1.0;
      /* up: Ni (A - i mu g5) ; down: Ni (A + i mu g5) */
      memcpy(T, A, sizeof(T));
      for (int q = 0; q < 6; q++) T[2 * (7 * q) + 1] -= g5 * mu;
      cmat_mul(6, Ni, T, fcl_u + i * 144 + b * 72);
      memcpy(T, A, sizeof(T));
      for (int q = 0; q < 6; q++) T[2 * (7 * q) + 1] += g5 * mu;
      cmat_mul(6, Ni, T, fcl_d + i * 144 + b * 72);
      for (int q = 0; q < 72; q++) C[q] = eps * Ni[q];
      orc_clover_pack_block(C, b, coff + i * 72);
    }
}

/* clover_term (codegen dslash_common.cc:1346-1434): block b acts on spins
 * 2b, 2b+1 with index 3*(spin-2b)+colour */
void orc_clover_apply(long nsites, const double *clov, const double *in, double *out)
{
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nsites; i++)
    for (int b = 0; b < 2; b++) {
      double M[72];
      orc_clover_unpack_block(clov + i * 72, b, M);
      const double *v = in + i * 24 + b * 12;
      double *o = out + i * 24 + b * 12;
      for (int r = 0; r < 6; r++) {
        cpx s = {0, 0};
        for (int k = 0; k < 6; k++)
          s = cadd(s, cmul(cmk(M[2 * (6 * r + k)], M[2 * (6 * r + k) + 1]), cmk(v[2 * k], v[2 * k + 1])));
        o[2 * r] = s.re;
        o[2 * r + 1] = s.im;
      }
    }
}

void orc_full_clover_apply(long nsites, const double *fcl, const double *in, double *out)
{
  for (long i = 0; i < nsites; i++)
    for (int b = 0; b < 2; b++) {
      const double *M = fcl + i * 144 + b * 72;
      const double *v = in + i * 24 + b * 12;
      double *o = out + i * 24 + b * 12;
      for (int r = 0; r < 6; r++) {
        cpx s = {0, 0};
        for (int k = 0; k < 6; k++)
          s = cadd(s, cmul(cmk(M[2 * (6 * r + k)], M[2 * (6 * r + k) + 1]), cmk(v[2 * k], v[2 * k + 1])));
        o[2 * r] = s.re;
        o[2 * r + 1] = s.im;
      }
    }
}

/* out = (alpha + i mu gamma_5) in  (complex_functors.h:428-465) */
void orc_twisted_apply(long nsites, double alpha, double mu, const double *in, double *out)
{
  for (long i = 0; i < nsites; i++)
    for (int sp = 0; sp < 4; sp++) {
      double g = (sp < 2) ? mu : -mu;
      for (int col = 0; col < 3; col++) {
        int k = i * 24 + (3 * sp + col) * 2;
        double re = in[k], im = in[k + 1];
        out[k] = alpha * re - g * im;
        out[k + 1] = alpha * im + g * re;
      }
    }
}

/* out += i*s*rho*gamma_5 in */
static void add_i_g5(long nsites, double srho, const double *in, double *out)
{
  for (long i = 0; i < nsites; i++)
    for (int sp = 0; sp < 4; sp++) {
      double g = (sp < 2) ? srho : -srho;
      for (int col = 0; col < 3; col++) {
        int k = i * 24 + (3 * sp + col) * 2;
        double re = in[k], im = in[k + 1];
        out[k] += -g * im;
        out[k + 1] += g * re;
      }
    }
}

/* ------------------------------------------------------------------------ */
/* even-odd operators                                                        */
/* ------------------------------------------------------------------------ */
int orc_op_nflav(const orc_op *op)
{
  return (op->kind == ORC_NDTMWILSON || op->kind == ORC_NDTMCLOVER) ? 2 : 1;
}

void orc_op_apply(const orc_op *op, double *res, const double *in, int isign, int target_cb)
{
  const int *L = op->L;
  long vh = orc_vh(L);
  long n = vh * 24;
  int ocb = 1 - target_cb;
  int nf = orc_op_nflav(op);
  double *tmp = (double *)calloc(2 * n, sizeof(double));
  double *tmp2 = (double *)calloc(2 * n, sizeof(double));
  double *h = (double *)calloc(n, sizeof(double));
  double alpha = 4.0 + op->mass;
  switch (op->kind) {
    case ORC_WILSON: {
      /* wilson.h:49-65: alpha*in - (0.25/alpha) D_oe D_eo in */
      orc_hop(L, op->u, in, tmp, isign, ocb, &op->hop);
      orc_hop(L, op->u, tmp, h, isign, target_cb, &op->hop);
      for (long k = 0; k < n; k++) res[k] = alpha * in[k] - (0.25 / alpha) * h[k];
    } break;
    case ORC_CLOVER: {
      /* clover.h:162-172: A in + i isign rho g5 in - 1/4 D A^-1 D in */
      orc_hop(L, op->u, in, h, isign, ocb, &op->hop);
      orc_clover_apply(vh, op->invclov[ocb], h, tmp);
      orc_hop(L, op->u, tmp, h, isign, target_cb, &op->hop);
      orc_clover_apply(vh, op->clov[target_cb], in, res);
      add_i_g5(vh, isign * op->rho, in, res);
      for (long k = 0; k < n; k++) res[k] -= 0.25 * h[k];
    } break;
    case ORC_TMWILSON: {
      /* twisted_mass.h:62-77, tm_dslash_body.h:20-45 */
      double den = alpha * alpha + op->mu * op->mu;
      orc_hop(L, op->u, in, h, isign, ocb, &op->hop);
      orc_twisted_apply(vh, alpha / den, -isign * op->mu / den, h, tmp);
      orc_hop(L, op->u, tmp, h, isign, target_cb, &op->hop);
      orc_twisted_apply(vh, alpha, isign * op->mu, in, res);
      for (long k = 0; k < n; k++) res[k] -= 0.25 * h[k];
    } break;
    case ORC_TMCLOVER: {
      /* twisted_mass_clover.h:92-102; clidx = isign==1 ? 0 : 1 */
      int clidx = (isign == 1) ? 0 : 1;
      orc_hop(L, op->u, in, h, isign, ocb, &op->hop);
      orc_full_clover_apply(vh, op->fcl[ocb][clidx], h, tmp);
      orc_hop(L, op->u, tmp, h, isign, target_cb, &op->hop);
      orc_clover_apply(vh, op->clov[target_cb], in, res);
      add_i_g5(vh, isign * op->rho, in, res);
      for (long k = 0; k < n; k++) res[k] -= 0.25 * h[k];
    } break;
    case ORC_NDTMWILSON: {
      /* ndtm_reuse_operator.h:65-108 */
      double muinv = 1.0 / (alpha * alpha + op->mu * op->mu - op->eps * op->eps);
      double are = alpha * muinv, aim = -isign * op->mu * muinv, e = op->eps * muinv;
      for (int f = 0; f < 2; f++) orc_hop(L, op->u, in + f * n, tmp + f * n, isign, ocb, &op->hop);
      /* (a + i b g5 tau3 + e tau1) */
      for (long i = 0; i < vh; i++)
        for (int sp = 0; sp < 4; sp++) {
          double bg = (sp < 2) ? aim : -aim;
          for (int col = 0; col < 3; col++) {
            long k = i * 24 + (3 * sp + col) * 2;
            double ur = tmp[k], ui = tmp[k + 1], dr = tmp[n + k], di = tmp[n + k + 1];
            tmp2[k] = are * ur - bg * ui + e * dr;
            tmp2[k + 1] = are * ui + bg * ur + e * di;
            tmp2[n + k] = are * dr + bg * di + e * ur;
            tmp2[n + k + 1] = are * di - bg * dr + e * ui;
          }
        }
      for (int f = 0; f < 2; f++) {
        double musign = 1.0 - 2.0 * f;
        orc_hop(L, op->u, tmp2 + f * n, h, isign, target_cb, &op->hop);
        orc_twisted_apply(vh, alpha, isign * musign * op->mu, in + f * n, res + f * n);
        for (long k = 0; k < n; k++) res[f * n + k] -= 0.25 * h[k] + op->eps * in[(1 - f) * n + k];
      }
    } break;
    case ORC_NDTMCLOVER: {
      /* ndtm_reuse_operator_clover.h:115-128, tm_clov_dslash_body.h:1184-1367 */
      for (int f = 0; f < 2; f++) orc_hop(L, op->u, in + f * n, tmp + f * n, isign, ocb, &op->hop);
      const double *fu = op->ndfcl[ocb][isign == 1 ? 0 : 1];
      const double *fd = op->ndfcl[ocb][isign == 1 ? 1 : 0];
      orc_full_clover_apply(vh, fu, tmp, tmp2);
      orc_full_clover_apply(vh, fd, tmp + n, tmp2 + n);
      orc_clover_apply(vh, op->ndcoff[ocb], tmp + n, h);
      for (long k = 0; k < n; k++) tmp2[k] += h[k];
      orc_clover_apply(vh, op->ndcoff[ocb], tmp, h);
      for (long k = 0; k < n; k++) tmp2[n + k] += h[k];
      for (int f = 0; f < 2; f++) {
        double rho = (f == 1 ? -1.0 : 1.0) * op->rho;
        orc_hop(L, op->u, tmp2 + f * n, h, isign, target_cb, &op->hop);
        orc_clover_apply(vh, op->clov[target_cb], in + f * n, res + f * n);
        add_i_g5(vh, isign * rho, in + f * n, res + f * n);
        for (long k = 0; k < n; k++) res[f * n + k] -= 0.25 * h[k] + op->eps * in[(1 - f) * n + k];
      }
    } break;
  }
  (void)nf;
  free(tmp);
  free(tmp2);
  free(h);
}

/* ------------------------------------------------------------------------ */
/* BLAS                                                                      */
/* ------------------------------------------------------------------------ */
double orc_norm2(long n, const double *x)
{
  double s = 0;
  for (long k = 0; k < n; k++) s += x[k] * x[k];
  return s;
}

/* <l, r> = sum conj(l) r  (complex_functors.h:11-64) */
void orc_cdot(long n, const double *l, const double *r, double out[2])
{
  double re = 0, im = 0;
  for (long k = 0; k < n; k += 2) {
    re += l[k] * r[k] + l[k + 1] * r[k + 1];
    im += l[k] * r[k + 1] - l[k + 1] * r[k];
  }
  out[0] = re;
  out[1] = im;
}

/* ------------------------------------------------------------------------ */
/* solvers                                                                   */
/* ------------------------------------------------------------------------ */
static void mdagm(const orc_op *op, double *mmp, double *mp, const double *p, int isign, int cb)
{
  orc_op_apply(op, mp, p, isign, cb);
  orc_op_apply(op, mmp, mp, -isign, cb);
}

/* invcg.h:88-456 (RELATIVE residual) */
void orc_cg(const orc_op *op, double *x, const double *rhs, double rsd_target,
            int max_iters, int isign, int cb, orc_solve_stats *st)
{
  long n = orc_vh(op->L) * 24 * orc_op_nflav(op);
  double *r = (double *)malloc(n * sizeof(double)), *p = (double *)malloc(n * sizeof(double));
  double *mp = (double *)malloc(n * sizeof(double)), *mmp = (double *)malloc(n * sizeof(double));
  double chi_sq = orc_norm2(n, rhs);
  double rsd_sq = rsd_target * rsd_target * chi_sq;
  st->mv_apps = 0;
  st->status = 0;
  mdagm(op, mmp, mp, x, isign, cb);
  st->mv_apps += 2;
  for (long k = 0; k < n; k++) r[k] = rhs[k] - mmp[k];
  double cp = orc_norm2(n, r);
  if (cp <= rsd_sq) {
    st->n_iters = 0;
    st->rsd_sq_final = cp / chi_sq;
    goto done;
  }
  memcpy(p, r, n * sizeof(double));
  for (int k = 0; k <= max_iters; ++k) {
    double c = cp;
    mdagm(op, mmp, mp, p, isign, cb);
    st->mv_apps += 2;
    double d = orc_norm2(n, mp);
    double a = c / d;
    for (long q = 0; q < n; q++) { r[q] -= a * mmp[q]; x[q] += a * p[q]; }
    cp = orc_norm2(n, r);
    if (cp < rsd_sq) {
      st->n_iters = k;
      mdagm(op, mmp, mp, x, isign, cb);
      st->mv_apps += 2;
      for (long q = 0; q < n; q++) r[q] = rhs[q] - mmp[q];
      st->rsd_sq_final = orc_norm2(n, r);
      goto done;
    }
    double b = cp / c;
    for (long q = 0; q < n; q++) p[q] = r[q] + b * p[q];
  }
  st->n_iters = max_iters;
  st->rsd_sq_final = cp;
  st->status = 5;
done:
  free(r); free(p); free(mp); free(mmp);
}

static void cdiv(double *o, const double *a, const double *b)
{
  double den = b[0] * b[0] + b[1] * b[1];
  double re = (a[0] * b[0] + a[1] * b[1]) / den;
  double im = (a[1] * b[0] - a[0] * b[1]) / den;
  o[0] = re; o[1] = im;
}
static void cmul2(double *o, const double *a, const double *b)
{
  double re = a[0] * b[0] - a[1] * b[1], im = a[0] * b[1] + a[1] * b[0];
  o[0] = re; o[1] = im;
}

/* invbicgstab.h:86-262 (RELATIVE residual); n_iters reported as there */
void orc_bicgstab(const orc_op *op, double *x, const double *rhs, double rsd_target,
                  int max_iters, int isign, int cb, orc_solve_stats *st)
{
  long n = orc_vh(op->L) * 24 * orc_op_nflav(op);
  double *r0 = (double *)malloc(n * sizeof(double)), *r = (double *)malloc(n * sizeof(double));
  double *p = (double *)calloc(n, sizeof(double)), *v = (double *)calloc(n, sizeof(double));
  double *t = (double *)malloc(n * sizeof(double));
  double rhs_sq = orc_norm2(n, rhs);
  double rsd_sq = rsd_target * rsd_target * rhs_sq;
  st->mv_apps = 0;
  st->status = 0;
  orc_op_apply(op, r0, x, isign, cb);
  st->mv_apps++;
  for (long q = 0; q < n; q++) r0[q] = rhs[q] - r0[q];
  double r_norm = orc_norm2(n, r0);
  if (r_norm < rsd_sq) {
    st->rsd_sq_final = r_norm;
    st->n_iters = 0;
    goto done;
  }
  memcpy(r, r0, n * sizeof(double));
  double rho_prev[2] = {1, 0}, alpha[2] = {1, 0}, omega[2] = {1, 0}, rho[2];
  int notconv = 1, k;
  for (k = 1; k <= max_iters && notconv; k++) {
    orc_cdot(n, r0, r, rho);
    if (rho[0] == 0 && rho[1] == 0) { st->status = 4; st->rsd_sq_final = r_norm; st->n_iters = k; goto done; }
    double t1[2], t2[2], beta[2];
    cdiv(t1, rho, rho_prev);
    cdiv(t2, alpha, omega);
    cmul2(beta, t1, t2);
    /* p = r + beta (p - omega v) */
    for (long q = 0; q < n; q += 2) {
      double tr = p[q] - (omega[0] * v[q] - omega[1] * v[q + 1]);
      double ti = p[q + 1] - (omega[0] * v[q + 1] + omega[1] * v[q]);
      p[q] = r[q] + beta[0] * tr - beta[1] * ti;
      p[q + 1] = r[q + 1] + beta[0] * ti + beta[1] * tr;
    }
    orc_op_apply(op, v, p, isign, cb);
    st->mv_apps++;
    double tmp[2];
    orc_cdot(n, r0, v, tmp);
    if (tmp[0] == 0 && tmp[1] == 0) { st->status = 4; st->n_iters = k; goto done; }
    cdiv(alpha, rho, tmp);
    rho_prev[0] = rho[0]; rho_prev[1] = rho[1];
    /* s = r - alpha v (stored in r) */
    for (long q = 0; q < n; q += 2) {
      double vr = v[q], vi = v[q + 1];
      r[q] -= alpha[0] * vr - alpha[1] * vi;
      r[q + 1] -= alpha[0] * vi + alpha[1] * vr;
    }
    orc_op_apply(op, t, r, isign, cb);
    st->mv_apps++;
    double t_norm = orc_norm2(n, t);
    if (t_norm == 0) { st->status = 4; st->rsd_sq_final = r_norm; st->n_iters = k; goto done; }
    orc_cdot(n, t, r, omega);
    omega[0] /= t_norm; omega[1] /= t_norm;
    /* x += omega r + alpha p ; r -= omega t */
    for (long q = 0; q < n; q += 2) {
      double rr = r[q], ri = r[q + 1];
      x[q] += omega[0] * rr - omega[1] * ri + alpha[0] * p[q] - alpha[1] * p[q + 1];
      x[q + 1] += omega[0] * ri + omega[1] * rr + alpha[0] * p[q + 1] + alpha[1] * p[q];
      r[q] = rr - (omega[0] * t[q] - omega[1] * t[q + 1]);
      r[q + 1] = ri - (omega[0] * t[q + 1] + omega[1] * t[q]);
    }
    r_norm = orc_norm2(n, r);
    if (r_norm < rsd_sq) notconv = 0;
  }
  st->rsd_sq_final = r_norm;
  st->n_iters = k;
  if (notconv) st->status = 5;
done:
  free(r0); free(r); free(p); free(v); free(t);
}

/* minvcg.h:77-387 */
void orc_minvcg(const orc_op *op, double *x, const double *rhs, int n_shift,
                const double *shifts, const double *rsd_target, int max_iters,
                int isign, int cb, int *n_iters, double *rsd_sq_final)
{
  long n = orc_vh(op->L) * 24 * orc_op_nflav(op);
  double *r = (double *)malloc(n * sizeof(double)), *p0 = (double *)malloc(n * sizeof(double));
  double *mp = (double *)malloc(n * sizeof(double)), *mmp = (double *)malloc(n * sizeof(double));
  double *p = (double *)malloc(n * n_shift * sizeof(double));
  double *zeta = (double *)malloc(n_shift * sizeof(double)), *zeta_prev = (double *)malloc(n_shift * sizeof(double));
  double *beta = (double *)malloc(n_shift * sizeof(double)), *rsd_sq = (double *)malloc(n_shift * sizeof(double));
  int *conv = (int *)calloc(n_shift, sizeof(int));
  double cp = orc_norm2(n, rhs);
  for (int s = 0; s < n_shift; s++) rsd_sq[s] = rsd_target[s] * rsd_target[s] * cp;
  memcpy(r, rhs, n * sizeof(double));
  memcpy(p0, rhs, n * sizeof(double));
  for (int s = 0; s < n_shift; s++) memcpy(p + s * n, rhs, n * sizeof(double));
  mdagm(op, mmp, mp, p0, isign, cb);
  double d = orc_norm2(n, mp);
  double b = -cp / d;
  for (long q = 0; q < n; q++) r[q] += b * mmp[q];
  double c = orc_norm2(n, r);
  for (int s = 0; s < n_shift; s++) {
    zeta_prev[s] = 1;
    zeta[s] = 1.0 / (1.0 - shifts[s] * b);
    beta[s] = b / (1.0 - shifts[s] * b);
    for (long q = 0; q < n; q++) x[s * n + q] = -beta[s] * rhs[q];
  }
  int convP = 0, k;
  *n_iters = 0;
  for (k = 1; k < max_iters && !convP; ++k) {
    double a = c / cp;
    for (long q = 0; q < n; q++) p0[q] = r[q] + a * p0[q];
    for (int s = 0; s < n_shift; s++)
      if (!conv[s]) {
        double as = a * zeta[s] * beta[s] / (zeta_prev[s] * b);
        for (long q = 0; q < n; q++) p[s * n + q] = zeta[s] * r[q] + as * p[s * n + q];
      }
    cp = c;
    double bp = b;
    mdagm(op, mmp, mp, p0, isign, cb);
    d = orc_norm2(n, mp);
    b = -cp / d;
    for (long q = 0; q < n; q++) r[q] += b * mmp[q];
    c = orc_norm2(n, r);
    convP = 1;
    for (int s = 0; s < n_shift; s++) {
      if (!conv[s]) {
        double z0 = zeta[s], z1 = zeta_prev[s];
        zeta_prev[s] = z0;
        zeta[s] = z0 * z1 * bp;
        zeta[s] /= (b * a * (z1 - z0) + z1 * bp * (1.0 - shifts[s] * b));
        beta[s] = b * zeta[s] / z0;
        for (long q = 0; q < n; q++) x[s * n + q] -= beta[s] * p[s * n + q];
        double css = c * zeta[s] * zeta[s];
        conv[s] = css < rsd_sq[s];
        if (conv[s]) rsd_sq_final[s] = css;
      }
      convP &= conv[s];
    }
    *n_iters = k;
  }
  free(r); free(p0); free(mp); free(mmp); free(p); free(zeta); free(zeta_prev);
  free(beta); free(rsd_sq); free(conv);
}

/* threads the parallel site loops use (1 without OpenMP) */
int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
