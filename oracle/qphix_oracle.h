/*
 * qphix_oracle.h -- CPU oracle for the even-odd Wilson / clover / twisted-mass
 * Dslash path, its fused BLAS and its Krylov solvers.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64 restatement of the
 * QPhiX algorithms (site loops of the hop and the clover apply split over
 * OpenMP threads, each site computed alone) that provides the reference
 * results of tests/ and of the smoke check, and the CPU baseline leg of
 * bench.py.  The product path (libqphix_amd.so) never calls it.
 *
 * Host layout (the "canonical" layout, the same one the C ABI packers accept;
 * it is the QDP++ checkerboarded-lexicographic order that
 * include/qphix/qdp_packer_parscalar.h:75,177 reads):
 *   site index   i(x,y,z,t) = cb*Vh + ((t*Nz + z)*Ny + y)*Nxh + x/2,
 *                cb = (x+y+z+t) & 1,  Vh = Nx*Ny*Nz*Nt/2, Nxh = Nx/2
 *   gauge        u[(i*4 + mu)*18 + (row*3 + col)*2 + reim]   U_mu(x), row-major
 *   cb spinor    psi[(i - cb*Vh)*24 + (spin*3 + col)*2 + reim]
 *   clover       per cb site 72 doubles in Chroma PrimitiveClovTriang order:
 *                diag[2][6], then offd[2][15][re,im]
 *                (tests/clover_term_qdp_w.h:325-442; qdp_packer_parscalar.h:196-216)
 *   full clover  per cb site 144 doubles: [block 2][row 6][col 6][re,im]
 *                (qdp_packer_parscalar.h:238-315)
 * Spin basis: DeGrand-Rossi (QDP++), gamma_5 = diag(1,1,-1,-1).
 */
#ifndef QPHIX_ORACLE_H
#define QPHIX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- geometry / indexing ------------------------------------------------ */
long orc_vh(const int L[4]);
long orc_site_index(const int L[4], int x, int y, int z, int t);
void orc_site_coords(const int L[4], long i, int c[4]);

/* ---- seeded fields ------------------------------------------------------ */
double orc_uniform(uint64_t seed, uint64_t counter);
void orc_gauss_pair(uint64_t seed, uint64_t counter, double g[2]);
/* U_mu(x) = reunit(exp(i*eps*A)), A traceless Hermitian from Gaussian G. */
void orc_gen_gauge(const int L[4], uint64_t seed, double eps, double *u);
/* Re, Im ~ N(0,1) per (site, spin, colour); cb in {0,1}. */
void orc_gen_spinor(const int L[4], uint64_t seed, int cb, double *psi);
void orc_su3_exp_i(const double A[18], double eps, double U[18]);
void orc_reunit(double U[18]);

/* ---- hopping term ------------------------------------------------------- */
typedef struct {
  double t_boundary, aniso_s, aniso_t;
  int use_tbc[4];
  double tbc[4][2];
} orc_hop_params;
/* out(x) = sum_mu c_f U_mu(x)(1 - isign g_mu) psi(x+mu)
 *              + c_b U_mu(x-mu)^+ (1 + isign g_mu) psi(x-mu),  x on target cb */
void orc_hop(const int L[4], const double *u, const double *psi, double *out,
             int isign, int cb, const orc_hop_params *p);
void orc_hop_masked(const int L[4], const double *u, const double *psi, double *out,
                    int isign, int cb, const orc_hop_params *p, unsigned mask);
void orc_dslash_dir(const int L[4], const double *u, const double *psi, double *out,
                    int cb, int dir, const orc_hop_params *p);
/* out = alpha*chi - beta*hop(psi) */
void orc_achimbdpsi(const int L[4], const double *u, const double *psi,
                    const double *chi, double *out, double alpha, double beta,
                    int isign, int cb, const orc_hop_params *p);
/* dense spin matrices (DeGrand-Rossi), [4][4][2] */
void orc_gamma(int mu, double g[32]);

/* ---- clover term -------------------------------------------------------- */
/* F_{mu nu}, mu<nu (6 planes) per site: f[(i*6 + k)*18 ...] */
void orc_field_strength(const int L[4], const double *u, double *f);
/* A = diag_mass + clover, Chroma packing, full lattice in cb-major order */
void orc_clover_build(const int L[4], const double *u, double csw_r, double csw_t,
                      double diag_mass, double *clov);
void orc_clover_unpack_block(const double *site72, int b, double M[72]);
void orc_clover_pack_block(const double M[72], int b, double *site72);
int orc_cmat_inverse(int n, const double *A, double *Ainv);
void orc_clover_invert(long nsites, const double *clov, double *invclov);
/* full blocks of (A + i*mu*g5)^-1 */
void orc_full_clover_tm_inverse(long nsites, const double *clov, double mu,
                                double *fcl);
/* doublet: fcl_u = (A^2+mu^2-eps^2)^-1 (A - i mu g5), fcl_d = ... (A + i mu g5),
 * c = eps (A^2+mu^2-eps^2)^-1  (packed Hermitian) */
void orc_ndclover_inverse(long nsites, const double *clov, double mu, double eps,
                          double *fcl_u, double *fcl_d, double *coff);
void orc_clover_apply(long nsites, const double *clov, const double *in, double *out);
void orc_full_clover_apply(long nsites, const double *fcl, const double *in,
                           double *out);

/* ---- operators (one parity; Schur complement) ---------------------------- */
enum { ORC_WILSON = 0, ORC_CLOVER = 1, ORC_TMWILSON = 2, ORC_TMCLOVER = 3,
       ORC_NDTMWILSON = 4, ORC_NDTMCLOVER = 5 };
typedef struct {
  int kind;
  int L[4];
  orc_hop_params hop;
  const double *u;           /* full canonical gauge */
  double mass;               /* Wilson/TM: alpha = 4 + mass */
  double mu, eps, rho;       /* twisted mass, doublet splitting, prec mass */
  const double *clov[2];     /* A per cb (Chroma packing)                   */
  const double *invclov[2];  /* A^-1 per cb                                 */
  const double *fcl[2][2];   /* [cb][0:+mu,1:-mu] full inverse blocks (TM)  */
  const double *ndfcl[2][2]; /* [cb][up,dn] doublet inverse blocks           */
  const double *ndcoff[2];   /* [cb] doublet flavour-off-diagonal block      */
} orc_op;
int orc_op_nflav(const orc_op *op);
/* res, in: nflav cb spinors, consecutive (flavour-major) */
void orc_op_apply(const orc_op *op, double *res, const double *in, int isign,
                  int target_cb);
void orc_twisted_apply(long nsites, double alpha, double mu, const double *in,
                       double *out);

/* ---- BLAS (reference real/complex functor semantics, n = reals) ---------- */
double orc_norm2(long n, const double *x);
void orc_cdot(long n, const double *l, const double *r, double out[2]);

/* ---- solvers -------------------------------------------------------------- */
typedef struct {
  int n_iters;
  double rsd_sq_final;
  unsigned long mv_apps;
  int status; /* 0 ok, 4 breakdown, 5 not converged */
} orc_solve_stats;
/* CG on M^+M; relative residual target like invcg.h */
void orc_cg(const orc_op *op, double *x, const double *rhs, double rsd_target,
            int max_iters, int isign, int cb, orc_solve_stats *st);
void orc_bicgstab(const orc_op *op, double *x, const double *rhs,
                  double rsd_target, int max_iters, int isign, int cb,
                  orc_solve_stats *st);
/* multishift CG on M^+M + sigma_s; x: n_shift consecutive solutions */
void orc_minvcg(const orc_op *op, double *x, const double *rhs, int n_shift,
                const double *shifts, const double *rsd_target, int max_iters,
                int isign, int cb, int *n_iters, double *rsd_sq_final);

/* threads of the OpenMP site loops (hop, clover apply); 1 without OpenMP */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
