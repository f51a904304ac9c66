/*
 * qphix_amd.h -- C ABI of the gfx950 (MI355X) implementation of the QPhiX
 * even-odd Dslash, fused BLAS and Krylov solvers.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (QPHIX_OK = 0) and is stream-ordered on the library's HIP stream
 * (qphix_set_stream). Entry points that return host scalars (reductions,
 * solvers) synchronise that stream. Field pointers are device pointers returned
 * by the qphix_alloc_* calls of the geometry they belong to; a geometry fixes
 * the lattice, the precision (the FT template argument of the reference) and
 * the gauge compression (compress12).
 *
 * Canonical host layouts accepted by the packers are those of the QDP++
 * packers (include/qphix/qdp_packer_parscalar.h):
 *   site index  i = cb*Vh + ((t*Nz + z)*Ny + y)*Nxh + x/2, cb = (x+y+z+t)&1
 *   spinor (one cb)  double[Vh][4 spin][3 colour][2]
 *   gauge  (full)    double[V][4 mu][3 row][3 col][2]     U_mu(x), no boundary phase
 *   clover (one cb)  double[Vh][72]  Chroma PrimitiveClovTriang: diag[2][6], offd[2][15][2]
 *   full clover      double[Vh][2 block][6][6][2]
 * The reference source lines cited below are in JeffersonLab/qphix.
 */
#ifndef QPHIX_AMD_H
#define QPHIX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (reference aborts / throws, dslash_body.h:64-95,
 *      geometry.h:151-154, minvcg.h:92-97, invbicgstab.h:163-230) ---------- */
enum {
  QPHIX_OK = 0,
  QPHIX_EDOMAIN = 1,     /* lattice / argument outside the supported domain   */
  QPHIX_ENOMEM = 2,      /* device allocation failed                           */
  QPHIX_EHIP = 3,        /* HIP runtime error                                  */
  QPHIX_EBREAKDOWN = 4,  /* solver breakdown (rho = 0, <r0,v> = 0, |t| = 0)     */
  QPHIX_ENOCONV = 5,     /* solver reached MaxIters                            */
  QPHIX_EINVAL = 6,      /* null / mismatched pointer or handle                */
  QPHIX_ENODEV = 7       /* no HIP device                                      */
};

/* precision of a geometry (Geometry<FT,...> FT = double / float / half;
 * arith_type.h:8-17 computes half in float). F16 (IEEE half) and BF16
 * (bfloat16) fields store a per-site FP32 scale, compute in FP32 and need
 * two-row gauge compression. */
enum { QPHIX_F64 = 0, QPHIX_F32 = 1, QPHIX_F16 = 2, QPHIX_BF16 = 3 };

typedef struct qphix_geom qphix_geom;
typedef struct qphix_dslash qphix_dslash;
typedef struct qphix_op qphix_op;
typedef struct qphix_solver qphix_solver;

const char *qphix_strerror(int status);
/* message of the last failing call on this host thread */
const char *qphix_last_error(void);

/* ---- device and stream --------------------------------------------------- */
int qphix_init(int device);
/* hip_stream is a hipStream_t; NULL selects the library's own stream */
int qphix_set_stream(void *hip_stream);
int qphix_sync(void);
/* HIP-event timer on the library stream */
int qphix_timer_start(void);
int qphix_timer_stop(double *ms);
/* Tuning hooks (no reference counterpart: the reference's blocking
 * parameters By/Bz/Sy/Sz play this role on the CPU). Hop kernels are
 * autotuned on first use per lattice unless QPHIX_TUNE=0; block is the
 * default work-group size (64..256, multiple of 64) when tuning is off. */
int qphix_set_hop_block(int block);
/* mode 0: no tuning; 1: autotune on first use; 2: force (block, bz, slabs):
 * bz z-planes per tile, slabs t-slabs (bz = 0: natural site order) */
int qphix_tune_config(int mode, int block, int bz, int slabs);
/* mode 2 with (y, z) tiles of by rows (0: all of y) by bz planes, and lds
 * bytes of LDS reserved per work-group (0..65536; the kernels do not use it,
 * it caps the work-groups resident per CU) */
int qphix_tune_traversal(int block, int bz, int by, int slabs, int lds);
/* multi-RHS tiling: reg sources per work-item (1, 2 or 4; each link and
 * clover block is loaded once for all of them; 8 selects the matrix-core
 * kernel for FP32 and FP16 (two-row) fields, 8 sources per work-group, other
 * precisions then fail with QPHIX_EHIP) and rt sources per launch (1..32, a
 * multiple of reg);
 * reg = 0 returns to autotuning. Clears the tuned multi-RHS configurations. */
int qphix_tune_rhs(int reg, int rt);
const char *qphix_tune_summary(void);

/* ---- geometry: replaces Geometry<FT,V,S,c12>(latt, By, Bz, NCores, Sy, Sz,
 *      PadXY, PadXYZ, MinCt) (geometry.h:130-235). Blocking / threading
 *      arguments have no meaning on the GPU; all four extents must be even. */
int qphix_geom_create(const int latt[4], int prec, int compress12, qphix_geom **g);
int qphix_geom_destroy(qphix_geom *g);
/* bytes[0..3] = spinor, gauge (one cb), clover, full clover allocation sizes */
int qphix_geom_info(const qphix_geom *g, int latt[4], int *prec, int *compress12,
                    long *vh, size_t bytes[4]);

/* ---- allocation: allocCBFourSpinor / allocCBGauge / allocCBClov /
 *      allocCBFullClov / free (geometry.h:270-398). Memory is zeroed. ------- */
int qphix_alloc_spinor(const qphix_geom *g, void **p);
int qphix_alloc_gauge(const qphix_geom *g, void **p);
int qphix_alloc_clover(const qphix_geom *g, void **p);
int qphix_alloc_full_clover(const qphix_geom *g, void **p);
int qphix_free(void *p);

/* ---- packers: qdp_pack_spinor / qdp_unpack_spinor / qdp_pack_gauge /
 *      qdp_pack_clover / qdp_pack_full_clover
 *      (qdp_packer_parscalar.h:32-146, 148-227, 230-327, 329-461) ---------- */
int qphix_pack_spinor(const qphix_geom *g, void *dst, const double *src, int cb);
int qphix_unpack_spinor(const qphix_geom *g, double *dst, const void *src, int cb);
/* fills both checkerboard gauge arrays (8 links per site incl. back links) */
int qphix_pack_gauge(const qphix_geom *g, void *u_cb0, void *u_cb1, const double *src);
int qphix_pack_clover(const qphix_geom *g, void *dst, const double *src, int cb);
int qphix_unpack_clover(const qphix_geom *g, double *dst, const void *src, int cb);
int qphix_pack_full_clover(const qphix_geom *g, void *dst, const double *src, int cb);

/* ---- precision conversion: convert() (blas_new_c.h:852-1031) for spinors
 *      with a scale factor; gauge / clover copies between geometries of the
 *      same lattice (what the mixed-precision callers build by re-packing). */
int qphix_convert_spinor(const qphix_geom *g_out, void *out, double scale,
                         const qphix_geom *g_in, const void *in);
int qphix_convert_gauge(const qphix_geom *g_out, void *out, const qphix_geom *g_in,
                        const void *in);
int qphix_convert_clover(const qphix_geom *g_out, void *out, const qphix_geom *g_in,
                         const void *in, int full);

/* ---- seeded field preparation on the device (replaces the QDP++ RNG set-up
 *      and tests/clover_term_qdp_w.h:114-195, 325-442, tests/mesfield.cc:24-64).
 *      gauge: U_mu(x) = reunit(exp(i*eps*A)), A traceless Hermitian with
 *      unit-variance Gaussian G: A = (G + G^+)/2 - tr(G + G^+)/6.  */
int qphix_gen_gauge(const qphix_geom *g, void *u_cb0, void *u_cb1, uint64_t seed, double eps);
/* Re, Im ~ N(0,1) per site, spin, colour */
int qphix_gen_spinor(const qphix_geom *g, void *dst, uint64_t seed, int cb);
enum {
  QPHIX_CLOV_A = 0,       /* A = diag_mass + clover term (packed Hermitian)  */
  QPHIX_CLOV_AINV = 1,    /* A^-1 (packed Hermitian)                          */
  QPHIX_CLOV_TMINV = 2,   /* full blocks (A + i*mu*gamma_5)^-1                */
  QPHIX_CLOV_ND_UP = 3,   /* full blocks (A^2+mu^2-eps^2)^-1 (A - i mu g5)    */
  QPHIX_CLOV_ND_DN = 4,   /* full blocks (A^2+mu^2-eps^2)^-1 (A + i mu g5)    */
  QPHIX_CLOV_ND_OFF = 5   /* eps (A^2+mu^2-eps^2)^-1 (packed Hermitian)       */
};
/* Clover-leaf field strength and the Chroma clover term on checkerboard cb,
 * computed in FP64 from the packed gauge arrays. csw_r / csw_t are the clover
 * coefficients (isotropic: both c_sw), diag_mass = 4 + m. */
int qphix_clover_build(const qphix_geom *g, const void *u_cb0, const void *u_cb1,
                       double csw_r, double csw_t, double diag_mass, int cb, int what,
                       double mu, double eps, void *out);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, loops that rebuild the clover term
 * every molecular-dynamics step, harnesses that forbid hipMalloc): the same
 * build and the same results as qphix_clover_build, with no hipMalloc,
 * hipFree or synchronisation. Its FP64 scratch is one static device array of
 * 32768 sites (63 MiB, held by each device the library runs on until the
 * process ends); lattices of any size are built in chunks of that many sites.
 * qphix_init must have run (QPHIX_EINVAL otherwise). Stream-ordered on the
 * library stream; calls that overlap on one device share the scratch and race. */
int qphix_clover_build_noalloc(const qphix_geom *g, const void *u_cb0, const void *u_cb1,
                               double csw_r, double csw_t, double diag_mass, int cb, int what,
                               double mu, double eps, void *out);

/* ---- Dslash objects: Dslash / ClovDslash / TMDslash / TMClovDslash
 *      constructors (dslash_def.h:27-32, clover_dslash_def.h:25-31,
 *      tm_dslash_def.h:22-29, tm_clov_dslash_def.h:21-30) --------------- */
typedef struct {
  double t_boundary;      /* multiplies the hops across the t boundary       */
  double aniso_s;         /* spatial hop coefficient                          */
  double aniso_t;         /* temporal hop coefficient                         */
  int use_tbc[4];         /* per dimension: apply tbc_phases to every hop     */
  double tbc_phases[4][2];
  double mass;            /* TMDslash: Wilson mass (alpha = 4 + mass)         */
  double twisted_mass;    /* TMDslash: mu                                     */
  double prec_mass_rho;   /* ClovDslash / TMClovDslash: rho                   */
} qphix_dslash_params;
int qphix_dslash_create(const qphix_geom *g, const qphix_dslash_params *p, qphix_dslash **d);
int qphix_dslash_destroy(qphix_dslash *d);

/* Dslash::dslash, ::dslashAChiMinusBDPsi (dslash_def.h:38-51) */
int qphix_wilson_dslash(qphix_dslash *d, void *res, const void *psi, const void *u,
                        int isign, int cb);
int qphix_wilson_achimbdpsi(qphix_dslash *d, void *res, const void *psi, const void *chi,
                            const void *u, double alpha, double beta, int isign, int cb);
/* Dslash::dslashDir (dslash_def.h:54-58, dslash_body.h:1074-1137): one term of
 * the hop with isign = +1; dir = 2 mu takes the forward neighbour psi(x + mu),
 * dir = 2 mu + 1 the backward one */
int qphix_wilson_dslash_dir(qphix_dslash *d, void *res, const void *psi, const void *u, int cb,
                            int dir);
/* ClovDslash::dslash, ::dslashAChiMinusBDPsi, ::clovMult (clover_dslash_def.h:41-78) */
int qphix_clov_dslash(qphix_dslash *d, void *res, const void *psi, const void *u,
                      const void *invclov, int isign, int cb);
int qphix_clov_achimbdpsi(qphix_dslash *d, void *res, const void *psi, const void *chi,
                          const void *u, const void *clov, double beta, int isign, int cb);
int qphix_clov_mult(qphix_dslash *d, void *res, const void *chi, const void *clov);
/* res = A chi for the packed Hermitian clover blocks (full = 0) or the full
 * complex 6x6 blocks (full = 1) of one checkerboard: the site product of
 * clover_product.h:251-339 (there accumulated into res; here assigned) */
int qphix_clover_apply(const qphix_geom *g, void *res, const void *chi, const void *clov, int full);
/* TMDslash::dslash, ::dslashAChiMinusBDPsi, ::two_flav_AChiMinusBDPsi
 * (tm_dslash_def.h:35-58) */
int qphix_tm_dslash(qphix_dslash *d, void *res, const void *psi, const void *u, int isign,
                    int cb);
int qphix_tm_achimbdpsi(qphix_dslash *d, void *res, const void *psi, const void *chi,
                        const void *u, double alpha, double beta, int isign, int cb,
                        double musign);
int qphix_tm_two_flav_achimbdpsi(qphix_dslash *d, void *const res[2],
                                 const void *const psi[2], const void *const chi[2],
                                 const void *u, double alpha, double beta, double epsilon,
                                 int isign, int cb);
/* TMClovDslash::dslash, ::dslashAChiMinusBDPsi, ::two_flav_AChiMinusBDPsi,
 * ::two_flav_inverse_clover_term (tm_clov_dslash_def.h:44-109) */
int qphix_tmclov_dslash(qphix_dslash *d, void *res, const void *psi, const void *u,
                        const void *const invclov[2], int isign, int cb, int fl);
int qphix_tmclov_achimbdpsi(qphix_dslash *d, void *res, const void *psi, const void *chi,
                            const void *u, const void *clov, double beta, int isign, int cb,
                            int fl);
int qphix_tmclov_two_flav_achimbdpsi(qphix_dslash *d, void *const res[2],
                                     const void *const psi[2], const void *const chi[2],
                                     const void *u, const void *clov, double beta,
                                     double epsilon, int isign, int cb);
int qphix_tmclov_two_flav_inverse_clover_term(qphix_dslash *d, void *const res[2],
                                              const void *const psi[2],
                                              const void *const fcl[2],
                                              const void *cl_offdiag, int isign);

/* multi-right-hand-side forms (no reference counterpart): n right-hand sides
 * in launches of an autotuned right-hand-side tile; links and clover are read
 * from HBM once per tile and shared through L2 */
int qphix_wilson_dslash_multi(qphix_dslash *d, void *const res[], const void *const psi[], int n,
                              const void *u, int isign, int cb);
int qphix_wilson_achimbdpsi_multi(qphix_dslash *d, void *const res[], const void *const psi[],
                                  const void *const chi[], int n, const void *u, double alpha,
                                  double beta, int isign, int cb);
int qphix_clov_dslash_multi(qphix_dslash *d, void *const res[], const void *const psi[], int n,
                            const void *u, const void *invclov, int isign, int cb);
int qphix_clov_achimbdpsi_multi(qphix_dslash *d, void *const res[], const void *const psi[],
                                const void *const chi[], int n, const void *u, const void *clov,
                                double beta, int isign, int cb);

/* ---- even-odd operators: EvenOddLinearOperator::operator() (linearOp.h:20-35)
 *      for EvenOddWilsonOperator (wilson.h:49-65), EvenOddCloverOperator
 *      (clover.h:162-172), EvenOddTMWilsonOperator (twisted_mass.h:62-77),
 *      EvenOddTMCloverOperator (twisted_mass_clover.h:92-102),
 *      EvenOddNDTMWilsonReuseOperator (ndtm_reuse_operator.h:65-108),
 *      EvenOddNDTMCloverReuseOperator (ndtm_reuse_operator_clover.h:115-128) */
enum {
  QPHIX_OP_WILSON = 0,
  QPHIX_OP_CLOVER = 1,
  QPHIX_OP_TMWILSON = 2,
  QPHIX_OP_TMCLOVER = 3,
  QPHIX_OP_NDTMWILSON = 4,
  QPHIX_OP_NDTMCLOVER = 5
};
typedef struct {
  int kind;
  double mass;            /* Wilson kinds: alpha = 4 + mass                      */
  double twisted_mass;    /* TM kinds: mu                                        */
  double epsilon;         /* ND kinds: flavour splitting                         */
  double prec_mass_rho;   /* clover kinds: +i rho gamma_5 on the target block;   */
                          /* TM clover kinds pass the twisted mass here          */
  qphix_dslash_params hop;
  const void *u[2];              /* gauge per checkerboard                        */
  const void *clov[2];           /* A per cb (clover kinds; target cb used)       */
  const void *invclov[2];        /* A^-1 per cb (clover; other cb used)           */
  const void *fullclov[2][2];    /* TM clover: [cb][0: +mu, 1: -mu] inverse blocks */
                                 /* ND clover: [cb][0: up, 1: down] inverse blocks */
  const void *clov_odiag[2];     /* ND clover: flavour off-diagonal block per cb   */
} qphix_op_params;
int qphix_op_create(const qphix_geom *g, const qphix_op_params *p, qphix_op **op);
int qphix_op_destroy(qphix_op *op);
int qphix_op_nflav(const qphix_op *op);
/* res[f], in[f] for f < nflav */
int qphix_op_apply(qphix_op *op, void *const res[], const void *const in[], int isign,
                   int target_cb);

/* the operator on n single-flavour right-hand sides (no reference counterpart) */
int qphix_op_apply_multi(qphix_op *op, void *const res[], const void *const in[], int n, int isign,
                         int target_cb);
/* Multigrid-facing pieces (linearOp.h:26-35; wilson.h:67-86; clover.h:174-251).
 * Full spinors are [cb 0, cb 1] pairs. M_offdiag: res = -1/2 D in on target_cb
 * (Wilson, clover). M_diag_inv: clover A_cb^-1 in, Wilson in / (4 + m) (cb 0).
 * M_diag (clover): A_cb in_cb, cb = -1 for both. M_unprec (clover): the full
 * operator A in - 1/2 D in on both checkerboards. DslashDir (Wilson, clover):
 * dslashDir of in_(1-cb) into res_cb for both cb. Other operators:
 * QPHIX_EINVAL ("not yet implemented", as the reference prints). */
int qphix_op_m_offdiag(qphix_op *op, void *const res[], const void *const in[], int isign,
                       int target_cb);
int qphix_op_m_diag_inv(qphix_op *op, void *const res[], const void *const in[], int isign, int cb);
int qphix_op_m_diag(qphix_op *op, void *const res[2], const void *const in[2], int isign, int cb);
int qphix_op_m_unprec(qphix_op *op, void *const res[2], const void *const in[2], int isign);
int qphix_op_dslash_dir(qphix_op *op, void *const res[2], const void *const in[2], int dir);

/* ---- solvers: AbstractSolver::operator() (abs_solver.h:85-97) for InvCG
 *      (invcg.h:25-694), InvBiCGStab (invbicgstab.h:18-619),
 *      InvRichardsonMultiPrec (inv_richardson_multiprec.h:11-320) and
 *      AbstractMultiSolver (abs_solver.h:218-229) for MInvCG (minvcg.h:17-452) */
enum { QPHIX_RESID_ABSOLUTE = 0, QPHIX_RESID_RELATIVE = 1 };
typedef struct {
  int n_iters;
  double rsd_sq_final;      /* as the reference reports it (abs for CG/BiCGStab) */
  unsigned long site_flops;
  unsigned long mv_apps;
  int status;               /* QPHIX_OK, QPHIX_EBREAKDOWN or QPHIX_ENOCONV       */
} qphix_solve_stats;
int qphix_cg_create(qphix_op *M, int max_iters, qphix_solver **s);
int qphix_bicgstab_create(qphix_op *M, int max_iters, qphix_solver **s);
/* outer operator M_outer and inner solver (built on an operator of another
 * geometry / precision); mdagm = 1 solves M^+M x = rhs */
int qphix_richardson_create(qphix_op *M_outer, qphix_solver *inner, double delta,
                            int max_iters, int mdagm, qphix_solver **s);
int qphix_minvcg_create(qphix_op *M, int max_iters, int max_shifts, qphix_solver **s);
/* InvMR / InvMRSmoother (invmr.h:19-348): minimal residual with
 * over-relaxation omega; smoother = 1 runs max_iters iterations without a
 * residual test and reports rsd_sq_final = -1 */
int qphix_mr_create(qphix_op *M, int max_iters, double omega, int smoother, qphix_solver **s);
/* UnprecSolverWrapper (unprec_solver_wrapper.h:16-100) and
 * AbstractUnprecSolver::operator() (abs_solver.h:324-376), single flavour:
 * source preparation on solve_cb, the even-odd solve, and reconstruction of
 * the other checkerboard; x and rhs are [cb 0, cb 1] pairs */
int qphix_unprec_source_prepare(qphix_op *M, void *const out[], const void *const in_cb[],
                                const void *const in_other_cb[], int isign, int solve_cb);
int qphix_unprec_solution_reconstruct(qphix_op *M, void *const out_other_cb[],
                                      const void *const in_other_cb[], const void *const out_cb[],
                                      int isign, int solve_cb);
int qphix_unprec_solve(qphix_solver *eo, void *const x[2], const void *const rhs[2],
                       double rsd_target, int isign, int verbose, int solve_cb, int resid_type,
                       qphix_solve_stats *st);
int qphix_solver_destroy(qphix_solver *s);
int qphix_solve(qphix_solver *s, void *const x[], const void *const rhs[], double rsd_target,
                int isign, int verbose, int target_cb, int resid_type,
                qphix_solve_stats *st);
/* Block BLAS over sets of spinor fields (no reference counterpart; used by the
 * block solver). FP64 or FP32 fields, set sizes 1..32, synchronous.
 * Gram: out[(i * nb + j) * 2 + {0,1}] = sum_sites conj(a_i) . b_j (FP64).
 * Combine: o_j = y_j + sum_i a_i mat_ij, mat row-major na x nb complex
 * (interleaved re, im); y may be NULL (no addition); o_j may alias a_i, y_j. */
int qphix_block_gram(const qphix_geom *g, const void *const a[], int na, const void *const b[], int nb,
                     double *out);
int qphix_block_combine(const qphix_geom *g, const void *const a[], int na, const double *mat,
                        const void *const y[], void *const o[], int nb);
/* Block CG on M^+M (no reference counterpart): the n <= 32 right-hand sides of
 * one qphix_solve_multi call share a block Krylov space; the residual block is
 * kept orthonormal by a Cholesky QR each iteration. FP64 or FP32 fields,
 * single flavour. n_iters is the number of block iterations, rsd_sq_final the
 * relative recursive residual of each right-hand side. */
int qphix_block_cg_create(qphix_op *M, int max_iters, qphix_solver **s);
/* With a CG solver: CG on M^+M for n right-hand sides in lockstep, one launch
 * per step for all unconverged right-hand sides (no reference counterpart; per
 * right-hand side the iteration is that of invcg.h:88-456). With a block CG
 * solver: the block iteration above. st[n] */
int qphix_solve_multi(qphix_solver *s, void *const x[], const void *const rhs[], int n,
                      double rsd_target, int isign, int verbose, int target_cb, int resid_type,
                      qphix_solve_stats *st);
/* How qphix_minvcg_solve runs its iterations (no reference counterpart; the
 * reference keeps the recurrences on the host, minvcg.h:279-384). mode 0: FP64
 * scalars and the zeta / beta recurrences on the host, one host round trip per
 * iteration; 1: all of them in device memory, updated by device kernels
 * between the vector sweeps, no host round trip; 2 (the default): as 1, with
 * one iteration captured in a HIP graph and replayed. All three give
 * bit-identical solutions and counters. */
int qphix_minvcg_set_mode(qphix_solver *s, int mode);
/* The same control for CG and BiCGStab solvers (no reference counterpart;
 * the reference keeps their scalars on the host, invcg.h:293-343,
 * invbicgstab.h:150-256) and for multishift CG (as qphix_minvcg_set_mode). CG
 * and BiCGStab default to 0; modes 1 and 2 apply to single-flavour operators
 * and silent solves (the others run mode 0) and give bit-identical solutions,
 * counters and statuses. */
int qphix_solver_set_mode(qphix_solver *s, int mode);
/* x[shift * nflav + f]; rsd_sq_final[n_shift] */
int qphix_minvcg_solve(qphix_solver *s, void *const x[], const void *const rhs[],
                       int n_shift, const double *shifts, const double *rsd_target,
                       int isign, int verbose, int target_cb, int *n_iters,
                       double *rsd_sq_final, unsigned long *site_flops,
                       unsigned long *mv_apps);

/* ---- batched solves (no reference counterpart): several single-flavour
 *      systems of one operator kind and precision, possibly on lattices of
 *      different sizes, advanced together. Each step is one persistent launch
 *      over all unconverged systems; scalars stay on the device in FP64; with
 *      graph_steps > 0 one CG iteration is captured in a HIP graph and replayed
 *      graph_steps times between host convergence checks. Systems given in
 *      pairs of consecutive entries with the same operator (several
 *      right-hand sides per lattice) run their hops as a register tile of two
 *      sources when the batch has at least 2^18 sites per parity in all
 *      (QPHIX_BATCH_TILE=0 / 1 forces it off / on). */
typedef struct qphix_batch qphix_batch;
int qphix_batch_create(qphix_op *const ops[], int nsys, qphix_batch **b);
int qphix_batch_destroy(qphix_batch *b);
/* CG on M^+M per system (invcg.h iteration); relative targets rsd_target[nsys] */
int qphix_batch_cg(qphix_batch *b, void *const x[], const void *const rhs[], const double rsd_target[],
                   int max_iters, int isign, int target_cb, int graph_steps, qphix_solve_stats st[]);
/* FP64 defect correction (outer[i] on the same lattice as the batch's
 * operator i) around the batched inner CG on M^+M, relaxation delta. The
 * outer loop's work spinors live in the batch object from the first call to
 * qphix_batch_destroy, so calls on one batch must not overlap. */
int qphix_batch_richardson(qphix_batch *inner, qphix_op *const outer[], void *const x[],
                           const void *const rhs[], double rsd_target, double delta, int max_outer,
                           int max_inner, int isign, int target_cb, int graph_steps,
                           qphix_solve_stats st[]);

/* ---- fused BLAS: blas_new_c.h functions with num_flav = 1 (line numbers).
 *      Scalars are doubles, applied in the geometry's arithmetic type;
 *      reductions accumulate in FP64 and return on the host. ------------- */
int qphix_copy_spinor(const qphix_geom *g, void *res, const void *src);                 /* 18 */
int qphix_zero_spinor(const qphix_geom *g, void *res);                                  /* 49 */
int qphix_axy(const qphix_geom *g, double alpha, const void *x, void *y);              /* 70: y = a x */
int qphix_ax(const qphix_geom *g, double alpha, void *x);                              /* 102 */
int qphix_ypeqx(const qphix_geom *g, const void *x, void *y);                          /* 114: y += x */
int qphix_ymeqx(const qphix_geom *g, const void *x, void *y);                          /* 126: y -= x */
int qphix_aypx(const qphix_geom *g, double alpha, const void *x, void *y);             /* 139: y = a y + x */
int qphix_twisted_mass(const qphix_geom *g, const double apimu[2], const void *x, void *y); /* 157 */
int qphix_two_flav_twisted_mass(const qphix_geom *g, const double apimu[2], double epsilon,
                                const void *const x[2], void *const y[2]);               /* 176 */
int qphix_axpy(const qphix_geom *g, double alpha, const void *x, void *y);             /* 214 */
int qphix_axpby(const qphix_geom *g, double alpha, const void *x, double beta, void *y); /* 247 */
int qphix_norm2_spinor(const qphix_geom *g, const void *x, double *n2);                 /* 278 */
int qphix_norm2_multi(const qphix_geom *g, const void *const x[], int n, double *n2);  /* n norms, one launch */
int qphix_axpy_norm2(const qphix_geom *g, double alpha, const void *x, void *y, double *n2); /* 312 */
int qphix_xmy_norm2(const qphix_geom *g, void *res, const void *x, const void *y, double *n2); /* 345: res = x - y */
int qphix_xmy2_norm2(const qphix_geom *g, const void *x, void *y, double *n2);        /* 360: y = y - x */
int qphix_rmammp_norm2_rxpap(const qphix_geom *g, void *r, double a, const void *mmp,
                             void *x, const void *p, double *cp);                        /* 397 */
int qphix_richardson_rxupdate_norm_r(const qphix_geom *g, void *x, void *r, const void *dx,
                                     const void *dr, double *n2);                        /* 435 */
int qphix_bicgstab_xmy(const qphix_geom *g, const void *x, void *y);                   /* 489: y = x - y */
int qphix_inner_product(const qphix_geom *g, const void *l, const void *r, double out[2]); /* 520 */
int qphix_inner_product_norm(const qphix_geom *g, const void *l, const void *r,
                             double out[3]);                                             /* 573 */
int qphix_bicgstab_p_update(const qphix_geom *g, const void *r, void *p, const void *v,
                            const double beta[2], const double omega[2]);               /* 627 */
int qphix_bicgstab_s_update(const qphix_geom *g, const double alpha[2], void *s,
                            const void *v);                                              /* 679 */
int qphix_caxpy(const qphix_geom *g, const double alpha[2], const void *x, void *y);   /* 696 */
int qphix_bicgstab_rxupdate(const qphix_geom *g, void *x, void *r, const void *t,
                            const void *p, const double omega[2], const double alpha[2],
                            double *r_norm);                                             /* 736 */
int qphix_mr_rxupdate(const qphix_geom *g, void *x, void *r, const void *mr,
                      const double a[2]);                                                /* 796 */

/* Reductions into device memory (no reference counterpart: the reference's
 * BLAS return host doubles; these are for callers that capture solver steps
 * in HIP graphs). The same operations and values as norm2Spinor (278),
 * axpyNorm2 (312), xmyNorm2Spinor (345), innerProduct (520) and
 * innerProductNorm (573) above, with the 1, 1, 1, 2 or 3 results written to
 * d_out (device memory) by kernels on the library stream: no copy, no memset,
 * no synchronisation. Calls share the library's partial-sum buffer, so two of
 * them must not run concurrently on different streams. */
int qphix_norm2_spinor_dev(const qphix_geom *g, const void *x, double *d_out);
int qphix_axpy_norm2_dev(const qphix_geom *g, double alpha, const void *x, void *y, double *d_out);
int qphix_xmy_norm2_dev(const qphix_geom *g, void *res, const void *x, const void *y, double *d_out);
int qphix_inner_product_dev(const qphix_geom *g, const void *l, const void *r, double *d_out);
int qphix_inner_product_norm_dev(const qphix_geom *g, const void *l, const void *r, double *d_out);

/* Streaming calibration (no reference counterpart; measurement only): mean
 * time of a kernel that moves a hop kernel's per-site traffic without its
 * stencil, one work-item per site of g reading words_in 16-byte words and
 * writing words_out in the site-blocked field layout with the hop kernels'
 * load / store policies. (48, 6) is the FP32 c12 clover hop (links 24, clover
 * 18, spinor 6 in; 6 out), (60, 6) FP32 c18, (30, 6) the FP32 c12 Wilson hop,
 * (66, 6) the FP32 c12 twisted-mass clover hop (full blocks 36 words),
 * (120, 12) the FP64 c18 clover hop. Allocates and frees its own buffers;
 * synchronous. */
int qphix_stream_calibrate(const qphix_geom *g, int words_in, int words_out, int reps, double *ms);

/* ---- Chroma LinOpSysSolver surface (README.markdown:163-256): the
 *      <InvertParam> group of QPHIX_CLOVER_INVERTER and
 *      QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER mapped onto this library.
 *      The clover term follows tests/clover_term_qdp_w.h:156-180:
 *        spatial hop coefficient  aniso_s = nu / xi_0 (anisoP), else 1
 *        temporal hop coefficient aniso_t = 1, t_boundary = -1 if AntiPeriodicT
 *        csw_r = clovCoeffR / xi_0 (anisoP), csw_t = clovCoeffT
 *        diag_mass = 1 + 3 aniso_s + Mass, Mass = 1/(2 Kappa) - 4 from Kappa
 *      and the system solved is the even-odd clover operator on the odd
 *      checkerboard, A_oo - 1/4 D_oe A_ee^-1 D_eo (clover.h:162-172). */
enum { QPHIX_INV_CLOVER = 0, QPHIX_INV_CLOVER_ITER_REFINE = 1 };
enum { QPHIX_SOLVER_CG = 0, QPHIX_SOLVER_BICGSTAB = 1 };
typedef struct {
  int inv_type;              /* <invType>                                         */
  int solver_type;           /* <SolverType>: CG solves M^+M psi = M^+ chi          */
  int max_iter;              /* <MaxIter> (outer and inner for iterative refinement) */
  double rsd_target;         /* <RsdTarget>, relative                             */
  double delta;              /* <Delta>: inner residual reduction (refinement)    */
  double mass;               /* <Mass>, or from <Kappa>                           */
  double clov_coeff_r;       /* <clovCoeffR> (or <clovCoeff>)                     */
  double clov_coeff_t;       /* <clovCoeffT> (or <clovCoeff>)                     */
  int aniso_p;               /* <AnisoParam><anisoP>                              */
  int t_dir;                 /* <t_dir>: 3 is the only direction supported        */
  double xi_0, nu;           /* <xi_0>, <nu>                                      */
  int antiperiodic_t;        /* <AntiPeriodicT>                                   */
  int verbose;               /* <Verbose>: one summary line per solve on stderr   */
  int tune;                  /* <Tune>: autotune the operator's kernels at create */
  double rsd_tolerance_factor; /* <RsdToleranceFactor>: a solve whose true relative */
                               /* residual exceeds RsdTarget times this fails       */
  /* <NCores> <ThreadsPerCore> <By> <Bz> <Sy> <Sz> <PadXY> <PadXYZ> <MinCt>:
   * host threading and cache blocking of the reference; read and checked,
   * without effect here (the hop kernels are autotuned instead) */
  int ncores, threads_per_core, by, bz, sy, sz, pad_xy, pad_xyz, min_ct;
  /* precisions, configure options of the reference build
   * (--enable-qphix-solver-compress12, --enable-qphix-solver-inner-type);
   * XML: <OuterPrecision> d|f, <InnerPrecision> h|f|d, <Compress12> (both),
   * <InnerCompress12> (inner only; half precision needs it) */
  int prec, compress12, inner_prec, inner_compress12;
} qphix_invert_param;
/* defaults: BiCGStab, MaxIter 1000, RsdTarget 1e-7, Delta 0.1, isotropic,
 * c_sw 1, periodic t, RsdToleranceFactor 5, outer F64, inner F32, no
 * compression */
int qphix_invert_param_default(qphix_invert_param *p);
/* reads the first <InvertParam> element of an XML document into p (starting
 * from the defaults). Malformed XML and missing or unreadable tags:
 * QPHIX_EINVAL; unsupported values (invType, t_dir, both Mass and Kappa, ...):
 * QPHIX_EDOMAIN. qphix_last_error names the tag and line. */
int qphix_invert_param_read_xml(const char *xml, qphix_invert_param *p);
/* the coefficients above: out = {aniso_s, aniso_t, t_boundary, csw_r, csw_t,
 * diag_mass} (host only) */
int qphix_invert_param_coeffs(const qphix_invert_param *p, double out[6]);

typedef struct qphix_syssolver qphix_syssolver;
typedef struct {
  int n_count;               /* solver iterations (outer ones for refinement)     */
  double resid;              /* true |chi - M psi| on the odd checkerboard        */
  double rel_resid;          /* resid / |chi|                                     */
  unsigned long mv_apps;     /* operator applications incl. the residual check    */
  double seconds;            /* wall time of the solve                            */
  int status;                /* QPHIX_OK, QPHIX_ENOCONV, QPHIX_EBREAKDOWN         */
} qphix_syssolver_results;
/* builds the gauge field (canonical host layout above), the clover term A on
 * the odd and A^-1 on the even checkerboard in FP64 on the device, the
 * operators and the solver(s) */
int qphix_syssolver_create(const int latt[4], const qphix_invert_param *p, const double *u,
                           qphix_syssolver **s);
int qphix_syssolver_destroy(qphix_syssolver *s);
/* psi, chi: canonical full-lattice spinors (double[2*Vh][4][3][2]); the odd
 * half of psi is the initial guess and receives the solution, the even half is
 * not touched. The call returns QPHIX_OK when the solve ran; r->status tells
 * whether it met RsdTarget * RsdToleranceFactor. */
int qphix_syssolver_solve(qphix_syssolver *s, double *psi, const double *chi,
                          qphix_syssolver_results *r);
/* the same on device spinors of the outer geometry (odd checkerboard) */
int qphix_syssolver_solve_dev(qphix_syssolver *s, void *psi, const void *chi,
                              qphix_syssolver_results *r);
int qphix_syssolver_geom(const qphix_syssolver *s, const qphix_geom **g);

#ifdef __cplusplus
}
#endif
#endif /* QPHIX_AMD_H */
