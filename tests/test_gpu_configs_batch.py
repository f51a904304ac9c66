"""The batched-source and batched-system workloads at their stated sizes.

* 32^3x64 clover (c_sw = 1, kappa = 0.1240, t antiperiodic): multishift CG in
  FP64 with 12 shifts sigma_k = 10^(-4 + 4k/11) to 1e-8 each, every shift's
  true residual |b - (M^+M + sigma) x| / |b| recomputed with the FP64
  operator, the device-scalar (graph) and host-scalar loops bit-identical;
  block CG and lockstep CG on M^+M in FP32 for 12 point sources (one per spin
  and colour at a seeded site) and 4 Gaussian sources to 1e-6, true residuals
  recomputed in FP64. (The FP64 operators are pinned to the oracle at 8^4.)
* The uneven twisted-mass clover batch on 8^4, 12^3x24, 16^3x32 and 24^3x48
  (own gauge seeds, mu ~ U[0.001, 0.01], kappa = 0.1240): FP16-storage inner
  CG on M^+M in batched, graph-replayed launches inside FP64 refinement to
  1e-10, true residuals recomputed with the FP64 operators.
* The non-degenerate twisted-mass clover doublet operator on 16^3x32 (mu =
  0.01, eps = 0.005) against the oracle in FP64 (1e-13) and FP32 (2e-6).
"""
import ctypes as C
import random

import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

MASS = 1.0 / (2 * 0.1240) - 4.0


def rel_dev(g, a, b):
    d = g.spinor()
    return (q.xmy_norm2(g, d, a, b) / q.norm2(g, b)) ** 0.5


class Clover32:
    L = [32, 32, 32, 64]

    def __init__(self):
        self.g64 = q.Geometry(self.L, q.F64, False)
        self.u64 = q.gen_gauge(self.g64, 2, 0.4)
        self.A = q.clover_build(self.g64, self.u64, 1.0, MASS, 1, q.CLOV_A)
        self.Ai = q.clover_build(self.g64, self.u64, 1.0, MASS, 0, q.CLOV_AINV)
        self.M64 = q.Operator(self.g64, q.OP_CLOVER, self.u64, clov=[None, self.A], invclov=[self.Ai, None],
                              t_boundary=-1.0)


@pytest.fixture(scope="module")
def c32():
    return Clover32()


def mdagm(M, g, x, sigma=0.0):
    t1, t2 = g.spinor(), g.spinor()
    M(t1, x, 1, 1)
    M(t2, t1, -1, 1)
    if sigma:
        q.call("qphix_axpy", g.h, C.c_double(sigma), x.p, t2.p)
    return t2


def test_multishift_32x32x32x64_12_shifts(c32):
    g = c32.g64
    shifts = [10 ** (-4 + 4 * k / 11) for k in range(12)]
    b = q.gen_spinor(g, 500 + 12, 1)
    sols = {}
    for mode in (0, 2):
        xs = [g.spinor() for _ in shifts]
        nit, fin, mv = q.MInvCG(c32.M64, 5000, len(shifts), mode=mode)(xs, b, shifts, [1e-8] * len(shifts))
        sols[mode] = (nit, fin, mv, [x.download(1) for x in xs], xs)
    assert sols[0][:3] == sols[2][:3]
    for a, c in zip(sols[0][3], sols[2][3]):
        assert np.array_equal(a, c)
    for sg, x in zip(shifts, sols[2][4]):
        r = rel_dev(g, mdagm(c32.M64, g, x, sg), b)
        assert r < 1.5e-8, (sg, r)


def test_block_and_lockstep_cg_32x32x32x64(c32):
    g64 = c32.g64
    g = q.Geometry(c32.L, q.F32, True)
    u = q.convert_gauge(g, g64, c32.u64)
    M32 = q.Operator(g, q.OP_CLOVER, u, clov=[None, q.convert_clover(g, g64, c32.A)],
                     invclov=[q.convert_clover(g, g64, c32.Ai), None], t_boundary=-1.0)
    site = random.Random(12).randrange(g.vh)
    e = np.zeros(g.vh * 24)
    rhs64 = []
    for k in range(12):
        e[site * 24 + 2 * k] = 1.0
        rhs64.append(g64.spinor().upload(e, 1))
        e[site * 24 + 2 * k] = 0.0
    rhs64 += [q.gen_spinor(g64, 600 + k, 1) for k in range(4)]
    rhs = []
    for b in rhs64:
        f = g.spinor()
        q.convert_spinor(g, f, 1.0, g64, b)
        rhs.append(f)
    for solver in (q.InvBlockCG(M32, 3000), q.InvCG(M32, 5000)):
        xs = [g.spinor() for _ in rhs]
        sts = solver.solve_multi(xs, rhs, 1e-6)
        assert all(st.status == q.OK for st in sts)
        for x, b in zip(xs, rhs64):
            x64 = g64.spinor()
            q.convert_spinor(g64, x64, 1.0, g, x)
            assert rel_dev(g64, mdagm(c32.M64, g64, x64), b) < 2e-6


def test_uneven_batch_tmclover_config_lattices():
    lats = [[8, 8, 8, 8], [12, 12, 12, 24], [16, 16, 16, 32], [24, 24, 24, 48]]
    rng = random.Random(5)
    keep, ops64, ops16, gs, rhs = [], [], [], [], []
    for i, L in enumerate(lats):
        mu = rng.uniform(0.001, 0.01)
        g64, g16 = q.Geometry(L, q.F64, False), q.Geometry(L, q.F16, True)
        u = q.gen_gauge(g64, 5 + i, 0.4)
        A = q.clover_build(g64, u, 1.0, MASS, 1, q.CLOV_A)
        fp = q.clover_build(g64, u, 1.0, MASS, 0, q.CLOV_TMINV, mu)
        fm = q.clover_build(g64, u, 1.0, MASS, 0, q.CLOV_TMINV, -mu)
        u16 = q.convert_gauge(g16, g64, u)
        A16 = q.convert_clover(g16, g64, A)
        fp16, fm16 = q.convert_clover(g16, g64, fp, full=True), q.convert_clover(g16, g64, fm, full=True)
        kw = dict(twisted_mass=mu, prec_mass_rho=mu, t_boundary=-1.0)
        M64 = q.Operator(g64, q.OP_TMCLOVER, u, clov=[None, A], fullclov=[[fp, fm], None], **kw)
        M16 = q.Operator(g16, q.OP_TMCLOVER, u16, clov=[None, A16], fullclov=[[fp16, fm16], None], **kw)
        keep += [g16, u, A, fp, fm, u16, A16, fp16, fm16]
        for r in range(2):
            ops64.append(M64)
            ops16.append(M16)
            gs.append(g64)
            rhs.append(q.gen_spinor(g64, 100 * i + r, 1))
    x = [g.spinor() for g in gs]
    sts = q.Batch(ops16).richardson(ops64, x, rhs, 1e-10, delta=1e-2, graph_steps=8)
    for i, st in enumerate(sts):
        assert st.status == q.OK, (i, st.n_iters)
        assert rel_dev(gs[i], mdagm(ops64[i], gs[i], x[i]), rhs[i]) < 2e-10, i


@pytest.mark.parametrize("prec,c12,tol", [(q.F64, False, 1e-13), (q.F32, True, 2e-6)], ids=["f64", "f32c12"])
def test_doublet_tmclover_16x16x16x32(prec, c12, tol):
    L = [16, 16, 16, 32]
    mu, eps = 0.01, 0.005
    vh = L[0] * L[1] * L[2] * L[3] // 2
    u = oracle.gen_gauge(L, 9, 0.4)
    clov = oracle.clover_build(L, u, 1.0, MASS)
    cl = [clov[:vh * 72], clov[vh * 72:]]
    nd = [oracle.ndclover_inverse(c, mu, eps) for c in cl]
    Mo = oracle.Operator(oracle.NDTMCLOVER, L, u, oracle.hop_params(-1.0), clov=cl, ndfcl=[n[:2] for n in nd],
                         ndcoff=[n[2] for n in nd], mu=mu, eps=eps, rho=mu)
    g = q.Geometry(L, prec, c12)
    ud = q.pack_gauge(g, u)
    cld = [g.clover().upload(c) for c in cl]
    ndf = [[g.full_clover().upload(n[0]), g.full_clover().upload(n[1])] for n in nd]
    ndo = [g.clover().upload(n[2]) for n in nd]
    M = q.Operator(g, q.OP_NDTMCLOVER, ud, clov=cld, fullclov=ndf, clov_odiag=ndo, twisted_mass=mu,
                   epsilon=eps, t_boundary=-1.0)
    for isign in (1, -1):
        xh = [oracle.gen_spinor(L, 20 + f, 1) for f in range(2)]
        xd = [g.spinor().upload(h, 1) for h in xh]
        y = [g.spinor(), g.spinor()]
        M(y, xd, isign, 1)
        got = np.concatenate([v.download(1) for v in y])
        ref = Mo(np.concatenate([v.download(1) for v in xd]), isign, 1)
        assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < tol, isign
