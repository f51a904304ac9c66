"""Even-odd operators and Krylov solvers on the GPU, checked against the CPU
oracle: operators to the Dslash tolerances, solvers by their true residual
recomputed in FP64 with the oracle operator, and by the solution of the
oracle's own solver on the same system (tests/testDslashFull.cc:445-932,
tests/testMInvCG.cc:63-190 do the same against QDP++)."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L8 = [8, 8, 8, 8]
VH = 2048
MU, EPS, MASS = 0.07, 0.03, 0.1


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


class World:
    """Fields of one 8^4 configuration in the oracle and on the device."""

    def __init__(self):
        self.u = oracle.gen_gauge(L8, 4, 0.4)
        clov = oracle.clover_build(L8, self.u, 1.0, MASS)
        self.clov = [clov[:VH * 72], clov[VH * 72:]]
        self.inv = [oracle.clover_invert(c) for c in self.clov]
        self.fcl = [[oracle.full_clover_tm_inverse(c, MU), oracle.full_clover_tm_inverse(c, -MU)]
                    for c in self.clov]
        self.nd = [oracle.ndclover_inverse(c, MU, EPS) for c in self.clov]
        hp = oracle.hop_params(-1.0, 1.0, 1.0)
        self.orc = {
            "wilson": oracle.Operator(oracle.WILSON, L8, self.u, hp, mass=MASS),
            "clover": oracle.Operator(oracle.CLOVER, L8, self.u, hp, clov=self.clov, invclov=self.inv,
                                      rho=0.01),
            "tmwilson": oracle.Operator(oracle.TMWILSON, L8, self.u, hp, mass=MASS, mu=MU),
            "tmclover": oracle.Operator(oracle.TMCLOVER, L8, self.u, hp, clov=self.clov,
                                        fcl=self.fcl, mu=MU, rho=MU),
            "ndtmwilson": oracle.Operator(oracle.NDTMWILSON, L8, self.u, hp, mass=MASS, mu=MU,
                                          eps=EPS),
            "ndtmclover": oracle.Operator(oracle.NDTMCLOVER, L8, self.u, hp, clov=self.clov, mu=MU,
                                          eps=EPS, rho=MU, ndfcl=[[n[0], n[1]] for n in self.nd],
                                          ndcoff=[n[2] for n in self.nd]),
        }
        self.dev = {}

    def device(self, prec, c12):
        key = (prec, c12)
        if key in self.dev:
            return self.dev[key]
        g = q.Geometry(L8, prec, c12)
        ud = q.pack_gauge(g, self.u)
        cl = [g.clover().upload(c) for c in self.clov]
        inv = [g.clover().upload(c) for c in self.inv]
        fcl = [[g.full_clover().upload(f) for f in pair] for pair in self.fcl]
        ndf = [[g.full_clover().upload(n[0]), g.full_clover().upload(n[1])] for n in self.nd]
        ndo = [g.clover().upload(n[2]) for n in self.nd]
        kw = dict(t_boundary=-1.0)
        ops = {
            "wilson": q.Operator(g, q.OP_WILSON, ud, mass=MASS, **kw),
            "clover": q.Operator(g, q.OP_CLOVER, ud, clov=cl, invclov=inv, prec_mass_rho=0.01, **kw),
            "tmwilson": q.Operator(g, q.OP_TMWILSON, ud, mass=MASS, twisted_mass=MU, **kw),
            "tmclover": q.Operator(g, q.OP_TMCLOVER, ud, clov=cl, fullclov=fcl, twisted_mass=MU,
                                   prec_mass_rho=MU, **kw),
            "ndtmwilson": q.Operator(g, q.OP_NDTMWILSON, ud, mass=MASS, twisted_mass=MU,
                                     epsilon=EPS, **kw),
            "ndtmclover": q.Operator(g, q.OP_NDTMCLOVER, ud, clov=cl, fullclov=ndf, clov_odiag=ndo,
                                     twisted_mass=MU, epsilon=EPS, **kw),
        }
        self.dev[key] = (g, ops, (ud, cl, inv, fcl, ndf, ndo))
        return self.dev[key]


@pytest.fixture(scope="module")
def world():
    return World()


def flavours(g, nf, seed, cb):
    hs = [oracle.gen_spinor(L8, seed + f, cb) for f in range(nf)]
    ds = [g.spinor().upload(h, cb) for h in hs]
    return ds, [d.download(cb) for d in ds]


@pytest.mark.parametrize("name", ["wilson", "clover", "tmwilson", "tmclover", "ndtmwilson",
                                  "ndtmclover"])
@pytest.mark.parametrize("prec,c12,tol", [(q.F64, False, 1e-13), (q.F32, True, 2e-6)],
                         ids=["f64", "f32c12"])
def test_operator(world, name, prec, c12, tol):
    g, ops, _ = world.device(prec, c12)
    M, Mo = ops[name], world.orc[name]
    nf = M.nflav
    for isign in (1, -1):
        for cb in (1, 0):
            x, xh = flavours(g, nf, 90, cb)
            res = [g.spinor() for _ in range(nf)]
            M(res, x, isign, cb)
            got = np.concatenate([r.download(cb) for r in res])
            ref = Mo(np.concatenate(xh), isign, cb)
            assert rel(got, ref) < tol, (name, isign, cb)


def test_cg_wilson_f64(world):
    """Config: 8^4 Wilson, FP64, antiperiodic t, CG on M^+M to 1e-12."""
    g, ops, _ = world.device(q.F64, False)
    b, bh = flavours(g, 1, 100, 1)
    x = g.spinor()
    st = q.InvCG(ops["wilson"], 1000)(x, b, 1e-12)
    assert st.status == q.OK
    Mo = world.orc["wilson"]
    xh = x.download(1)
    r = bh[0] - Mo(Mo(xh, 1, 1), -1, 1)
    assert np.linalg.norm(r) / np.linalg.norm(bh[0]) < 1.5e-12
    xo, sto = Mo.cg(bh[0], 1e-12, 1000)
    assert abs(st.n_iters - sto.n_iters) <= 1
    assert rel(xh, xo) < 1e-10


@pytest.mark.parametrize("name", ["clover", "tmclover"])
def test_bicgstab_f64(world, name):
    g, ops, _ = world.device(q.F64, False)
    b, bh = flavours(g, 1, 110, 1)
    x = g.spinor()
    st = q.InvBiCGStab(ops[name], 1000)(x, b, 1e-11)
    assert st.status == q.OK
    Mo = world.orc[name]
    xh = x.download(1)
    assert np.linalg.norm(bh[0] - Mo(xh, 1, 1)) / np.linalg.norm(bh[0]) < 1.5e-11
    xo, sto = Mo.bicgstab(bh[0], 1e-11, 1000)
    assert abs(st.n_iters - sto.n_iters) <= 2
    assert rel(xh, xo) < 1e-9


def test_bicgstab_clover_f32(world):
    """FP32 fields, FP64 reductions, true residual recomputed in FP64."""
    g, ops, _ = world.device(q.F32, True)
    b, bh = flavours(g, 1, 120, 1)
    x = g.spinor()
    st = q.InvBiCGStab(ops["clover"], 1000)(x, b, 1e-6)
    assert st.status == q.OK
    Mo = world.orc["clover"]
    assert np.linalg.norm(bh[0] - Mo(x.download(1), 1, 1)) / np.linalg.norm(bh[0]) < 5e-6


def test_cg_nd_clover_f64(world):
    g, ops, _ = world.device(q.F64, False)
    b, bh = flavours(g, 2, 130, 1)
    x = [g.spinor(), g.spinor()]
    st = q.InvCG(ops["ndtmclover"], 2000)(x, b, 1e-10)
    assert st.status == q.OK
    Mo = world.orc["ndtmclover"]
    xh = np.concatenate([v.download(1) for v in x])
    B = np.concatenate(bh)
    assert np.linalg.norm(B - Mo(Mo(xh, 1, 1), -1, 1)) / np.linalg.norm(B) < 1.5e-10


def test_minvcg_f64(world):
    g, ops, _ = world.device(q.F64, False)
    b, bh = flavours(g, 1, 140, 1)
    shifts = [1e-4, 1e-3, 1e-2, 0.1, 1.0]
    xs = [[g.spinor()] for _ in shifts]
    nit, fin, mv = q.MInvCG(ops["clover"], 2000, 8)(xs, b, shifts, [1e-10] * len(shifts))
    assert nit > 0
    Mo = world.orc["clover"]
    xo, nito, _ = Mo.minvcg(bh[0], shifts, [1e-10] * len(shifts), 2000)
    assert abs(nit - nito) <= 1
    for k, s in enumerate(shifts):
        xh = xs[k][0].download(1)
        r = bh[0] - Mo(Mo(xh, 1, 1), -1, 1) - s * xh
        assert np.linalg.norm(r) / np.linalg.norm(bh[0]) < 2e-10, (k, s)


def test_minvcg_too_many_shifts(world):
    g, ops, _ = world.device(q.F64, False)
    b, _ = flavours(g, 1, 141, 1)
    with pytest.raises(q.QphixError) as e:
        q.MInvCG(ops["wilson"], 100, 2)([[g.spinor()]] * 3, b, [0.1, 0.2, 0.3], [1e-8] * 3)
    assert e.value.code == q.EDOMAIN


@pytest.mark.parametrize("inner_prec", [q.F32, q.F16, q.BF16], ids=["f32", "f16", "bf16"])
def test_richardson_mixed_clover(world, inner_prec):
    """FP64 defect correction around a low-precision BiCGStab (config 3 at 8^4)."""
    g, ops, _ = world.device(q.F64, False)
    gi, opsi, _ = world.device(inner_prec, True)
    inner = q.InvBiCGStab(opsi["clover"], 500)
    solver = q.InvRichardsonMultiPrec(ops["clover"], inner, 1e-2 if inner_prec in (q.F16, q.BF16) else 1e-4,
                                      200)
    b, bh = flavours(g, 1, 150, 1)
    x = g.spinor()
    st = solver(x, b, 1e-10)
    assert st.status == q.OK
    Mo = world.orc["clover"]
    xh = x.download(1)
    assert np.linalg.norm(bh[0] - Mo(xh, 1, 1)) / np.linalg.norm(bh[0]) < 1.5e-10
    xo, _ = Mo.bicgstab(bh[0], 1e-12, 1000)
    assert rel(xh, xo) < 1e-8


def test_richardson_mdagm_cg_f32(world):
    g, ops, _ = world.device(q.F64, False)
    gi, opsi, _ = world.device(q.F32, True)
    solver = q.InvRichardsonMultiPrec(ops["wilson"], q.InvCG(opsi["wilson"], 1000), 1e-4, 100,
                                      mdagm=True)
    b, bh = flavours(g, 1, 160, 1)
    x = g.spinor()
    st = solver(x, b, 1e-11)
    assert st.status == q.OK
    Mo = world.orc["wilson"]
    xh = x.download(1)
    assert np.linalg.norm(bh[0] - Mo(Mo(xh, 1, 1), -1, 1)) / np.linalg.norm(bh[0]) < 1.5e-11


def test_bicgstab_breakdown_zero_rhs(world):
    """rhs = 0 with x0 = 0 converges at iteration 0 (invbicgstab.h:128-134)."""
    g, ops, _ = world.device(q.F64, False)
    st = q.InvBiCGStab(ops["clover"], 10)(g.spinor(), g.spinor(), 1e-8,
                                          resid=q.RESID_ABSOLUTE)
    assert st.n_iters == 0 and st.status == q.OK


def test_not_converged_status(world):
    g, ops, _ = world.device(q.F64, False)
    b, _ = flavours(g, 1, 170, 1)
    st = q.InvCG(ops["wilson"], 3)(g.spinor(), b, 1e-14)
    assert st.status == q.ENOCONV and st.n_iters == 3
