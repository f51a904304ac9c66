"""Two workloads at their stated parameters.

* Plain Wilson on 8^4, FP64, kappa = 0.12 (m = 1/(2 kappa) - 4), t
  antiperiodic, gauge seed 1: the Dslash on both checkerboards with isign
  +-1 and the even-odd operator against the oracle to 1e-13, CG on M^+M to
  1e-12 (true residual from the oracle operator, iteration count within one of
  the oracle's CG, the oracle's solution).
* Clover on 32^3x64, FP32 fields with FP64 reductions, c_sw = 1, kappa =
  0.1240, t antiperiodic: the FP32 clover Dslash A^-1 D psi (two-row and full
  links) against the FP64 kernel on the same seeded fields (the FP64 kernel is
  pinned to the oracle by the 8^4 tests), both checkerboards and signs; and
  BiCGStab in FP32 on the even-odd clover operator to 1e-7, its true residual
  recomputed with the FP64 operator.
"""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_wilson_8x8x8x8_kappa_0p12():
    L = [8, 8, 8, 8]
    kappa = 0.12
    mass = 1.0 / (2 * kappa) - 4.0
    u = oracle.gen_gauge(L, 1, 0.4)
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    g = q.Geometry(L, q.F64, False)
    ud = q.pack_gauge(g, u)
    D = q.Dslash(g, t_boundary=-1.0)
    for cb in (0, 1):
        psih = oracle.gen_spinor(L, 2 + cb, 1 - cb)
        psi = g.spinor().upload(psih, 1 - cb)
        for isign in (1, -1):
            out = g.spinor()
            D.dslash(out, psi, ud[cb], isign, cb)
            ref = oracle.hop(L, u, psih, isign, cb, hp)
            assert rel(out.download(cb), ref) < 1e-13, (cb, isign)
    Mo = oracle.Operator(oracle.WILSON, L, u, hp, mass=mass)
    M = q.Operator(g, q.OP_WILSON, ud, mass=mass, t_boundary=-1.0)
    bh = oracle.gen_spinor(L, 5, 1)
    b = g.spinor().upload(bh, 1)
    for isign in (1, -1):
        o = g.spinor()
        M(o, b, isign, 1)
        assert rel(o.download(1), Mo(bh, isign, 1)) < 1e-13
    x = g.spinor()
    st = q.InvCG(M, 1000)(x, b, 1e-12)
    assert st.status == q.OK
    xh = x.download(1)
    r = bh - Mo(Mo(xh, 1, 1), -1, 1)
    assert np.linalg.norm(r) / np.linalg.norm(bh) < 1.5e-12
    xo, sto = Mo.cg(bh, 1e-12, 1000)
    assert abs(st.n_iters - sto.n_iters) <= 1
    assert rel(xh, xo) < 1e-10


class C32:
    L = [32, 32, 32, 64]
    mass = 1.0 / (2 * 0.1240) - 4.0

    def __init__(self):
        self.g64 = q.Geometry(self.L, q.F64, False)
        self.u64 = q.gen_gauge(self.g64, 2, 0.4)
        self.a64 = [q.clover_build(self.g64, self.u64, 1.0, self.mass, cb, q.CLOV_A) for cb in (0, 1)]
        self.ai64 = [q.clover_build(self.g64, self.u64, 1.0, self.mass, cb, q.CLOV_AINV) for cb in (0, 1)]


@pytest.fixture(scope="module")
def c32():
    return C32()


def rel_dev(g64, a, b):
    d = g64.spinor()
    return (q.xmy_norm2(g64, d, a, b) / q.norm2(g64, b)) ** 0.5


@pytest.mark.parametrize("c12", [True, False], ids=["c12", "c18"])
def test_clover_32x32x32x64_f32_dslash(c32, c12):
    g64 = c32.g64
    g = q.Geometry(c32.L, q.F32, c12)
    u = q.convert_gauge(g, g64, c32.u64)
    ai = [q.convert_clover(g, g64, c) for c in c32.ai64]
    D64, D32 = q.Dslash(g64, t_boundary=-1.0), q.Dslash(g, t_boundary=-1.0)
    for cb in (0, 1):
        p64 = q.gen_spinor(g64, 60 + cb, 1 - cb)
        p32 = g.spinor()
        q.convert_spinor(g, p32, 1.0, g64, p64)
        for isign in (1, -1):
            o64, o32 = g64.spinor(), g.spinor()
            D64.clov_dslash(o64, p64, c32.u64[cb], c32.ai64[cb], isign, cb)
            D32.clov_dslash(o32, p32, u[cb], ai[cb], isign, cb)
            back = g64.spinor()
            q.convert_spinor(g64, back, 1.0, g, o32)
            r = rel_dev(g64, back, o64)
            assert r < 2e-6, (cb, isign, r)


def test_clover_32x32x32x64_bicgstab_f32(c32):
    g64 = c32.g64
    g = q.Geometry(c32.L, q.F32, True)
    u = q.convert_gauge(g, g64, c32.u64)
    A, Ai = q.convert_clover(g, g64, c32.a64[1]), q.convert_clover(g, g64, c32.ai64[0])
    M32 = q.Operator(g, q.OP_CLOVER, u, clov=[None, A], invclov=[Ai, None], t_boundary=-1.0)
    M64 = q.Operator(g64, q.OP_CLOVER, c32.u64, clov=[None, c32.a64[1]], invclov=[c32.ai64[0], None],
                     t_boundary=-1.0)
    b64 = q.gen_spinor(g64, 8, 1)
    b = g.spinor()
    q.convert_spinor(g, b, 1.0, g64, b64)
    x = g.spinor()
    st = q.InvBiCGStab(M32, 5000)(x, b, 1e-7)
    assert st.status == q.OK
    x64, r = g64.spinor(), g64.spinor()
    q.convert_spinor(g64, x64, 1.0, g, x)
    M64(r, x64, 1, 1)
    tr = rel_dev(g64, r, b64)
    print(f"BiCGStab FP32 32^3x64: {st.n_iters} iterations, true residual (FP64) {tr:.3e}")
    assert tr < 3e-7, tr
