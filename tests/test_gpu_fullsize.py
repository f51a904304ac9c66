"""The hot path at the headline lattice, 48^3 x 96 (5,308,416 sites per
parity), through properties that do not depend on the lattice size:

* FP32 (c12 and c18) clover Dslash A^-1 D psi against the FP64 kernel, which
  the 8^4 tests pin to the oracle, on the same seeded fields;
* the free field: unit links and a plane wave e^{ip.x} phi (antiperiodic
  momentum in t) give D psi = e^{ip.x} sum_mu 2 c_mu (cos p_mu -+ i gamma_mu
  sin p_mu) phi, and A = 4 + m, in closed form;
* M(isign = -1) is the adjoint of M(isign = +1) for the even-odd clover
  operator;
* the register-tiled multi-source kernel is bit-identical to the single-source
  one;
* the mixed-precision solve (FP16-storage BiCGStab inside FP64 defect
  correction, kappa = 0.1245) reaches a true FP64 residual of 1e-10 and the
  FP64 BiCGStab solution.
Differences are formed on the device (FP64 reductions)."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L = [48, 48, 48, 96]
MASS = 1 / (2 * 0.1245) - 4
CSW = 1.0


class Full:
    def __init__(self):
        self.g64 = q.Geometry(L, q.F64, False)
        self.g32 = q.Geometry(L, q.F32, True)
        self.g32c18 = q.Geometry(L, q.F32, False)
        self.g16 = q.Geometry(L, q.F16, True)
        g = self.g64
        self.u64 = q.gen_gauge(g, 3, 0.4)
        self.a64 = [q.clover_build(g, self.u64, CSW, MASS, cb, q.CLOV_A) for cb in (0, 1)]
        self.ai64 = [q.clover_build(g, self.u64, CSW, MASS, cb, q.CLOV_AINV) for cb in (0, 1)]
        self.lo = {}
        for gl in (self.g32, self.g32c18, self.g16):
            u = q.convert_gauge(gl, g, self.u64)
            a = [q.convert_clover(gl, g, c) for c in self.a64]
            ai = [q.convert_clover(gl, g, c) for c in self.ai64]
            self.lo[id(gl)] = (u, a, ai)

    def fields(self, gl):
        if gl is self.g64:
            return self.u64, self.a64, self.ai64
        return self.lo[id(gl)]


@pytest.fixture(scope="module")
def full():
    return Full()


def rel_dev(g64, a64, b64):
    """|a - b| / |b| of two FP64 device spinors."""
    d = g64.spinor()
    return (q.xmy_norm2(g64, d, a64, b64) / q.norm2(g64, b64)) ** 0.5


def to64(full, gl, x):
    y = full.g64.spinor()
    q.convert_spinor(full.g64, y, 1.0, gl, x)
    return y


@pytest.mark.parametrize("which", ["f32c12", "f32c18"])
def test_f32_clover_dslash_matches_f64(full, which):
    g64 = full.g64
    gl = full.g32 if which == "f32c12" else full.g32c18
    u, _, ai = full.fields(gl)
    D64 = q.Dslash(g64, t_boundary=-1.0)
    Dl = q.Dslash(gl, t_boundary=-1.0)
    for cb in (0, 1):
        psi64 = q.gen_spinor(g64, 40 + cb, 1 - cb)
        psil = gl.spinor()
        q.convert_spinor(gl, psil, 1.0, g64, psi64)
        for isign in (1, -1):
            o64, ol = g64.spinor(), gl.spinor()
            D64.clov_dslash(o64, psi64, full.u64[cb], full.ai64[cb], isign, cb)
            Dl.clov_dslash(ol, psil, u[cb], ai[cb], isign, cb)
            r = rel_dev(g64, to64(full, gl, ol), o64)
            assert r < 2e-6, (cb, isign, r)


def site_coords(cb):
    nxh = L[0] // 2
    s = np.arange(L[0] * L[1] * L[2] * L[3] // 2, dtype=np.int64)
    xh = s % nxh
    r = s // nxh
    y = r % L[1]
    r //= L[1]
    z = r % L[2]
    t = r // L[2]
    return 2 * xh + ((y + z + t + cb) & 1), y, z, t


def plane_wave(p, phi, cb):
    x, y, z, t = site_coords(cb)
    e = np.exp(1j * (p[0] * x + p[1] * y + p[2] * z + p[3] * t))
    return (e[:, None, None] * phi[None]).view(np.float64).ravel()


def gammas():
    g = []
    for mu in range(4):
        a = np.empty(32)
        oracle.lib().orc_gamma(mu, oracle.ptr(a))
        g.append(a.view(np.complex128).reshape(4, 4))
    return g


@pytest.mark.parametrize("prec", ["f32c12", "f64"])
def test_free_field_plane_wave(full, prec):
    gl = full.g32 if prec == "f32c12" else full.g64
    tol = 2e-6 if prec == "f32c12" else 1e-13
    u = q.gen_gauge(gl, 1, 0.0)  # exp(0) = 1: unit links
    ai = [q.clover_build(gl, u, CSW, MASS, cb, q.CLOV_AINV) for cb in (0, 1)]
    cs, ct = 1.0, 0.8
    D = q.Dslash(gl, t_boundary=-1.0, aniso_s=cs, aniso_t=ct)
    p = [2 * np.pi * 1 / L[0], 2 * np.pi * 2 / L[1], 2 * np.pi * 3 / L[2], 2 * np.pi * 5.5 / L[3]]
    rng = np.random.default_rng(11)
    phi = rng.standard_normal((4, 3)) + 1j * rng.standard_normal((4, 3))
    gam = gammas()
    for cb in (0, 1):
        psi = gl.spinor().upload(plane_wave(p, phi, 1 - cb), 1 - cb)
        for isign in (1, -1):
            K = sum(2 * c * (np.cos(pm) * np.eye(4) - isign * 1j * np.sin(pm) * gam[mu])
                    for mu, (pm, c) in enumerate(zip(p, (cs, cs, cs, ct))))
            want = plane_wave(p, K @ phi / (4 + MASS), cb)
            out = gl.spinor()
            D.clov_dslash(out, psi, u[cb], ai[cb], isign, cb)
            got = out.download(cb)
            r = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert r < tol, (cb, isign, r)
            del got, want


def test_even_odd_operator_adjoint(full):
    gl = full.g32
    u, a, ai = full.fields(gl)
    M = q.Operator(gl, q.OP_CLOVER, u, clov=[None, a[1]], invclov=[ai[0], None], t_boundary=-1.0)
    chi, psi = q.gen_spinor(gl, 50, 1), q.gen_spinor(gl, 51, 1)
    mpsi, mchi = gl.spinor(), gl.spinor()
    M(mpsi, psi, -1)
    M(mchi, chi, 1)
    lhs = q.inner_product(gl, chi, mpsi)
    rhs = q.inner_product(gl, mchi, psi)
    scale = (q.norm2(gl, chi) * q.norm2(gl, mpsi)) ** 0.5
    assert abs(lhs - rhs) < 1e-6 * scale, (lhs, rhs)


def test_multi_source_tile_bit_identical(full):
    gl = full.g32
    u, _, ai = full.fields(gl)
    D = q.Dslash(gl, t_boundary=-1.0)
    psis = [q.gen_spinor(gl, 60 + i, 0) for i in range(4)]
    assert q.lib().qphix_tune_rhs(4, 4) == q.OK
    try:
        outs = [gl.spinor() for _ in psis]
        D.clov_dslash_multi(outs, psis, u[1], ai[1], 1, 1)
    finally:
        q.lib().qphix_tune_rhs(0, 0)
    for i, ps in enumerate(psis):
        one = gl.spinor()
        D.clov_dslash(one, ps, u[1], ai[1], 1, 1)
        assert q.xmy_norm2(gl, gl.spinor(), outs[i], one) == 0.0, i


def test_mixed_precision_solve(full):
    g64, g16 = full.g64, full.g16
    Mo = q.Operator(g64, q.OP_CLOVER, full.u64, clov=[None, full.a64[1]], invclov=[full.ai64[0], None],
                    t_boundary=-1.0)
    u16, a16, ai16 = full.fields(g16)
    Mi = q.Operator(g16, q.OP_CLOVER, u16, clov=[None, a16[1]], invclov=[ai16[0], None], t_boundary=-1.0)
    b = q.gen_spinor(g64, 70, 1)
    x = g64.spinor()
    st = q.InvRichardsonMultiPrec(Mo, q.InvBiCGStab(Mi, 1000), 1e-2, 100)(x, b, 1e-10)
    assert st.status == q.OK
    r = g64.spinor()
    Mo(r, x, 1)
    assert rel_dev(g64, r, b) < 1e-10
    x64 = g64.spinor()
    st = q.InvBiCGStab(Mo, 2000)(x64, b, 1e-10)
    assert st.status == q.OK
    assert rel_dev(g64, x, x64) < 1e-7
