"""Parity of the gfx950 Dslash kernel family against the CPU oracle, through
the C ABI, on identical seeded inputs.

Tolerances (relative norm ||a - b|| / ||b|| over the target checkerboard):
FP64 1e-13 and FP32 2e-6 (the project's parity bar, tighter than the
reference's tests/tolerance.h), FP16-with-scale 5e-3 (tests/tolerance.h:19-31).
The 8^4 FP64 lattice with antiperiodic time is the reference's CPU-runnable
configuration; 4x6x4x8 exercises non-cubic extents and row wrap-around.
"""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

TOL = {q.F64: 1e-13, q.F32: 2e-6, q.F16: 5e-3, q.BF16: 4e-2}
L8 = [8, 8, 8, 8]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


class Setup:
    def __init__(self, L, prec, c12, seed=1, csw=1.0, mass=0.1):
        self.L, self.prec, self.c12 = L, prec, c12
        self.vh = L[0] * L[1] * L[2] * L[3] // 2
        self.u = oracle.gen_gauge(L, seed, 0.4)
        self.g = q.Geometry(L, prec, c12)
        self.ud = q.pack_gauge(self.g, self.u)
        clov = oracle.clover_build(L, self.u, csw, mass)
        self.clov = [clov[:self.vh * 72], clov[self.vh * 72:]]
        self.inv = [oracle.clover_invert(c) for c in self.clov]

    def spinor(self, seed, cb):
        h = oracle.gen_spinor(self.L, seed, cb)
        d = self.g.spinor().upload(h, cb)
        # the reference values the device actually holds
        return d, d.download(cb)

    def clov_dev(self, host):
        return self.g.clover().upload(host)


@pytest.fixture(scope="module", params=[(q.F64, False), (q.F64, True), (q.F32, False),
                                        (q.F32, True), (q.F16, True), (q.BF16, True)],
                ids=["f64c18", "f64c12", "f32c18", "f32c12", "f16c12", "bf16c12"])
def s8(request):
    prec, c12 = request.param
    return Setup(L8, prec, c12)


@pytest.mark.parametrize("cb", [0, 1])
@pytest.mark.parametrize("isign", [1, -1])
def test_wilson_dslash(s8, cb, isign):
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    D = q.Dslash(s8.g, t_boundary=-1.0)
    psi, psih = s8.spinor(10 + cb, 1 - cb)
    out = s8.g.spinor()
    D.dslash(out, psi, s8.ud[cb], isign, cb)
    ref = oracle.hop(s8.L, s8.u, psih, isign, cb, hp)
    assert rel(out.download(cb), ref) < TOL[s8.prec]


@pytest.mark.parametrize("isign", [1, -1])
def test_wilson_aniso_tbc(isign):
    L = [4, 6, 4, 8]
    for prec, c12 in [(q.F64, False), (q.F64, True)]:
        s = Setup(L, prec, c12, seed=3)
        tbc = [None, np.exp(0.7j), None, np.exp(-0.3j)]
        hp = oracle.hop_params(-1.0, 1.3, 0.7, tbc)
        D = q.Dslash(s.g, t_boundary=-1.0, aniso_s=1.3, aniso_t=0.7, tbc=tbc)
        for cb in (0, 1):
            psi, psih = s.spinor(20 + cb, 1 - cb)
            out = s.g.spinor()
            D.dslash(out, psi, s.ud[cb], isign, cb)
            ref = oracle.hop(L, s.u, psih, isign, cb, hp)
            assert rel(out.download(cb), ref) < 1e-13


@pytest.mark.parametrize("isign", [1, -1])
def test_wilson_achimbdpsi(s8, isign):
    D = q.Dslash(s8.g, t_boundary=-1.0)
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    for cb in (0, 1):
        psi, psih = s8.spinor(30, 1 - cb)
        chi, chih = s8.spinor(31, cb)
        out = s8.g.spinor()
        D.achimbdpsi(out, psi, chi, s8.ud[cb], 4.2, 0.3, isign, cb)
        ref = oracle.achimbdpsi(s8.L, s8.u, psih, chih, 4.2, 0.3, isign, cb, hp)
        assert rel(out.download(cb), ref) < TOL[s8.prec]


@pytest.mark.parametrize("isign", [1, -1])
def test_clover_dslash_and_achimbdpsi(s8, isign):
    rho = 0.05
    D = q.Dslash(s8.g, t_boundary=-1.0, prec_mass_rho=rho)
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    tol = TOL[s8.prec] * (4 if s8.prec in (q.F16, q.BF16) else 1)
    for cb in (0, 1):
        psi, psih = s8.spinor(40, 1 - cb)
        chi, chih = s8.spinor(41, cb)
        inv = s8.clov_dev(s8.inv[cb])
        clov = s8.clov_dev(s8.clov[cb])
        invh = s8.g.clover().upload(s8.inv[cb]).download()
        clovh = s8.g.clover().upload(s8.clov[cb]).download()
        out = s8.g.spinor()
        D.clov_dslash(out, psi, s8.ud[cb], inv, isign, cb)
        ref = oracle.clover_apply(invh, oracle.hop(s8.L, s8.u, psih, isign, cb, hp))
        assert rel(out.download(cb), ref) < tol
        D.clov_achimbdpsi(out, psi, chi, s8.ud[cb], clov, 0.25, isign, cb)
        ac = oracle.clover_apply(clovh, chih)
        g5 = np.tile(np.repeat([1.0, 1.0, -1.0, -1.0], 6), s8.vh)
        ig5 = np.empty_like(chih)
        ig5[0::2] = -g5[0::2] * chih[1::2]
        ig5[1::2] = g5[1::2] * chih[0::2]
        ref = ac + isign * rho * ig5 - 0.25 * oracle.hop(s8.L, s8.u, psih, isign, cb, hp)
        assert rel(out.download(cb), ref) < tol
        D.clov_mult(out, chi, clov)
        assert rel(out.download(cb), ac) < tol


def test_tm_dslash(s8):
    mass, mu = 0.1, 0.07
    al = 4 + mass
    D = q.Dslash(s8.g, t_boundary=-1.0, mass=mass, twisted_mass=mu)
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    n = s8.vh
    for isign in (1, -1):
        for cb in (0, 1):
            psi, psih = s8.spinor(50, 1 - cb)
            chi, chih = s8.spinor(51, cb)
            out = s8.g.spinor()
            D.tm_dslash(out, psi, s8.ud[cb], isign, cb)
            h = oracle.hop(s8.L, s8.u, psih, isign, cb, hp)
            tw = np.empty_like(h)
            den = al * al + mu * mu
            oracle.lib().orc_twisted_apply(n, al / den, -isign * mu / den, oracle.ptr(h),
                                           oracle.ptr(tw))
            assert rel(out.download(cb), tw) < TOL[s8.prec]
            for musign in (1.0, -1.0):
                D.tm_achimbdpsi(out, psi, chi, s8.ud[cb], al, 0.25, isign, cb, musign)
                ac = np.empty_like(chih)
                oracle.lib().orc_twisted_apply(n, al, isign * musign * mu, oracle.ptr(chih),
                                               oracle.ptr(ac))
                assert rel(out.download(cb), ac - 0.25 * h) < TOL[s8.prec]


def test_tmclov_dslash_full_blocks(s8):
    mu = 0.06
    D = q.Dslash(s8.g, t_boundary=-1.0, prec_mass_rho=mu)
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    tol = TOL[s8.prec] * (4 if s8.prec in (q.F16, q.BF16) else 1)
    for cb in (0, 1):
        fp = oracle.full_clover_tm_inverse(s8.clov[cb], mu)
        fm = oracle.full_clover_tm_inverse(s8.clov[cb], -mu)
        fdev = [s8.g.full_clover().upload(fp), s8.g.full_clover().upload(fm)]
        psi, psih = s8.spinor(60, 1 - cb)
        out = s8.g.spinor()
        for isign in (1, -1):
            for fl in (0, 1):
                D.tmclov_dslash(out, psi, s8.ud[cb], fdev, isign, cb, fl)
                use = fp if (isign == 1) == (fl == 0) else fm
                ref = oracle.full_clover_apply(use, oracle.hop(s8.L, s8.u, psih, isign, cb, hp))
                assert rel(out.download(cb), ref) < tol


def test_two_flav_inverse_clover_term(s8):
    mu, eps = 0.05, 0.02
    D = q.Dslash(s8.g)
    tol = TOL[s8.prec] * (4 if s8.prec in (q.F16, q.BF16) else 1)
    fu, fd, co = oracle.ndclover_inverse(s8.clov[0], mu, eps)
    fdev = [s8.g.full_clover().upload(fu), s8.g.full_clover().upload(fd)]
    cdev = s8.g.clover().upload(co)
    pu, puh = s8.spinor(70, 0)
    pd, pdh = s8.spinor(71, 0)
    res = [s8.g.spinor(), s8.g.spinor()]
    for isign in (1, -1):
        D.two_flav_inverse_clover_term(res, [pu, pd], fdev, cdev, isign)
        a, b = (fu, fd) if isign == 1 else (fd, fu)
        ru = oracle.full_clover_apply(a, puh) + oracle.clover_apply(co, pdh)
        rd = oracle.full_clover_apply(b, pdh) + oracle.clover_apply(co, puh)
        assert rel(res[0].download(), ru) < tol
        assert rel(res[1].download(), rd) < tol


def test_gpu_field_generation_matches_oracle():
    L = [4, 4, 4, 8]
    g = q.Geometry(L, q.F64, False)
    for cb in (0, 1):
        s = q.gen_spinor(g, 77, cb)
        assert rel(s.download(cb), oracle.gen_spinor(L, 77, cb)) < 1e-14
    # gauge: compare through the hop on both parities
    ud = q.gen_gauge(g, 5, 0.4)
    u = oracle.gen_gauge(L, 5, 0.4)
    D = q.Dslash(g, t_boundary=-1.0)
    for cb in (0, 1):
        psih = oracle.gen_spinor(L, 8, 1 - cb)
        psi = g.spinor().upload(psih, 1 - cb)
        out = g.spinor()
        D.dslash(out, psi, ud[cb], 1, cb)
        ref = oracle.hop(L, u, psih, 1, cb, oracle.hop_params(-1.0))
        assert rel(out.download(cb), ref) < 1e-13


@pytest.mark.parametrize("c12", [False, True])
def test_gpu_clover_build_matches_oracle(c12):
    L = [4, 4, 4, 8]
    vh = 256
    g = q.Geometry(L, q.F64, c12)
    u = oracle.gen_gauge(L, 9, 0.4)
    ud = q.pack_gauge(g, u)
    csw, mass, mu, eps = 1.0, 0.05, 0.04, 0.015
    clov = oracle.clover_build(L, u, csw, mass)
    for cb in (0, 1):
        ch = clov[cb * vh * 72:(cb + 1) * vh * 72]
        a = q.clover_build(g, ud, csw, mass, cb, q.CLOV_A)
        assert rel(a.download(), ch) < 1e-13
        ai = q.clover_build(g, ud, csw, mass, cb, q.CLOV_AINV)
        assert rel(ai.download(), oracle.clover_invert(ch)) < 1e-12
        co = q.clover_build(g, ud, csw, mass, cb, q.CLOV_ND_OFF, mu, eps)
        assert rel(co.download(), oracle.ndclover_inverse(ch, mu, eps)[2]) < 1e-12
        # full blocks, compared through clov-type multiplication of a spinor
        x, xh = None, oracle.gen_spinor(L, 3, cb)
        x = g.spinor().upload(xh, cb)
        out = g.spinor()
        for what, ref in ((q.CLOV_TMINV, oracle.full_clover_tm_inverse(ch, mu)),
                          (q.CLOV_ND_UP, oracle.ndclover_inverse(ch, mu, eps)[0]),
                          (q.CLOV_ND_DN, oracle.ndclover_inverse(ch, mu, eps)[1])):
            f = q.clover_build(g, ud, csw, mass, cb, what, mu, eps)
            D = q.Dslash(g)
            D.two_flav_inverse_clover_term([out, g.spinor()], [x, g.spinor()], [f, f],
                                           g.clover(), 1)
            assert rel(out.download(cb), oracle.full_clover_apply(ref, xh)) < 1e-12


def test_gpu_anisotropic_clover_build_matches_oracle():
    """Separate spatial / temporal clover coefficients and a free diagonal
    (the anisotropic clover term of tests/clover_term_qdp_w.h:156-180)."""
    L = [4, 4, 4, 8]
    vh = 256
    g = q.Geometry(L, q.F64, False)
    u = oracle.gen_gauge(L, 10, 0.4)
    ud = q.pack_gauge(g, u)
    cr, ct, dm = 0.7, 1.3, 2.2
    clov = oracle.clover_build(L, u, cr, 0.0, csw_t=ct, diag_mass=dm)
    for cb in (0, 1):
        ch = clov[cb * vh * 72:(cb + 1) * vh * 72]
        a = q.clover_build(g, ud, cr, 0.0, cb, q.CLOV_A, csw_t=ct, diag_mass=dm)
        assert rel(a.download(), ch) < 1e-13
        ai = q.clover_build(g, ud, cr, 0.0, cb, q.CLOV_AINV, csw_t=ct, diag_mass=dm)
        assert rel(ai.download(), oracle.clover_invert(ch)) < 1e-12


FLUX_CASES = [
    ([4, 4, 4, 8], {(0, 1): (1, (1, -1, 0)), (2, 3): (-1, (0, 1, -1))}, 3),
    ([4, 6, 4, 8], {(0, 2): (2, (1, -1, 0)), (1, 3): (1, (1, 1, -2))}, 4),
    ([6, 4, 4, 4], {(0, 3): (-1, (1, 0, -1)), (1, 2): (1, (1, -1, 0))}, 5),
]


@pytest.mark.parametrize("c12", [False, True], ids=["c18", "c12"])
@pytest.mark.parametrize("L,fluxes,rot", FLUX_CASES, ids=["01+23", "02+13", "03+12"])
def test_gpu_clover_build_closed_form(L, fluxes, rot, c12):
    """The device clover build on constant abelian fluxes (tests/closed_form.py)
    against the closed form diag_mass - sum (c/2) sigma (x) F with
    F = i sin(B T): A and A^-1 packed, the twisted-mass inverse blocks
    (A + i mu g5)^-1 and the doublet blocks (tests/mesfield.cc:24-64,
    tests/clover_term_qdp_w.h:156-195, 325-442)."""
    import closed_form as cf
    u, Fk = cf.abelian_flux(L, fluxes, rot)
    vh = L[0] * L[1] * L[2] * L[3] // 2
    g = q.Geometry(L, q.F64, c12)
    ud = q.pack_gauge(g, u)
    cr, ct, dm, mu, eps = 1.25, 0.75, 4.1, 0.3, 0.1
    A = cf.clover_dense_closed_form(Fk, cr, ct, dm)
    blocks = [A[:6, :6], A[6:, 6:]]
    g5 = [1.0, -1.0]
    inv = [np.linalg.inv(B) for B in blocks]
    tminv = [np.linalg.inv(B + 1j * mu * s * np.eye(6)) for B, s in zip(blocks, g5)]
    nd = [np.linalg.inv(B @ B + (mu * mu - eps * eps) * np.eye(6)) for B in blocks]
    ndoff = cf.pack_triangular([eps * m for m in nd])
    kw = dict(csw_t=ct, diag_mass=dm)
    for cb in (0, 1):
        a = q.clover_build(g, ud, cr, 0.0, cb, q.CLOV_A, **kw).download().reshape(vh, 72)
        assert np.abs(a - cf.pack_triangular(blocks)[None]).max() < 1e-13
        ai = q.clover_build(g, ud, cr, 0.0, cb, q.CLOV_AINV, **kw).download().reshape(vh, 72)
        assert np.abs(ai - cf.pack_triangular(inv)[None]).max() < 1e-13
        co = q.clover_build(g, ud, cr, 0.0, cb, q.CLOV_ND_OFF, mu, eps, **kw).download()
        assert np.abs(co.reshape(vh, 72) - ndoff[None]).max() < 1e-13
        # full blocks, through F x on a random spinor
        xh = oracle.gen_spinor(L, 3, cb)
        x = g.spinor().upload(xh, cb)
        X = xh.view(np.complex128).reshape(vh, 12)
        out = g.spinor()
        for what, dense in ((q.CLOV_TMINV, tminv),
                            (q.CLOV_ND_UP, [m @ (B - 1j * mu * s * np.eye(6))
                                            for m, B, s in zip(nd, blocks, g5)]),
                            (q.CLOV_ND_DN, [m @ (B + 1j * mu * s * np.eye(6))
                                            for m, B, s in zip(nd, blocks, g5)])):
            f = q.clover_build(g, ud, cr, 0.0, cb, what, mu, eps, **kw)
            q.call("qphix_clover_apply", g.h, out.p, x.p, f.p, 1)
            want = np.concatenate([X[:, :6] @ dense[0].T, X[:, 6:] @ dense[1].T], axis=1)
            got = out.download(cb).view(np.complex128).reshape(vh, 12)
            assert np.abs(got - want).max() < 1e-13, what


@pytest.mark.parametrize("prec,c12", [(q.F64, False), (q.F32, True), (q.F16, True), (q.BF16, True)],
                         ids=["f64", "f32", "f16", "bf16"])
def test_clover_apply_entry_point(prec, c12):
    """qphix_clover_apply: A x with packed Hermitian blocks and F x with full
    complex blocks (the site product of clover_product.h)."""
    L = [4, 4, 4, 8]
    vh = 256
    g = q.Geometry(L, prec, c12)
    u = oracle.gen_gauge(L, 11, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.1)[:vh * 72]
    fcl = oracle.full_clover_tm_inverse(clov, 0.05)
    x = g.spinor().upload(oracle.gen_spinor(L, 12, 0), 0)
    xh = x.download(0)
    cd = g.clover().upload(clov)
    fd = g.full_clover().upload(fcl)
    out = g.spinor()
    tol = TOL[prec] * (4 if prec in (q.F16, q.BF16) else 1)
    q.call("qphix_clover_apply", g.h, out.p, x.p, cd.p, 0)
    assert rel(out.download(0), oracle.clover_apply(clov, xh)) < tol
    q.call("qphix_clover_apply", g.h, out.p, x.p, fd.p, 1)
    assert rel(out.download(0), oracle.full_clover_apply(fcl, xh)) < tol
    with pytest.raises(q.QphixError):
        q.call("qphix_clover_apply", g.h, out.p, x.p, fd.p, 2)


def test_precision_conversion():
    L = [4, 4, 4, 8]
    g64 = q.Geometry(L, q.F64, False)
    g32 = q.Geometry(L, q.F32, True)
    g16 = q.Geometry(L, q.F16, True)
    gb16 = q.Geometry(L, q.BF16, True)
    xh = oracle.gen_spinor(L, 4, 0)
    x = g64.spinor().upload(xh)
    for gl, tol in ((g32, 1e-7), (g16, 2e-3), (gb16, 1.6e-2)):
        y = gl.spinor()
        q.convert_spinor(gl, y, 0.5, g64, x)
        z = g64.spinor()
        q.convert_spinor(g64, z, 2.0, gl, y)
        assert rel(z.download(), xh) < tol
    # gauge conversion F64 c18 -> F32/F16 c12 preserves the hop to the storage precision
    u = oracle.gen_gauge(L, 2, 0.4)
    ud = q.pack_gauge(g64, u)
    for gl, tol in ((g32, 2e-6), (g16, 5e-3), (gb16, 4e-2)):
        ul = q.convert_gauge(gl, g64, ud)
        psi = gl.spinor()
        q.convert_spinor(gl, psi, 1.0, g64, g64.spinor().upload(oracle.gen_spinor(L, 6, 0)))
        out = gl.spinor()
        q.Dslash(gl).dslash(out, psi, ul[1], 1, 1)
        ref = oracle.hop(L, u, oracle.gen_spinor(L, 6, 0), 1, 1)
        assert rel(out.download(1), ref) < tol
