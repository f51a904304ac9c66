"""Lattice shapes at the edges of the index arithmetic, against the oracle:
extent 2 (the forward and backward neighbours are the same site), an odd
number of checkerboarded x sites (Nxh = 3, 5, 7), parities smaller than one
wavefront and parities that end in a partial wavefront (the cross-lane x / y
exchange must fall back to loads there), and a parity of 8,400 sites that spans
two whole 4096-site layout blocks and ends in a partial third (neighbours in
other blocks, odd Nxh, rows straddling block ends). Wilson, clover, clover
A chi - beta D psi and twisted-mass kernels in FP64, FP32 and FP16, both
checkerboards and both signs, and an FP64 solve on the smallest lattice."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

TOL = {q.F64: 1e-13, q.F32: 2e-6, q.F16: 5e-3, q.BF16: 4e-2}
SHAPES = [[2, 2, 2, 2], [6, 2, 2, 4], [2, 4, 2, 6], [10, 6, 2, 2], [14, 6, 4, 2], [6, 10, 6, 4],
          [10, 12, 14, 10]]
PRECS = [(q.F64, False), (q.F32, True), (q.F16, True), (q.BF16, True)]


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module", params=SHAPES, ids=lambda L: "x".join(map(str, L)))
def lat(request):
    L = request.param
    vh = L[0] * L[1] * L[2] * L[3] // 2
    u = oracle.gen_gauge(L, 21, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.1)
    cl = [clov[:vh * 72], clov[vh * 72:]]
    return L, vh, u, cl, [oracle.clover_invert(c) for c in cl]


@pytest.mark.parametrize("prec,c12", PRECS, ids=["f64", "f32c12", "f16c12", "bf16c12"])
def test_kernels(lat, prec, c12):
    L, vh, u, cl, inv = lat
    g = q.Geometry(L, prec, c12)
    ud = q.pack_gauge(g, u)
    hp = oracle.hop_params(-1.0, 0.9, 1.1)
    D = q.Dslash(g, t_boundary=-1.0, aniso_s=0.9, aniso_t=1.1)
    Dtm = q.Dslash(g, t_boundary=-1.0, aniso_s=0.9, aniso_t=1.1, mass=0.1, twisted_mass=0.05)
    for cb in (0, 1):
        psi = g.spinor().upload(oracle.gen_spinor(L, 30 + cb, 1 - cb), 1 - cb)
        chi = g.spinor().upload(oracle.gen_spinor(L, 40 + cb, cb), cb)
        psih, chih = psi.download(1 - cb), chi.download(cb)  # the values the device holds
        invd, cld = g.clover().upload(inv[cb]), g.clover().upload(cl[cb])
        for isign in (1, -1):
            hop = oracle.hop(L, u, psih, isign, cb, hp)
            out = g.spinor()
            D.dslash(out, psi, ud[cb], isign, cb)
            assert rel(out.download(cb), hop) < TOL[prec], ("wilson", cb, isign)
            D.clov_dslash(out, psi, ud[cb], invd, isign, cb)
            assert rel(out.download(cb), oracle.clover_apply(inv[cb], hop)) < TOL[prec], ("clover", cb, isign)
            D.clov_achimbdpsi(out, psi, chi, ud[cb], cld, 0.25, isign, cb)
            ref = oracle.clover_apply(cl[cb], chih) - 0.25 * hop
            assert rel(out.download(cb), ref) < TOL[prec], ("clover achimbdpsi", cb, isign)
            # TMDslash: (alpha + i isign mu g5)^-1 D psi, alpha = 4 + m whatever the
            # anisotropy (tm_dslash_body.h:31-34)
            Dtm.tm_dslash(out, psi, ud[cb], isign, cb)
            al, mu = 4.1, 0.05
            den = al * al + mu * mu
            tw = np.empty_like(hop)
            oracle.lib().orc_twisted_apply(vh, al / den, -isign * mu / den, oracle.ptr(hop), oracle.ptr(tw))
            assert rel(out.download(cb), tw) < TOL[prec], ("tm", cb, isign)


def test_solve_smallest_lattice():
    L = [2, 2, 2, 2]
    vh = 8
    u = oracle.gen_gauge(L, 22, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.1)
    cl = [clov[:vh * 72], clov[vh * 72:]]
    inv = [oracle.clover_invert(c) for c in cl]
    g = q.Geometry(L, q.F64)
    ud = q.pack_gauge(g, u)
    M = q.Operator(g, q.OP_CLOVER, ud, clov=[None, g.clover().upload(cl[1])],
                   invclov=[g.clover().upload(inv[0]), None], t_boundary=-1.0)
    Mo = oracle.Operator(oracle.CLOVER, L, u, oracle.hop_params(-1.0), clov=cl, invclov=inv)
    bh = oracle.gen_spinor(L, 23, 1)
    # BiCGStab solves M x = b
    x, b = g.spinor(), g.spinor().upload(bh, 1)
    st = q.InvBiCGStab(M, 200)(x, b, 1e-12)
    assert st.status == q.OK
    assert rel(Mo(x.download(1), 1, 1), bh) < 1e-11
    # CG solves M^+M x = b
    x = g.spinor()
    st = q.InvCG(M, 200)(x, b, 1e-12)
    assert st.status == q.OK
    assert rel(Mo(Mo(x.download(1), 1, 1), -1, 1), bh) < 1e-11
