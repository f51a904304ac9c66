"""Pins the CPU oracle (oracle/qphix_oracle.c) before its results are trusted
as the reference for the HIP path.

Each check is chosen so that a plausible slip in the oracle (a swapped
projector sign, U instead of U^+, a link taken at x instead of x-mu, a
mis-indexed clover entry, a wrong twisted-mass sign) fails at least one of:
  * the reference's own spin projection / reconstruction tables
    (tests/golden/spin_tables.json, codegen/dslash_common.cc:848-990);
  * Clifford algebra of the gamma matrices;
  * free-field plane waves (closed-form eigenvalues, antiperiodic shift);
  * gauge covariance;
  * gamma_5-hermiticity / M(isign=-1) == M(isign=+1)^+ on dense matrices;
  * the even-odd Schur-complement identity against the full operator;
  * clover structure (Hermitian, unit gauge, sigma.F proportionality,
    A A^-1 = 1), and dense linear solves for the Krylov solvers.
"""
import json
import math
import os

import numpy as np
import pytest
import scipy.linalg

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
L4 = [4, 4, 4, 4]


def cplx(a):
    return a.view(np.complex128)


def gammas(orc):
    g = []
    for mu in range(4):
        a = np.empty(32)
        orc.lib().orc_gamma(mu, orc.ptr(a))
        g.append(a.view(np.complex128).reshape(4, 4))
    return g


def test_spin_tables_match_reference(orc):
    tab = json.load(open(os.path.join(HERE, "golden", "spin_tables.json")))
    g = gammas(orc)
    rng = np.random.default_rng(0)
    psi = rng.normal(size=4) + 1j * rng.normal(size=4)
    for name, s in (("plus", 1.0), ("minus", -1.0)):
        for mu in range(4):
            full = (np.eye(4) + s * g[mu]) @ psi
            h = []
            for a, b, c in tab["proj"][name][str(mu)]:
                h.append(psi[a] + complex(*c) * psi[b])
            assert np.allclose(full[:2], h, atol=1e-15)
            (i2, c2), (i3, c3) = tab["recon"][name][str(mu)]
            assert np.allclose(full[2], complex(*c2) * h[i2], atol=1e-15)
            assert np.allclose(full[3], complex(*c3) * h[i3], atol=1e-15)


def test_clifford_algebra(orc):
    g = gammas(orc)
    for m in range(4):
        assert np.allclose(g[m], g[m].conj().T)
        for n in range(4):
            ac = g[m] @ g[n] + g[n] @ g[m]
            assert np.allclose(ac, 2.0 * np.eye(4) * (m == n))
    g5 = g[0] @ g[1] @ g[2] @ g[3]
    assert np.allclose(g5, np.diag([1, 1, -1, -1]))


def test_rng_moments(orc):
    vals = np.array([orc.lib().orc_uniform(7, k) for k in range(20000)])
    assert 0 < vals.min() and vals.max() < 1
    assert abs(vals.mean() - 0.5) < 0.01
    psi = orc.gen_spinor([4, 4, 4, 8], 11, 0)
    assert abs(psi.mean()) < 0.05 and abs(psi.std() - 1.0) < 0.05


def test_su3_exp_matches_expm(orc):
    rng = np.random.default_rng(3)
    for _ in range(5):
        G = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
        A = 0.5 * (G + G.conj().T)
        A -= np.trace(A).real / 3 * np.eye(3)
        U = np.empty(18)
        orc.lib().orc_su3_exp_i(orc.ptr(np.ascontiguousarray(A).view(np.float64).ravel().copy()),
                                0.4, orc.ptr(U))
        ref = scipy.linalg.expm(1j * 0.4 * A)
        assert np.allclose(U.view(np.complex128).reshape(3, 3), ref, atol=1e-14)


def test_gauge_is_su3(orc):
    u = orc.gen_gauge([4, 4, 4, 4], 1, 0.4)
    U = u.view(np.complex128).reshape(-1, 3, 3)
    eye = np.einsum("nij,nkj->nik", U, U.conj())
    assert np.abs(eye - np.eye(3)).max() < 1e-14
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-14
    # not trivially the identity
    assert np.abs(U - np.eye(3)).mean() > 0.1


def unit_gauge(L):
    V = L[0] * L[1] * L[2] * L[3]
    u = np.zeros((V, 4, 3, 3), dtype=np.complex128)
    u[:, :, range(3), range(3)] = 1
    return u.view(np.float64).ravel().copy()


def coords_of_cb(L, cb):
    vh = L[0] * L[1] * L[2] * L[3] // 2
    out = np.empty((vh, 4), dtype=np.int64)
    nxh = L[0] // 2
    for j in range(vh):
        r = j
        xh = r % nxh; r //= nxh
        y = r % L[1]; r //= L[1]
        z = r % L[2]; t = r // L[2]
        out[j] = (2 * xh + ((y + z + t + cb) & 1), y, z, t)
    return out


@pytest.mark.parametrize("isign", [1, -1])
@pytest.mark.parametrize("tb", [1.0, -1.0])
def test_plane_wave(orc, isign, tb):
    L = [4, 6, 4, 8]
    g = gammas(orc)
    u = unit_gauge(L)
    n = np.array([1, 2, 3, 1])
    p = 2 * math.pi * n / np.array(L, dtype=float)
    if tb < 0:
        p[3] += math.pi / L[3]
    rng = np.random.default_rng(1)
    eta = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    aniso_s, aniso_t = 1.0, 0.7
    for cb in (0, 1):
        src = coords_of_cb(L, 1 - cb)
        ph = np.exp(1j * src @ p)
        psi = (ph[:, None, None] * eta[None]).astype(np.complex128)
        out = orc.hop(L, u, psi.view(np.float64).ravel().copy(), isign, cb,
                      orc.hop_params(tb, aniso_s, aniso_t))
        K = np.zeros((4, 4), dtype=np.complex128)
        for mu in range(4):
            c = aniso_t if mu == 3 else aniso_s
            K += c * (np.exp(1j * p[mu]) * (np.eye(4) - isign * g[mu])
                      + np.exp(-1j * p[mu]) * (np.eye(4) + isign * g[mu]))
        tgt = coords_of_cb(L, cb)
        expect = np.exp(1j * tgt @ p)[:, None, None] * (K @ eta)[None]
        assert np.abs(cplx(out).reshape(-1, 4, 3) - expect).max() < 1e-12


def test_dslash_dir_plane_wave_and_sum(orc):
    """dslashDir (dslash_body.h:1074-1137): on the free field a plane wave picks
    up exactly one term of the hop kernel per direction, dir = 2mu the forward
    term c e^{ip_mu} (1 - g_mu), dir = 2mu+1 the backward one; on a random
    gauge field the eight directions add up to the isign = +1 hop."""
    L = [4, 6, 4, 8]
    g = gammas(orc)
    u = unit_gauge(L)
    p = 2 * math.pi * np.array([1, 2, 3, 1]) / np.array(L, dtype=float)
    p[3] += math.pi / L[3]  # antiperiodic in t
    rng = np.random.default_rng(4)
    eta = rng.normal(size=(4, 3)) + 1j * rng.normal(size=(4, 3))
    hp = orc.hop_params(-1.0, 1.0, 0.7)
    for cb in (0, 1):
        src = coords_of_cb(L, 1 - cb)
        psi = (np.exp(1j * src @ p)[:, None, None] * eta[None]).view(np.float64).ravel().copy()
        tgt = coords_of_cb(L, cb)
        for d in range(8):
            mu, fwd = d // 2, d % 2 == 0
            c = 0.7 if mu == 3 else 1.0
            K = c * (np.exp(1j * p[mu]) * (np.eye(4) - g[mu]) if fwd
                     else np.exp(-1j * p[mu]) * (np.eye(4) + g[mu]))
            out = orc.dslash_dir(L, u, psi, cb, d, hp)
            expect = np.exp(1j * tgt @ p)[:, None, None] * (K @ eta)[None]
            assert np.abs(cplx(out).reshape(-1, 4, 3) - expect).max() < 1e-12, d
    ur = orc.gen_gauge(L, 9, 0.4)
    psi = orc.gen_spinor(L, 10, 0)
    tot = sum(orc.dslash_dir(L, ur, psi, 1, d, hp) for d in range(8))
    assert np.abs(tot - orc.hop(L, ur, psi, 1, 1, hp)).max() < 1e-13


def test_dslash_dir_point_source(orc):
    """dir = 2mu moves a point source at y to the site y - mu (the result at x
    reads psi(x + mu)); dir = 2mu + 1 to y + mu."""
    L = [4, 4, 4, 8]
    vh = 4 * 4 * 4 * 8 // 2
    u = orc.gen_gauge(L, 3, 0.4)
    ycb = 0
    j = 37
    yc = coords_of_cb(L, ycb)[j]
    psi = np.zeros(vh * 24)
    psi[j * 24 + 5] = 1.0
    tgt = coords_of_cb(L, 1 - ycb)
    for d in range(8):
        mu, fwd = d // 2, d % 2 == 0
        out = orc.dslash_dir(L, u, psi, 1 - ycb, d)
        nz = np.nonzero(np.abs(out.reshape(vh, 24)).sum(axis=1))[0]
        assert len(nz) == 1, d
        x = tgt[nz[0]].copy()
        x[mu] = (x[mu] + (1 if fwd else -1)) % L[mu]
        assert np.array_equal(x, yc), d


def random_su3(rng, n):
    G = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    Q, R = np.linalg.qr(G)
    d = np.diagonal(R, axis1=1, axis2=2)
    Q = Q * (d / np.abs(d))[:, None, :]
    det = np.linalg.det(Q)
    Q[:, :, 0] /= det[:, None]
    return Q


def full_index(L, c):
    x, y, z, t = c
    cb = (x + y + z + t) & 1
    vh = L[0] * L[1] * L[2] * L[3] // 2
    return cb * vh + ((t * L[2] + z) * L[1] + y) * (L[0] // 2) + x // 2


def gauge_transform(L, u, gmat):
    V = L[0] * L[1] * L[2] * L[3]
    U = cplx(u).reshape(V, 4, 3, 3)
    out = np.empty_like(U)
    for i in range(V):
        cb = i // (V // 2)
        c = coords_of_cb(L, cb)[i - cb * (V // 2)]
        for mu in range(4):
            d = c.copy(); d[mu] = (d[mu] + 1) % L[mu]
            j = full_index(L, d)
            out[i, mu] = gmat[i] @ U[i, mu] @ gmat[j].conj().T
    return out.view(np.float64).ravel().copy()


def test_gauge_covariance_hop(orc):
    L = [4, 4, 4, 4]
    V = 256
    rng = np.random.default_rng(5)
    u = orc.gen_gauge(L, 3, 0.4)
    gm = random_su3(rng, V)
    u2 = gauge_transform(L, u, gm)
    p = orc.hop_params(-1.0, 1.3, 0.8, tbc=[None, np.exp(0.3j), None, None])
    for cb in (0, 1):
        psi = orc.gen_spinor(L, 9, 1 - cb)
        P = cplx(psi).reshape(-1, 4, 3)
        gs = gm[(1 - cb) * 128:(2 - cb) * 128]
        psi2 = np.einsum("nij,nsj->nsi", gs, P).copy().view(np.float64).ravel()
        out = cplx(orc.hop(L, u, psi, 1, cb, p)).reshape(-1, 4, 3)
        out2 = cplx(orc.hop(L, u2, psi2.copy(), 1, cb, p)).reshape(-1, 4, 3)
        gt = gm[cb * 128:(cb + 1) * 128]
        assert np.abs(np.einsum("nij,nsj->nsi", gt, out) - out2).max() < 1e-12


def dense(fn, n):
    cols = []
    for k in range(n):
        e = np.zeros(n, dtype=np.complex128)
        e[k] = 1
        cols.append(cplx(fn(e.view(np.float64).copy())))
    return np.stack(cols, axis=1)


def test_gamma5_hermiticity_hop(orc):
    """<chi, D_{-1} psi> = <D_{+1} chi, psi> between parities."""
    L = [4, 4, 4, 4]
    u = orc.gen_gauge(L, 2, 0.4)
    p = orc.hop_params(-1.0, 1.0, 1.0)
    for cb in (0, 1):
        D_plus = dense(lambda v: orc.hop(L, u, v, 1, cb, p), 128 * 12)
        D_minus_other = dense(lambda v: orc.hop(L, u, v, -1, 1 - cb, p), 128 * 12)
        assert np.abs(D_plus.conj().T - D_minus_other).max() < 1e-13


def test_aniso_and_tboundary_only_touch_time(orc):
    L = [4, 4, 4, 4]
    u = orc.gen_gauge(L, 2, 0.4)
    psi = orc.gen_spinor(L, 4, 0)
    a = orc.hop(L, u, psi, 1, 1, orc.hop_params(1.0, 1.0, 1.0))
    b = orc.hop(L, u, psi, 1, 1, orc.hop_params(1.0, 1.0, 2.0))
    c = orc.hop(L, u, psi, 1, 1, orc.hop_params(1.0, 2.0, 1.0))
    d = orc.hop(L, u, psi, 1, 1, orc.hop_params(1.0, 2.0, 2.0))
    assert np.allclose(d, 2 * a, atol=1e-13)
    assert np.allclose((b - a) + (c - a), a, atol=1e-13)
    # antiperiodic t: only t=0 and t=Nt-1 target sites change
    e = orc.hop(L, u, psi, 1, 1, orc.hop_params(-1.0, 1.0, 1.0))
    diff = np.abs(e - a).reshape(-1, 24).max(axis=1)
    t = coords_of_cb(L, 1)[:, 3]
    assert np.all(diff[(t > 0) & (t < L[3] - 1)] < 1e-14)
    assert np.all(diff[(t == 0) | (t == L[3] - 1)] > 1e-3)


def test_achimbdpsi(orc):
    L = [4, 4, 4, 4]
    u = orc.gen_gauge(L, 2, 0.4)
    psi = orc.gen_spinor(L, 4, 0)
    chi = orc.gen_spinor(L, 5, 1)
    h = orc.hop(L, u, psi, -1, 1)
    r = orc.achimbdpsi(L, u, psi, chi, 2.5, 0.3, -1, 1)
    assert np.allclose(r, 2.5 * chi - 0.3 * h, atol=1e-13)


# ----------------------------------------------------------------- clover
def pauli_sigma(g):
    """sigma_{mu nu} = (1/2)[g_mu, g_nu] (anti-Hermitian convention)."""
    s = {}
    for m in range(4):
        for n in range(m + 1, 4):
            s[(m, n)] = 0.5 * (g[m] @ g[n] - g[n] @ g[m])
    return s


def test_clover_unit_gauge(orc):
    L = [4, 4, 4, 4]
    u = unit_gauge(L)
    clov = orc.clover_build(L, u, 1.0, 0.1)
    site = clov.reshape(-1, 72)
    assert np.allclose(site[:, :12], 4.1)
    assert np.allclose(site[:, 12:], 0.0)


def test_clover_sigma_f_structure(orc):
    """A - diag_mass is one common multiple of sum_{mu<nu} sigma_{mu nu} (x) F_{mu nu}.

    Chroma's packed formula is pinned by requiring the six plane coefficients to
    be identical and the fit residual to vanish: any mis-assigned E/B entry or
    sign breaks either."""
    L = [4, 4, 4, 4]
    u = orc.gen_gauge(L, 8, 0.4)
    csw = 1.0
    clov = orc.clover_build(L, u, csw, 0.0)
    f = np.empty(2 * 128 * 6 * 18)
    orc.lib().orc_field_strength((orc.C.c_int * 4)(*L), orc.ptr(u), orc.ptr(f))
    F = cplx(f).reshape(-1, 6, 3, 3)
    g = gammas(orc)
    sig = pauli_sigma(g)
    planes = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    rows, rhs = [], []
    for i in range(0, 256, 17):
        A = np.zeros((12, 12), dtype=np.complex128)
        for b in range(2):
            A[6 * b:6 * b + 6, 6 * b:6 * b + 6] = orc.clover_dense(clov.reshape(-1, 72)[i], b)
        A -= 4.0 * np.eye(12)
        rows.append(np.stack([np.kron(sig[pl], F[i, k]).ravel()
                              for k, pl in enumerate(planes)], axis=1))
        rhs.append(A.ravel())
    M = np.concatenate(rows)
    y = np.concatenate(rhs)
    coef, *_ = np.linalg.lstsq(M, y, rcond=None)
    assert np.abs(M @ coef - y).max() < 1e-12
    assert np.allclose(coef, coef[0], atol=1e-12)
    # Sheikholeslami-Wohlert normalisation: with the leaf average F = i F_h
    # (F_h Hermitian) and [g_mu,g_nu]/2 = -i sigma_h, the fit coefficient -c_sw/2
    # is A = (4+m) - (c_sw/4) sum_{mu,nu} sigma_h F_h, the textbook clover term
    # of the Wilson operator (4+m) - D/2.
    assert abs(coef[0] + 0.5 * csw) < 1e-12


def test_clover_plane_coefficients(orc):
    """Anisotropic clover term: csw_r weighs the planes without t and csw_t the
    planes with t (clovCoeffR / clovCoeffT, tests/clover_term_qdp_w.h:156-166),
    and diag_mass is the whole diagonal. The sigma (x) F fit of the isotropic
    test, with one free coefficient per plane, must return -csw/2 of the
    plane's kind on each."""
    L = [4, 4, 4, 4]
    u = orc.gen_gauge(L, 9, 0.4)
    cr, ct, dm = 1.25, 0.5, 3.0
    clov = orc.clover_build(L, u, cr, 0.0, csw_t=ct, diag_mass=dm)
    f = np.empty(2 * 128 * 6 * 18)
    orc.lib().orc_field_strength((orc.C.c_int * 4)(*L), orc.ptr(u), orc.ptr(f))
    F = cplx(f).reshape(-1, 6, 3, 3)
    sig = pauli_sigma(gammas(orc))
    planes = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
    rows, rhs = [], []
    for i in range(0, 256, 17):
        A = np.zeros((12, 12), dtype=np.complex128)
        for b in range(2):
            A[6 * b:6 * b + 6, 6 * b:6 * b + 6] = orc.clover_dense(clov.reshape(-1, 72)[i], b)
        A -= dm * np.eye(12)
        rows.append(np.stack([np.kron(sig[pl], F[i, k]).ravel()
                              for k, pl in enumerate(planes)], axis=1))
        rhs.append(A.ravel())
    M = np.concatenate(rows)
    y = np.concatenate(rhs)
    coef, *_ = np.linalg.lstsq(M, y, rcond=None)
    assert np.abs(M @ coef - y).max() < 1e-12
    want = [-0.5 * (ct if 3 in pl else cr) for pl in planes]
    assert np.allclose(coef, want, atol=1e-12), coef


# constant abelian fluxes through two disjoint planes, rotated by a constant
# SU(3) matrix (tests/closed_form.py): lattice, {plane: (n, T)}, rotation seed
FLUX_CASES = [
    ([4, 4, 4, 8], {(0, 1): (1, (1, -1, 0)), (2, 3): (-1, (0, 1, -1))}, 3),
    ([4, 6, 4, 8], {(0, 2): (2, (1, -1, 0)), (1, 3): (1, (1, 1, -2))}, 4),
    ([6, 4, 4, 4], {(0, 3): (-1, (1, 0, -1)), (1, 2): (1, (1, -1, 0))}, 5),
    ([4, 4, 4, 4], {(0, 1): (1, (1, -1, 0))}, None),
]


@pytest.mark.parametrize("L,fluxes,rot", FLUX_CASES)
def test_field_strength_closed_form(orc, L, fluxes, rot):
    """F_{mu nu} of a constant abelian flux is i sin(B T) on every site and
    zero on the planes without flux (tests/mesfield.cc:24-64: four leaves,
    (Q - Q^+)/8).  A reversed leaf halves it, a wrong factor scales it."""
    import closed_form as cf
    u, Fk = cf.abelian_flux(L, fluxes, rot)
    V = L[0] * L[1] * L[2] * L[3]
    f = np.empty(V * 6 * 18)
    orc.lib().orc_field_strength((orc.C.c_int * 4)(*L), orc.ptr(u), orc.ptr(f))
    F = cplx(f).reshape(V, 6, 3, 3)
    for k, pl in enumerate(cf.PLANES):
        assert np.abs(F[:, k] - Fk[pl][None]).max() < 1e-14, pl
    assert sum(np.abs(Fk[pl]).max() > 0.1 for pl in cf.PLANES) == len(fluxes)


@pytest.mark.parametrize("L,fluxes,rot", FLUX_CASES)
def test_clover_closed_form(orc, L, fluxes, rot):
    """The packed clover term of the abelian flux equals
    diag_mass - sum_{mu<nu} (c/2) sigma_{mu nu} (x) F_{mu nu} with the closed-form
    F, c = c_sw^T on the planes with t and c_sw^R on the others
    (tests/clover_term_qdp_w.h:156-195, 325-442).  The closed form is block
    diagonal in the chiral basis (spins 0,1 and 2,3)."""
    import closed_form as cf
    u, Fk = cf.abelian_flux(L, fluxes, rot)
    cr, ct, dm = 1.25, 0.75, 4.1
    clov = orc.clover_build(L, u, cr, 0.0, csw_t=ct, diag_mass=dm).reshape(-1, 72)
    A = cf.clover_dense_closed_form(Fk, cr, ct, dm)
    assert np.abs(A[:6, 6:]).max() < 1e-15 and np.abs(A[6:, :6]).max() < 1e-15
    want = cf.pack_triangular([A[:6, :6], A[6:, 6:]])
    assert np.abs(clov - want[None]).max() < 1e-14
    if rot is not None:
        assert np.abs(want[12:]).max() > 0.05  # off-diagonal entries are exercised
    for i in range(0, clov.shape[0], 37):
        for b in range(2):
            assert np.allclose(cf.unpack_triangular(clov[i])[b], orc.clover_dense(clov[i], b),
                               atol=1e-15)


def test_clover_covariance_and_inverse(orc):
    L = [4, 4, 4, 4]
    rng = np.random.default_rng(2)
    u = orc.gen_gauge(L, 8, 0.4)
    gm = random_su3(rng, 256)
    u2 = gauge_transform(L, u, gm)
    c1 = orc.clover_build(L, u, 1.0, 0.1).reshape(-1, 72)
    c2 = orc.clover_build(L, u2, 1.0, 0.1).reshape(-1, 72)
    for i in range(0, 256, 13):
        for b in range(2):
            A1 = orc.clover_dense(c1[i], b)
            A2 = orc.clover_dense(c2[i], b)
            G = np.kron(np.eye(2), gm[i])
            assert np.abs(G @ A1 @ G.conj().T - A2).max() < 1e-12
            assert np.allclose(A1, A1.conj().T)
    inv = orc.clover_invert(c1.ravel()).reshape(-1, 72)
    for i in range(0, 256, 7):
        for b in range(2):
            assert np.abs(orc.clover_dense(c1[i], b) @ orc.clover_dense(inv[i], b)
                          - np.eye(6)).max() < 1e-13


def test_clover_apply_matches_dense(orc):
    L = [4, 4, 4, 4]
    u = orc.gen_gauge(L, 8, 0.4)
    clov = orc.clover_build(L, u, 1.0, 0.1)[:128 * 72]
    x = orc.gen_spinor(L, 1, 0)
    y = cplx(orc.clover_apply(clov, x)).reshape(-1, 2, 6)
    X = cplx(x).reshape(-1, 2, 6)
    for i in range(128):
        for b in range(2):
            assert np.allclose(orc.clover_dense(clov.reshape(-1, 72)[i], b) @ X[i, b], y[i, b])


def test_tm_and_nd_clover_inverses(orc):
    L = [4, 4, 4, 4]
    u = orc.gen_gauge(L, 8, 0.4)
    clov = orc.clover_build(L, u, 1.0, 0.1)[:128 * 72]
    mu, eps = 0.05, 0.02
    fp = cplx(orc.full_clover_tm_inverse(clov, mu)).reshape(-1, 2, 6, 6)
    fu, fd, co = orc.ndclover_inverse(clov, mu, eps)
    fu = cplx(fu).reshape(-1, 2, 6, 6)
    fd = cplx(fd).reshape(-1, 2, 6, 6)
    for i in range(0, 128, 9):
        for b in range(2):
            A = orc.clover_dense(clov.reshape(-1, 72)[i], b)
            g5 = 1 if b == 0 else -1
            assert np.abs((A + 1j * mu * g5 * np.eye(6)) @ fp[i, b] - np.eye(6)).max() < 1e-13
            B = np.block([[A + 1j * mu * g5 * np.eye(6), -eps * np.eye(6)],
                          [-eps * np.eye(6), A - 1j * mu * g5 * np.eye(6)]])
            C = orc.clover_dense(co.reshape(-1, 72)[i], b)
            Binv = np.block([[fu[i, b], C], [C, fd[i, b]]])
            assert np.abs(B @ Binv - np.eye(12)).max() < 1e-13


# --------------------------------------------------------------- operators
def make_ops(orc, L, seed=3):
    u = orc.gen_gauge(L, seed, 0.4)
    hp = orc.hop_params(-1.0, 1.0, 1.0)
    vh = L[0] * L[1] * L[2] * L[3] // 2
    clov = orc.clover_build(L, u, 1.0, 0.05)
    cl = [clov[:vh * 72], clov[vh * 72:]]
    inv = [orc.clover_invert(c) for c in cl]
    mu, eps = 0.07, 0.03
    fcl = [[orc.full_clover_tm_inverse(c, mu), orc.full_clover_tm_inverse(c, -mu)] for c in cl]
    nd = [orc.ndclover_inverse(c, mu, eps) for c in cl]
    return {
        "wilson": orc.Operator(orc.WILSON, L, u, hp, mass=0.1),
        "clover": orc.Operator(orc.CLOVER, L, u, hp, clov=cl, invclov=inv),
        "clover_rho": orc.Operator(orc.CLOVER, L, u, hp, clov=cl, invclov=inv, rho=0.02),
        "tmwilson": orc.Operator(orc.TMWILSON, L, u, hp, mass=0.1, mu=mu),
        "tmclover": orc.Operator(orc.TMCLOVER, L, u, hp, clov=cl, fcl=fcl, mu=mu, rho=mu),
        "ndtmwilson": orc.Operator(orc.NDTMWILSON, L, u, hp, mass=0.1, mu=mu, eps=eps),
        "ndtmclover": orc.Operator(orc.NDTMCLOVER, L, u, hp, clov=cl, mu=mu, eps=eps, rho=mu,
                                   ndfcl=[[n[0], n[1]] for n in nd],
                                   ndcoff=[n[2] for n in nd]),
    }, u, cl, (mu, eps)


@pytest.fixture(scope="module")
def ops44(orc):
    return make_ops(orc, [4, 4, 4, 4])


@pytest.mark.parametrize("name", ["wilson", "clover_rho", "tmwilson", "tmclover",
                                  "ndtmwilson", "ndtmclover"])
def test_operator_dagger(orc, ops44, name):
    ops = ops44[0]
    op = ops[name]
    n = 128 * 12 * op.nflav
    Mp = dense(lambda v: op(v, 1, 1), n)
    Mm = dense(lambda v: op(v, -1, 1), n)
    assert np.abs(Mp.conj().T - Mm).max() < 1e-12


def site_diag_apply(kind, x, cb, fields, isign, inverse):
    """The site-diagonal block A of the unpreconditioned operator, from first
    principles (numpy), used for the Schur-complement identity."""
    u, cl, (mu, eps), mass = fields
    X = cplx(x).reshape(-1, 2, 2, 6) if kind.startswith("nd") else cplx(x).reshape(-1, 2, 6)
    out = np.empty_like(X)
    nsite = X.shape[0]
    for i in range(nsite):
        for b in range(2):
            g5 = 1 if b == 0 else -1
            if kind in ("wilson", "tmwilson", "ndtmwilson"):
                A = (4 + mass) * np.eye(6)
            else:
                A = oracle.clover_dense(cl[cb].reshape(-1, 72)[i], b)
            if kind.startswith("tm"):
                A = A + 1j * isign * mu * g5 * np.eye(6)
            if kind.startswith("nd"):
                A2 = np.block([[A + 1j * isign * mu * g5 * np.eye(6), -eps * np.eye(6)],
                               [-eps * np.eye(6), A - 1j * isign * mu * g5 * np.eye(6)]])
                v = np.concatenate([X[i, 0, b], X[i, 1, b]])
                r = np.linalg.solve(A2, v) if inverse else A2 @ v
                out[i, 0, b], out[i, 1, b] = r[:6], r[6:]
            else:
                out[i, b] = np.linalg.solve(A, X[i, b]) if inverse else A @ X[i, b]
    return out.view(np.float64).ravel().copy()


@pytest.mark.parametrize("name", ["wilson", "clover", "tmwilson", "tmclover",
                                  "ndtmwilson", "ndtmclover"])
@pytest.mark.parametrize("isign", [1, -1])
def test_schur_complement(orc, ops44, name, isign):
    """M_oo psi_o equals the odd row of the full operator (A - D/2) applied to
    (psi_e, psi_o) with psi_e = A_ee^-1 (D_eo psi_o)/2, whose even row is zero.
    The Wilson normalisation divides the full operator by alpha."""
    ops, u, cl, me = ops44
    L = [4, 4, 4, 4]
    op = ops[name]
    mass = 0.1 if "wilson" in name else 0.0
    fields = (u, cl, me, mass)
    nf = op.nflav
    hp = orc.hop_params(-1.0, 1.0, 1.0)
    rng = np.random.default_rng(4)
    po = rng.normal(size=128 * 24 * nf)
    n = 128 * 24

    def hopf(v, cb):
        return np.concatenate([orc.hop(L, u, v[f * n:(f + 1) * n].copy(), isign, cb, hp)
                               for f in range(nf)])
    kind = name
    if nf == 2:
        # flavour-major arrays -> per-site flavour interleave for site_diag_apply
        def sd(v, cb, inv):
            w = v.reshape(2, 128, 24).transpose(1, 0, 2).copy().ravel()
            r = site_diag_apply(kind, w, cb, fields, isign, inv)
            return r.reshape(128, 2, 24).transpose(1, 0, 2).copy().ravel()
    else:
        def sd(v, cb, inv):
            return site_diag_apply(kind, v, cb, fields, isign, inv)
    pe = sd(0.5 * hopf(po, 0), 0, True)
    even_row = sd(pe, 0, False) - 0.5 * hopf(po, 0)
    assert np.abs(even_row).max() < 1e-12
    odd_row = sd(po, 1, False) - 0.5 * hopf(pe, 1)
    got = op(po, isign, 1)
    assert np.abs(got - odd_row).max() < 1e-12


# ----------------------------------------------------------------- solvers
def test_cg_matches_dense(orc, ops44):
    op = ops44[0]["clover"]
    n = 128 * 12
    M = dense(lambda v: op(v, 1, 1), n)
    b = orc.gen_spinor([4, 4, 4, 4], 21, 1)
    x, st = op.cg(b, 1e-12, 500)
    xr = np.linalg.solve(M.conj().T @ M, cplx(b))
    assert st.status == 0
    assert np.abs(cplx(x) - xr).max() / np.abs(xr).max() < 1e-9
    assert st.rsd_sq_final / orc.lib().orc_norm2(b.size, orc.ptr(b)) < 1e-22


def test_bicgstab_matches_dense(orc, ops44):
    op = ops44[0]["tmclover"]
    n = 128 * 12
    M = dense(lambda v: op(v, 1, 1), n)
    b = orc.gen_spinor([4, 4, 4, 4], 22, 1)
    x, st = op.bicgstab(b, 1e-11, 500)
    xr = np.linalg.solve(M, cplx(b))
    assert st.status == 0
    assert np.abs(cplx(x) - xr).max() / np.abs(xr).max() < 1e-9


def test_minvcg_matches_dense(orc, ops44):
    op = ops44[0]["wilson"]
    n = 128 * 12
    M = dense(lambda v: op(v, 1, 1), n)
    MM = M.conj().T @ M
    b = orc.gen_spinor([4, 4, 4, 4], 23, 1)
    shifts = [1e-3, 0.01, 0.1, 1.0]
    x, nit, fin = op.minvcg(b, shifts, [1e-11] * 4, 500)
    assert nit > 0
    for s, sh in enumerate(shifts):
        xr = np.linalg.solve(MM + sh * np.eye(n), cplx(b))
        assert np.abs(cplx(x[s]) - xr).max() / np.abs(xr).max() < 1e-8


def test_cg_nd_operator(orc, ops44):
    op = ops44[0]["ndtmclover"]
    b = np.concatenate([orc.gen_spinor([4, 4, 4, 4], 24, 1), orc.gen_spinor([4, 4, 4, 4], 25, 1)])
    x, st = op.cg(b, 1e-10, 1000)
    r = b - op(op(x, 1, 1), -1, 1)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-9
