"""Multi-right-hand-side kernels and the lockstep multi-RHS CG: every right-hand
side must give exactly the single-RHS result (same arithmetic per site), and
match the oracle; the tuned right-hand-side tile is exercised on a lattice
large enough to be tuned (16^3x64)."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("L,prec,c12,n", [([8, 8, 8, 8], q.F64, False, 5),
                                          ([16, 16, 16, 64], q.F32, True, 6)],
                         ids=["8^4-f64-n5", "16^3x64-f32c12-n6"])
def test_multi_clover_dslash_equals_single(L, prec, c12, n):
    vh = L[0] * L[1] * L[2] * L[3] // 2
    u = oracle.gen_gauge(L, 31, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.05)
    g = q.Geometry(L, prec, c12)
    ud = q.pack_gauge(g, u)
    inv = g.clover().upload(oracle.clover_invert(clov[:vh * 72]))
    A = g.clover().upload(clov[vh * 72:])
    D = q.Dslash(g, t_boundary=-1.0, prec_mass_rho=0.02)
    psi = [g.spinor().upload(oracle.gen_spinor(L, 40 + r, 1), 1) for r in range(n)]
    chi = [g.spinor().upload(oracle.gen_spinor(L, 60 + r, 1), 1) for r in range(n)]
    for isign in (1, -1):
        outs = [g.spinor() for _ in range(n)]
        D.clov_dslash_multi(outs, psi, ud[0], inv, isign, 0)
        for r in range(n):
            one = g.spinor()
            D.clov_dslash(one, psi[r], ud[0], inv, isign, 0)
            assert np.array_equal(outs[r].download(0), one.download(0)), r
        if prec == q.F64:
            ref = oracle.clover_apply(inv.download(),
                                      oracle.hop(L, u, psi[2].download(1), isign, 0,
                                                 oracle.hop_params(-1.0)))
            assert rel(outs[2].download(0), ref) < 1e-13
        # achimbdpsi on the other parity, with chi
        res = [g.spinor() for _ in range(n)]
        D.clov_achimbdpsi_multi(res, outs, chi, ud[1], A, 0.25, isign, 1)
        for r in (0, n - 1):
            one = g.spinor()
            D.clov_achimbdpsi(one, outs[r], chi[r], ud[1], A, 0.25, isign, 1)
            assert np.array_equal(res[r].download(1), one.download(1))
    print(q.tune_summary())


def test_multi_operator_and_cg():
    L = [8, 8, 8, 8]
    vh = 2048
    u = oracle.gen_gauge(L, 32, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.05)
    cl = [clov[:vh * 72], clov[vh * 72:]]
    g = q.Geometry(L, q.F64, False)
    ud = q.pack_gauge(g, u)
    cld = [g.clover().upload(c) for c in cl]
    invd = [g.clover().upload(oracle.clover_invert(c)) for c in cl]
    M = q.Operator(g, q.OP_CLOVER, ud, clov=cld, invclov=invd, t_boundary=-1.0)
    n = 7
    # point sources (spin x colour at one site) and Gaussian sources
    rhs_h = []
    for r in range(4):
        e = np.zeros(vh * 24)
        e[(123 * 24) + 2 * (3 * (r % 4) + r % 3)] = 1.0
        rhs_h.append(e)
    rhs_h += [oracle.gen_spinor(L, 70 + r, 1) for r in range(n - 4)]
    rhs = [g.spinor().upload(h, 1) for h in rhs_h]
    # operator: multi equals single
    outs = [g.spinor() for _ in range(n)]
    M.apply_multi(outs, rhs, 1, 1)
    for r in range(n):
        one = g.spinor()
        M(one, rhs[r], 1, 1)
        assert np.array_equal(outs[r].download(1), one.download(1))
    # lockstep CG vs one solve per right-hand side
    x = [g.spinor() for _ in range(n)]
    sts = q.InvCG(M, 1000).solve_multi(x, rhs, 1e-10)
    hp = oracle.hop_params(-1.0)
    Mo = oracle.Operator(oracle.CLOVER, L, u, hp, clov=cl, invclov=[invd[0].download(), invd[1].download()])
    for r in range(n):
        assert sts[r].status == q.OK
        xs = g.spinor()
        st1 = q.InvCG(M, 1000)(xs, rhs[r], 1e-10)
        assert st1.n_iters == sts[r].n_iters
        assert np.array_equal(xs.download(1), x[r].download(1))
        xh = x[r].download(1)
        res = rhs_h[r] - Mo(Mo(xh, 1, 1), -1, 1)
        assert np.linalg.norm(res) / np.linalg.norm(rhs_h[r]) < 1.5e-10
    norms = q.norm2_multi(g, rhs)
    for r in range(n):
        assert abs(norms[r] - rhs_h[r] @ rhs_h[r]) <= 1e-13 * norms[r]


@pytest.fixture
def forced_rhs_tiling():
    yield
    q.call("qphix_tune_rhs", 0, 0)


@pytest.mark.parametrize("reg,rt", [(1, 4), (2, 2), (2, 6), (4, 4), (4, 8), (8, 8), (8, 16)])
@pytest.mark.parametrize("prec,c12", [(q.F64, False), (q.F32, True), (q.F16, True), (q.BF16, True)],
                         ids=["f64", "f32c12", "f16c12", "bf16c12"])
def test_register_tiled_multi_rhs(forced_rhs_tiling, reg, rt, prec, c12):
    """Each forced tiling (register tile reg, rt sources per launch; n = 5 or,
    for the matrix-core kernel, 11, so that a remainder goes through the
    one-per-work-item kernel) gives every right-hand side its single-RHS
    result: exactly for the vector kernels (same expressions, same rounding),
    to FP32 rounding for the FP32 matrix-core kernel (its dot products are
    summed in the matrix unit's order) and to 3e-3 for the FP16 one (its
    operands, the projected half spinors and the links with their rebuilt
    third row, are rounded to half), for the Wilson, clover and
    twisted-mass-clover members of the kernel family. FP64 and BF16 fields
    have no matrix-core form."""
    L = [4, 4, 4, 8]
    mfma = reg == 8
    n = 11 if mfma else 5
    if mfma and prec not in (q.F32, q.F16):
        q.call("qphix_tune_rhs", reg, rt)
        g = q.Geometry(L, prec, c12)
        D = q.Dslash(g, t_boundary=-1.0)
        ud = q.pack_gauge(g, oracle.gen_gauge(L, 33, 0.4))
        ps = [g.spinor() for _ in range(n)]
        with pytest.raises(q.QphixError):
            D.dslash_multi([g.spinor() for _ in range(n)], ps, ud[0], 1, 0)
        return
    vh = L[0] * L[1] * L[2] * L[3] // 2
    u = oracle.gen_gauge(L, 33, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.05)
    g = q.Geometry(L, prec, c12)
    ud = q.pack_gauge(g, u)
    inv = g.clover().upload(oracle.clover_invert(clov[:vh * 72]))
    A = g.clover().upload(clov[vh * 72:])
    f = g.full_clover().upload(oracle.full_clover_tm_inverse(clov[:vh * 72], 0.03))
    q.call("qphix_tune_rhs", reg, rt)
    D = q.Dslash(g, t_boundary=-1.0, prec_mass_rho=0.02)
    psi = [g.spinor().upload(oracle.gen_spinor(L, 70 + r, 1), 1) for r in range(n)]
    chi = [g.spinor().upload(oracle.gen_spinor(L, 80 + r, 0), 0) for r in range(n)]

    def check(multi, single, cb):
        outs = [g.spinor() for _ in range(n)]
        multi(outs)
        for r in range(n):
            one = g.spinor()
            single(one, r)
            got, exp = outs[r].download(cb), one.download(cb)
            if mfma:
                tol = 2e-6 if prec == q.F32 else 3e-3
                assert np.linalg.norm(got - exp) <= tol * np.linalg.norm(exp), (r, reg, rt)
            else:
                assert np.array_equal(got, exp), (r, reg, rt)

    for isign in (1, -1):
        check(lambda o: D.clov_dslash_multi(o, psi, ud[0], inv, isign, 0),
              lambda o, r: D.clov_dslash(o, psi[r], ud[0], inv, isign, 0), 0)
        check(lambda o: D.clov_achimbdpsi_multi(o, psi, chi, ud[0], A, 0.25, isign, 0),
              lambda o, r: D.clov_achimbdpsi(o, psi[r], chi[r], ud[0], A, 0.25, isign, 0), 0)
        check(lambda o: D.dslash_multi(o, psi, ud[0], isign, 0),
              lambda o, r: D.dslash(o, psi[r], ud[0], isign, 0), 0)
    # twisted-mass clover Schur operator (full-block inverse kernel, A chi + i rho g5 chi)
    fm = g.full_clover().upload(oracle.full_clover_tm_inverse(clov[:vh * 72], -0.03))
    M = q.Operator(g, q.OP_TMCLOVER, ud, clov=[None, A], fullclov=[[f, fm], None], twisted_mass=0.03,
                   prec_mass_rho=0.03, t_boundary=-1.0)
    for isign in (1, -1):
        check(lambda o: M.apply_multi(o, psi, isign, 1), lambda o, r: M(o, psi[r], isign, 1), 1)


@pytest.mark.parametrize("reg,rt", [(2, 2), (4, 4)])
def test_f16_subnormal_halves_tile_equals_single(forced_rhs_tiling, reg, rt):
    """FP16 storage with components far below their site's largest one (1e-6 of
    it, subnormal halves once normalised by the per-site scale): the register
    tiles still give every source its single-source result bit for bit, and
    both agree with the oracle on the values the device holds. Such halves are
    planted on purpose: how their half-to-single conversions are compiled is
    where the single-source and tiled kernels could part."""
    L = [4, 4, 4, 8]
    n = 5
    vh = L[0] * L[1] * L[2] * L[3] // 2
    u = oracle.gen_gauge(L, 35, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.05)
    inv_h = oracle.clover_invert(clov[:vh * 72])
    g = q.Geometry(L, q.F16, True)
    ud = q.pack_gauge(g, u)
    inv = g.clover().upload(inv_h)
    A = g.clover().upload(clov[vh * 72:])
    rng = np.random.default_rng(7)

    def planted(seed, cb):
        h = oracle.gen_spinor(L, seed, cb).reshape(vh, 24)
        h[rng.random((vh, 24)) < 0.3] *= 1e-6
        return h.ravel()

    psi = [g.spinor().upload(planted(90 + r, 1), 1) for r in range(n)]
    chi = [g.spinor().upload(planted(100 + r, 0), 0) for r in range(n)]
    held = psi[0].download(1)
    tiny = np.abs(held.reshape(vh, 24)) / np.abs(held.reshape(vh, 24)).max(axis=1, keepdims=True)
    assert np.count_nonzero((tiny > 0) & (tiny < 2.0 ** -14)) > 100  # subnormal halves are stored
    q.call("qphix_tune_rhs", reg, rt)
    D = q.Dslash(g, t_boundary=-1.0)
    hp = oracle.hop_params(-1.0)
    for isign in (1, -1):
        for kind in ("clover", "achimbdpsi", "wilson"):
            outs = [g.spinor() for _ in range(n)]
            if kind == "clover":
                D.clov_dslash_multi(outs, psi, ud[0], inv, isign, 0)
            elif kind == "achimbdpsi":
                D.clov_achimbdpsi_multi(outs, psi, chi, ud[0], A, 0.25, isign, 0)
            else:
                D.dslash_multi(outs, psi, ud[0], isign, 0)
            for r in range(n):
                one = g.spinor()
                if kind == "clover":
                    D.clov_dslash(one, psi[r], ud[0], inv, isign, 0)
                elif kind == "achimbdpsi":
                    D.clov_achimbdpsi(one, psi[r], chi[r], ud[0], A, 0.25, isign, 0)
                else:
                    D.dslash(one, psi[r], ud[0], isign, 0)
                got = one.download(0)
                assert np.array_equal(outs[r].download(0), got), (kind, isign, r)
            hop = oracle.hop(L, u, psi[n - 1].download(1), isign, 0, hp)
            if kind == "clover":
                ref = oracle.clover_apply(inv_h, hop)
            elif kind == "achimbdpsi":
                ref = oracle.clover_apply(clov[vh * 72:], chi[n - 1].download(0)) - 0.25 * hop
            else:
                ref = hop
            assert rel(got, ref) < 5e-3, (kind, isign, rel(got, ref))
