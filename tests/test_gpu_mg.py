"""Multigrid-facing pieces of the operator API (linearOp.h, wilson.h,
clover.h:174-251), dslashDir, InvMR / InvMRSmoother and the unpreconditioned
solve wrapper (unprec_solver_wrapper.h, abs_solver.h:271-376), checked against
the oracle: kernels to the Dslash tolerances, solves by the residual of the
full unpreconditioned operator recomputed in FP64."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L = [8, 8, 8, 8]
VH = 2048
MASS = 0.1


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def world():
    u = oracle.gen_gauge(L, 61, 0.4)
    clov = oracle.clover_build(L, u, 1.0, MASS)
    cl = [clov[:VH * 72], clov[VH * 72:]]
    inv = [oracle.clover_invert(c) for c in cl]
    return u, cl, inv


def device(world, prec, c12, kind=q.OP_CLOVER):
    u, cl, inv = world
    g = q.Geometry(L, prec, c12)
    ud = q.pack_gauge(g, u)
    cld = [g.clover().upload(c) for c in cl]
    invd = [g.clover().upload(c) for c in inv]
    kw = dict(t_boundary=-1.0, aniso_t=0.8)
    if kind == q.OP_CLOVER:
        M = q.Operator(g, kind, ud, clov=cld, invclov=invd, **kw)
    else:
        M = q.Operator(g, kind, ud, mass=MASS, **kw)
    return g, M, (ud, cld, invd)


HP = oracle.hop_params(-1.0, 1.0, 0.8)


@pytest.mark.parametrize("prec,c12,tol", [(q.F64, False, 1e-13), (q.F32, True, 2e-6)], ids=["f64", "f32c12"])
def test_dslash_dir(world, prec, c12, tol):
    u = world[0]
    g = q.Geometry(L, prec, c12)
    ud = q.pack_gauge(g, u)
    D = q.Dslash(g, t_boundary=-1.0, aniso_t=0.8)
    for cb in (0, 1):
        psi_h = oracle.gen_spinor(L, 70 + cb, 1 - cb)
        psi = g.spinor().upload(psi_h, 1 - cb)
        for d in range(8):
            out = g.spinor()
            D.dslash_dir(out, psi, ud[cb], cb, d)
            ref = oracle.dslash_dir(L, u, psi_h, cb, d, HP)
            assert rel(out.download(cb), ref) < tol, (cb, d)


def test_clover_operator_pieces(world):
    u, cl, inv = world
    g, M, keep = device(world, q.F64, False)
    inp = [oracle.gen_spinor(L, 80 + c, c) for c in (0, 1)]
    din = [g.spinor().upload(inp[c], c) for c in (0, 1)]
    for isign in (1, -1):
        for t in (0, 1):
            r = g.spinor()
            M.m_offdiag(r, din[1 - t], isign, t)
            assert rel(r.download(t), -0.5 * oracle.hop(L, u, inp[1 - t], isign, t, HP)) < 1e-13
        # M_unprec = A in - 1/2 D in on both checkerboards
        res = [g.spinor(), g.spinor()]
        M.m_unprec(res, din, isign)
        for c in (0, 1):
            ref = oracle.clover_apply(cl[c], inp[c]) - 0.5 * oracle.hop(L, u, inp[1 - c], isign, c, HP)
            assert rel(res[c].download(c), ref) < 1e-13
    for c in (0, 1):
        r = g.spinor()
        M.m_diag_inv(r, din[c], 1, c)
        assert rel(r.download(c), oracle.clover_apply(inv[c], inp[c])) < 1e-13
    res = [g.spinor(), g.spinor()]
    M.m_diag(res, din, 1)
    for c in (0, 1):
        assert rel(res[c].download(c), oracle.clover_apply(cl[c], inp[c])) < 1e-13
    for d in (0, 5, 7):
        res = [g.spinor(), g.spinor()]
        M.dslash_dir(res, din, d)
        for c in (0, 1):
            assert rel(res[c].download(c), oracle.dslash_dir(L, u, inp[1 - c], c, d, HP)) < 1e-13


def test_wilson_pieces_and_unsupported(world):
    g, M, keep = device(world, q.F64, False, q.OP_WILSON)
    x = oracle.gen_spinor(L, 90, 0)
    r = g.spinor()
    M.m_diag_inv(r, g.spinor().upload(x, 0), 1, 0)
    assert rel(r.download(0), x / (4 + MASS)) < 1e-15
    with pytest.raises(q.QphixError):
        M.m_diag_inv(r, g.spinor(), 1, 1)
    with pytest.raises(q.QphixError):
        M.m_unprec([g.spinor(), g.spinor()], [g.spinor(), g.spinor()], 1)
    Mt = q.Operator(g, q.OP_TMWILSON, keep[0], mass=MASS, twisted_mass=0.05, t_boundary=-1.0)
    with pytest.raises(q.QphixError):
        Mt.m_offdiag(r, g.spinor(), 1, 1)


def test_mr_step_and_solve(world):
    u = world[0]
    g, M, keep = device(world, q.F64, False, q.OP_WILSON)
    Mo = oracle.Operator(oracle.WILSON, L, u, HP, mass=MASS)
    b_h = oracle.gen_spinor(L, 95, 1)
    b = g.spinor().upload(b_h, 1)
    # one smoother step from x = 0: x = omega <Mb, b> / |Mb|^2 b  (invmr.h:153-176)
    omega = 1.1
    x = g.spinor()
    st = q.InvMRSmoother(M, 1, omega)(x, b, 1e-10)
    assert st.n_iters == 1 and st.rsd_sq_final == -1
    mb = Mo(b_h, 1, 1)
    mbc, bc = mb[0::2] + 1j * mb[1::2], b_h[0::2] + 1j * b_h[1::2]
    a = omega * np.vdot(mbc, bc) / np.vdot(mbc, mbc).real
    xe = a * bc
    xg = x.download(1)
    assert np.abs(xg[0::2] + 1j * xg[1::2] - xe).max() < 1e-13 * np.abs(xe).max()
    # InvMR to 1e-10, true residual in FP64
    x = g.spinor()
    st = q.InvMR(M, 500, 1.0)(x, b, 1e-10)
    assert st.status == q.OK and st.n_iters > 1
    assert rel(Mo(x.download(1), 1, 1), b_h) < 2e-10


@pytest.mark.parametrize("solve_cb", [1, 0])
def test_unprec_solve(world, solve_cb):
    u, cl, inv = world
    g, M, keep = device(world, q.F64, False)
    b = [oracle.gen_spinor(L, 100 + c, c) for c in (0, 1)]
    db = [g.spinor().upload(b[c], c) for c in (0, 1)]
    x = [g.spinor(), g.spinor()]
    st = q.unprec_solve(q.InvBiCGStab(M, 1000), x, db, 1e-11, 1, solve_cb)
    assert st.status == q.OK
    xs = [x[c].download(c) for c in (0, 1)]
    num = den = 0.0
    for c in (0, 1):
        r = b[c] - (oracle.clover_apply(cl[c], xs[c]) - 0.5 * oracle.hop(L, u, xs[1 - c], 1, c, HP))
        num += r @ r
        den += b[c] @ b[c]
    assert (num / den) ** 0.5 < 1e-9
