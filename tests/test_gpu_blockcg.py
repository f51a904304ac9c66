"""Block CG on M^+M (point and Gaussian sources sharing one block Krylov space)
and the block BLAS it runs on: Gram matrices and matrix combinations against
numpy on downloaded fields, solutions by their true residual recomputed in FP64
with the oracle operator, by agreement with one CG solve per right-hand side,
and by taking fewer iterations than the single-RHS solves."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L = [8, 8, 8, 8]
VH = 2048


def setup(prec, c12, seed=41):
    u = oracle.gen_gauge(L, seed, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.05)
    cl = [clov[:VH * 72], clov[VH * 72:]]
    inv = [oracle.clover_invert(c) for c in cl]
    g = q.Geometry(L, prec, c12)
    ud = q.pack_gauge(g, u)
    cld = [g.clover().upload(c) for c in cl]
    invd = [g.clover().upload(c) for c in inv]
    M = q.Operator(g, q.OP_CLOVER, ud, clov=cld, invclov=invd, t_boundary=-1.0)
    Mo = oracle.Operator(oracle.CLOVER, L, u, oracle.hop_params(-1.0), clov=cl,
                         invclov=[invd[0].download(), invd[1].download()])
    return g, M, Mo, (ud, cld, invd)


def sources(n_point, n_gauss):
    out = []
    for r in range(n_point):  # unit source on spin r // 3, colour r % 3 at one site
        e = np.zeros(VH * 24)
        e[(321 * 24) + 2 * (3 * ((r // 3) % 4) + r % 3)] = 1.0
        out.append(e)
    return out + [oracle.gen_spinor(L, 90 + r, 1) for r in range(n_gauss)]


@pytest.mark.parametrize("prec,tol,true_tol", [(q.F64, 1e-10, 2e-10), (q.F32, 1e-6, 5e-6)],
                         ids=["f64", "f32"])
def test_block_cg(prec, tol, true_tol):
    g, M, Mo, keep = setup(prec, False)
    rhs_h = sources(4, 4)
    n = len(rhs_h)
    rhs = [g.spinor().upload(h, 1) for h in rhs_h]
    x = [g.spinor() for _ in range(n)]
    sts = q.InvBlockCG(M, 500).solve_multi(x, rhs, tol)
    single_iters = []
    for r in range(n):
        assert sts[r].status == q.OK, (r, sts[r].n_iters, sts[r].rsd_sq_final)
        assert sts[r].rsd_sq_final <= tol * tol
        xh = x[r].download(1)
        res = rhs_h[r] - Mo(Mo(xh, 1, 1), -1, 1)
        assert np.linalg.norm(res) / np.linalg.norm(rhs_h[r]) < true_tol, r
        if prec == q.F64 and r in (0, 5):
            xs = g.spinor()
            st1 = q.InvCG(M, 1000)(xs, rhs[r], tol)
            single_iters.append(st1.n_iters)
            assert np.linalg.norm(xs.download(1) - xh) / np.linalg.norm(xh) < 1e-8
    if single_iters:
        assert sts[0].n_iters < min(single_iters)


def test_block_blas_gram_and_combine():
    g, M, Mo, keep = setup(q.F64, False)
    n = 5
    vs = [oracle.gen_spinor(L, 200 + r, 1) for r in range(n)]
    ws = [oracle.gen_spinor(L, 300 + r, 1) for r in range(n)]
    dv = [g.spinor().upload(v, 1) for v in vs]
    dw = [g.spinor().upload(w, 1) for w in ws]
    G = q.block_gram(g, dv, dw)
    cv = [v[0::2] + 1j * v[1::2] for v in vs]
    cw = [w[0::2] + 1j * w[1::2] for w in ws]
    ref = np.array([[np.vdot(a, b) for b in cw] for a in cv])
    assert np.abs(G - ref).max() <= 1e-12 * np.abs(ref).max()
    rng = np.random.default_rng(5)
    Mx = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    q.block_combine(g, dv, Mx, dw, dw)  # w_j += sum_i v_i M_ij, in place
    for j in range(n):
        got = dw[j].download(1)
        exp = cw[j] + sum(cv[i] * Mx[i, j] for i in range(n))
        assert np.abs(got[0::2] + 1j * got[1::2] - exp).max() < 1e-12


def test_block_cg_rejects():
    g16, M16, _, keep = setup(q.F16, True)
    rhs = [q.gen_spinor(g16, 1, 1)]
    with pytest.raises(q.QphixError):
        q.InvBlockCG(M16, 10).solve_multi([g16.spinor()], rhs, 1e-6)
