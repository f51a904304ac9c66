"""Batched solves of systems on lattices of different sizes (one persistent
launch per step, device FP64 scalars, HIP-graph replay): each solution must
match the per-system solver and have the requested true residual, recomputed
by the oracle in FP64; graph replay must give exactly the eager result."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

LATS = [[8, 8, 8, 8], [4, 4, 4, 8], [6, 6, 6, 6]]
MU = [0.004, 0.007, 0.009]


def build(L, seed, mu, prec, c12):
    vh = L[0] * L[1] * L[2] * L[3] // 2
    u = oracle.gen_gauge(L, seed, 0.4)
    clov = oracle.clover_build(L, u, 1.0, 0.0322581)
    cl = [clov[:vh * 72], clov[vh * 72:]]
    fcl = [[oracle.full_clover_tm_inverse(c, mu), oracle.full_clover_tm_inverse(c, -mu)] for c in cl]
    g = q.Geometry(L, prec, c12)
    ud = q.pack_gauge(g, u)
    cld = [g.clover().upload(c) for c in cl]
    fd = [[g.full_clover().upload(f) for f in pair] for pair in fcl]
    op = q.Operator(g, q.OP_TMCLOVER, ud, clov=cld, fullclov=fd, twisted_mass=mu,
                    prec_mass_rho=mu, t_boundary=-1.0)
    orc = oracle.Operator(oracle.TMCLOVER, L, u, oracle.hop_params(-1.0), clov=cl, fcl=fcl,
                          mu=mu, rho=mu)
    return g, op, orc, (ud, cld, fd)


@pytest.fixture(scope="module")
def systems():
    out = []
    for i, L in enumerate(LATS):
        out.append(build(L, 50 + i, MU[i], q.F64, False))
    return out


def true_resid(orc, x, b):
    r = b - orc(orc(x, 1, 1), -1, 1)
    return np.linalg.norm(r) / np.linalg.norm(b)


def test_batch_cg_matches_single_and_graph(systems):
    ops = [s[1] for s in systems for _ in range(2)]
    gs = [s[0] for s in systems for _ in range(2)]
    orcs = [s[2] for s in systems for _ in range(2)]
    rhs_h = [oracle.gen_spinor(LATS[i // 2], 80 + i, 1) for i in range(len(ops))]
    rhs = [g.spinor().upload(h, 1) for g, h in zip(gs, rhs_h)]
    B = q.Batch(ops)
    res = {}
    for steps in (0, 4):
        x = [g.spinor() for g in gs]
        sts = B.cg(x, rhs, 1e-10, 2000, graph_steps=steps)
        res[steps] = [v.download(1) for v in x]
        for i, st in enumerate(sts):
            assert st.status == q.OK, (i, st.n_iters)
            assert true_resid(orcs[i], res[steps][i], rhs_h[i]) < 2e-10
    for a, b in zip(res[0], res[4]):
        assert np.array_equal(a, b)
    # the per-system solver on the same system
    for i in (0, 3):
        xs = gs[i].spinor()
        st1 = q.InvCG(ops[i], 2000)(xs, rhs[i], 1e-10)
        assert st1.status == q.OK
        xs_h = xs.download(1)
        assert np.linalg.norm(xs_h - res[0][i]) / np.linalg.norm(xs_h) < 1e-8


def test_batch_richardson_f16_inner(systems):
    inner = [build(L, 50 + i, MU[i], q.F16, True) for i, L in enumerate(LATS)]
    ops_in = [s[1] for s in inner for _ in range(2)]
    ops_out = [s[1] for s in systems for _ in range(2)]
    gs = [s[0] for s in systems for _ in range(2)]
    orcs = [s[2] for s in systems for _ in range(2)]
    rhs_h = [oracle.gen_spinor(LATS[i // 2], 90 + i, 1) for i in range(len(ops_in))]
    rhs = [g.spinor().upload(h, 1) for g, h in zip(gs, rhs_h)]
    x = [g.spinor() for g in gs]
    B = q.Batch(ops_in)
    sts = B.richardson(ops_out, x, rhs, 1e-10, delta=1e-2, graph_steps=8)
    for i, st in enumerate(sts):
        assert st.status == q.OK, (i, st.n_iters, st.rsd_sq_final)
        assert true_resid(orcs[i], x[i].download(1), rhs_h[i]) < 2e-10


@pytest.fixture
def force_batch_tile():
    import os
    os.environ["QPHIX_BATCH_TILE"] = "1"
    yield
    del os.environ["QPHIX_BATCH_TILE"]


def test_batch_register_tile_matches_untiled(systems, force_batch_tile):
    """Four right-hand sides per lattice, grouped (each pair of systems shares
    its operator, so the hops run as a register tile of two sources) and
    interleaved (no pair shares an operator: one source per work-item): both
    orders give bit-identical solutions, and the true residuals hold."""
    n_rhs = 4
    grouped = [(li, r) for li in range(len(LATS)) for r in range(n_rhs)]
    inter = [(li, r) for r in range(n_rhs) for li in range(len(LATS))]
    rhs_h = {(li, r): oracle.gen_spinor(LATS[li], 300 + 10 * li + r, 1) for li, r in grouped}
    out = {}
    for name, order in (("grouped", grouped), ("interleaved", inter)):
        ops = [systems[li][1] for li, _ in order]
        gs = [systems[li][0] for li, _ in order]
        rhs = [g.spinor().upload(rhs_h[key], 1) for g, key in zip(gs, order)]
        x = [g.spinor() for g in gs]
        sts = q.Batch(ops).cg(x, rhs, 1e-10, 2000, graph_steps=4)
        for st, key in zip(sts, order):
            assert st.status == q.OK, (name, key)
        out[name] = {key: v.download(1) for key, v in zip(order, x)}
    for key in grouped:
        li = key[0]
        assert np.array_equal(out["grouped"][key], out["interleaved"][key]), key
        assert true_resid(systems[li][2], out["grouped"][key], rhs_h[key]) < 2e-10
