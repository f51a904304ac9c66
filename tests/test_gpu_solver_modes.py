"""BiCGStab and CG with their scalars in device memory (mode 1, eager
launches; mode 2, each iteration replayed from a HIP graph) against the loops
with host scalars (mode 0, the reference's control flow, invbicgstab.h:150-256,
invcg.h:293-343): solutions, iteration counts, final residuals, counters and
statuses bit-identical for FP64, FP32, FP16 and BF16 storage, converging,
stopped at MaxIters and (BiCGStab) breaking down; the FP64 solutions checked
against the oracle's operator; the mixed-precision defect correction with the
device-scalar inner solver equal to the one with the host-scalar inner solver."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L8 = [8, 8, 8, 8]
VH = 2048
MASS = 0.1


@pytest.fixture(scope="module")
def fields():
    u = oracle.gen_gauge(L8, 7, 0.4)
    clov = oracle.clover_build(L8, u, 1.0, MASS)
    clov = [clov[:VH * 72], clov[VH * 72:]]
    inv = [oracle.clover_invert(c) for c in clov]
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    out = {"orc": oracle.Operator(oracle.CLOVER, L8, u, hp, clov=clov, invclov=inv, rho=0.01)}
    for prec, c12 in ((q.F64, False), (q.F32, True), (q.F16, True), (q.BF16, True)):
        g = q.Geometry(L8, prec, c12)
        ud = q.pack_gauge(g, u)
        cl = [g.clover().upload(c) for c in clov]
        iv = [g.clover().upload(c) for c in inv]
        out[prec] = (g, q.Operator(g, q.OP_CLOVER, ud, clov=cl, invclov=iv, prec_mass_rho=0.01,
                                   t_boundary=-1.0), (ud, cl, iv))
    return out


def solve(M, g, b, mode, max_iters, tol):
    x = g.spinor()
    st = q.InvBiCGStab(M, max_iters).set_mode(mode)(x, b, tol)
    return (st.status, st.n_iters, st.rsd_sq_final, st.site_flops, st.mv_apps), x.download(1)


@pytest.mark.parametrize("prec", [q.F64, q.F32, q.F16, q.BF16], ids=["f64", "f32", "f16", "bf16"])
def test_bicgstab_device_scalars_bit_identical(fields, prec):
    g, M, _ = fields[prec]
    bh = oracle.gen_spinor(L8, 170, 1)
    b = g.spinor().upload(bh, 1)
    tol = {q.F64: 1e-11, q.F32: 1e-6, q.F16: 1e-2, q.BF16: 1e-2}[prec]
    res = [solve(M, g, b, mode, 1000, tol) for mode in (0, 1, 2, 2)]
    assert res[0][0][0] == q.OK and res[0][0][1] > 1
    for stats, x in res[1:]:
        assert stats == res[0][0]
        assert np.array_equal(x, res[0][1])
    if prec == q.F64:
        B = b.download(1)
        r = B - fields["orc"](res[0][1], 1, 1)
        assert np.linalg.norm(r) / np.linalg.norm(B) < 1.5e-11


def test_bicgstab_device_scalars_max_iters_and_breakdown(fields):
    g, M, _ = fields[q.F64]
    b = g.spinor().upload(oracle.gen_spinor(L8, 171, 1), 1)
    res = [solve(M, g, b, mode, 4, 1e-14) for mode in (0, 1, 2)]
    assert res[0][0][0] == q.ENOCONV
    for stats, x in res[1:]:
        assert stats == res[0][0]
        assert np.array_equal(x, res[0][1])
    # a zero source: <r0, r> = 0 on the first iteration is a breakdown
    z = g.spinor()
    res = [solve(M, g, z, mode, 100, 1e-10) for mode in (0, 1, 2)]
    assert res[0][0][0] == q.EBREAKDOWN
    for stats, _ in res[1:]:
        assert stats == res[0][0]


@pytest.mark.parametrize("inner_prec", [q.F16, q.BF16], ids=["f16", "bf16"])
def test_mixed_solve_with_device_scalar_inner(fields, inner_prec):
    g, M64, _ = fields[q.F64]
    gi, Mi, _ = fields[inner_prec]
    bh = oracle.gen_spinor(L8, 172, 1)
    b = g.spinor().upload(bh, 1)
    out = []
    for mode in (0, 2):
        x = g.spinor()
        inner = q.InvBiCGStab(Mi, 500).set_mode(mode)
        st = q.InvRichardsonMultiPrec(M64, inner, 1e-2, 200)(x, b, 1e-10)
        assert st.status == q.OK
        out.append(((st.n_iters, st.rsd_sq_final, st.site_flops, st.mv_apps), x.download(1)))
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1], out[1][1])
    r = bh - fields["orc"](out[1][1], 1, 1)
    assert np.linalg.norm(r) / np.linalg.norm(bh) < 1.5e-10


def test_graph_mode_tunes_before_capture():
    """A solve from x = 0 on a lattice whose hop kernels were never tuned (the
    first operator application happens inside the loop, and tunes): the graph
    mode runs that first iteration eagerly and captures the next one; the
    result equals the host-scalar loop's bit for bit."""
    L = [16, 16, 32, 32]  # 131072 sites per parity: autotuned
    g64 = q.Geometry(L, q.F64, False)
    u64 = q.gen_gauge(g64, 41, 0.4)
    A = q.clover_build(g64, u64, 1.0, MASS, 1, q.CLOV_A)
    Ai = q.clover_build(g64, u64, 1.0, MASS, 0, q.CLOV_AINV)
    g = q.Geometry(L, q.F32, True)
    M = q.Operator(g, q.OP_CLOVER, q.convert_gauge(g, g64, u64), clov=[None, q.convert_clover(g, g64, A)],
                   invclov=[q.convert_clover(g, g64, Ai), None], t_boundary=-1.0)
    b = q.gen_spinor(g, 42, 1)
    out = []
    for mode in (2, 0):
        x = g.spinor()
        st = q.InvBiCGStab(M, 500).set_mode(mode)(x, b, 1e-6)
        out.append(((st.status, st.n_iters, st.rsd_sq_final, st.site_flops, st.mv_apps), x))
    assert out[0][0] == out[1][0] and out[0][0][0] == q.OK
    assert np.array_equal(out[0][1].download(1), out[1][1].download(1))


def cg_solve(M, g, b, mode, max_iters, tol):
    x = g.spinor()
    st = q.InvCG(M, max_iters).set_mode(mode)(x, b, tol)
    return (st.status, st.n_iters, st.rsd_sq_final, st.site_flops, st.mv_apps), x.download(1)


@pytest.mark.parametrize("prec", [q.F64, q.F32, q.F16, q.BF16], ids=["f64", "f32", "f16", "bf16"])
def test_cg_device_scalars_bit_identical(fields, prec):
    g, M, _ = fields[prec]
    bh = oracle.gen_spinor(L8, 180, 1)
    b = g.spinor().upload(bh, 1)
    tol = {q.F64: 1e-11, q.F32: 1e-6, q.F16: 1e-2, q.BF16: 1e-2}[prec]
    res = [cg_solve(M, g, b, mode, 1000, tol) for mode in (0, 1, 2, 2)]
    assert res[0][0][0] == q.OK and res[0][0][1] > 1
    for stats, x in res[1:]:
        assert stats == res[0][0]
        assert np.array_equal(x, res[0][1])
    if prec == q.F64:
        B = b.download(1)
        Mo = fields["orc"]
        r = B - Mo(Mo(res[0][1], 1, 1), -1, 1)
        assert np.linalg.norm(r) / np.linalg.norm(B) < 1.5e-11
    # stopped at MaxIters
    res = [cg_solve(M, g, b, mode, 3, 1e-14) for mode in (0, 1, 2)]
    assert res[0][0][0] == q.ENOCONV
    for stats, x in res[1:]:
        assert stats == res[0][0]
        assert np.array_equal(x, res[0][1])
