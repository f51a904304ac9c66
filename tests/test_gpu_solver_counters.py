"""The solvers' integer outputs (site_flops, mv_apps) against the reference's
counting, evaluated at the returned n_iters.  The reference counts BLAS site
flops in units of 12 * num_flav and every operator application, including the
initial one on a zero start:
  CG        invcg.h:180,210,224,294,308,337,416 -> converged in 0-based iteration k:
            mv_apps = 2k + 4, site_flops = num_flav * (312 + 240 k)
  BiCGStab  invbicgstab.h:158-240 -> n_iters = iterations + 1:
            mv_apps = 2 n - 1, site_flops = num_flav * (72 + 960 (n - 1))
  MR        invmr.h:95-182 (terminating on the residual): mv_apps = 1 + k,
            site_flops = num_flav * (120 + 384 k)
  MInvCG    minvcg.h:139-307: 47 + 47 + 95 + 24 n_shift, then per iteration 190
            plus 120 per shift not yet converged (not scaled by the flavours)
  Richardson inv_richardson_multiprec.h:55-228: outer 47 + 71 + 96 per outer
            iteration + 71, plus num_flav times the inner solvers' site flops
Solver callers compute GFLOP/s from these (tests/testDslashFull.cc:582-584).
"""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L8 = [8, 8, 8, 8]
VH = 2048
MASS, MU, EPS = 0.1, 0.07, 0.03


@pytest.fixture(scope="module")
def fields():
    u = oracle.gen_gauge(L8, 4, 0.4)
    clov = oracle.clover_build(L8, u, 1.0, MASS)
    clov = [clov[:VH * 72], clov[VH * 72:]]
    inv = [oracle.clover_invert(c) for c in clov]
    out = {}
    for prec, c12 in ((q.F64, False), (q.F32, True)):
        g = q.Geometry(L8, prec, c12)
        ud = q.pack_gauge(g, u)
        cl = [g.clover().upload(c) for c in clov]
        iv = [g.clover().upload(c) for c in inv]
        out[prec] = dict(
            g=g, keep=(ud, cl, iv),
            wilson=q.Operator(g, q.OP_WILSON, ud, mass=MASS, t_boundary=-1.0),
            clover=q.Operator(g, q.OP_CLOVER, ud, clov=cl, invclov=iv, t_boundary=-1.0),
            ndtmwilson=q.Operator(g, q.OP_NDTMWILSON, ud, mass=MASS, twisted_mass=MU, epsilon=EPS,
                                  t_boundary=-1.0))
    return out


def rhs(g, seed, nf=1):
    return [g.spinor().upload(oracle.gen_spinor(L8, seed + f, 1), 1) for f in range(nf)]


def cg_counts(k, nf):
    return 2 * k + 4, nf * (312 + 240 * k)


def bicgstab_counts(n, nf):
    return 2 * n - 1, nf * (72 + 960 * (n - 1))


@pytest.mark.parametrize("name,nf", [("wilson", 1), ("ndtmwilson", 2)])
def test_cg_counters(fields, name, nf):
    f = fields[q.F64]
    g = f["g"]
    b = rhs(g, 300, nf)
    for x0 in (None, 0.25):  # zero start, and a nonzero start guess
        x = [g.spinor() for _ in range(nf)]
        if x0 is not None:
            for xf in x:
                q.copy(g, xf, b[0])
        st = q.InvCG(f[name], 1000)(x, b, 1e-10)
        assert st.status == q.OK and st.n_iters > 3
        assert (st.mv_apps, st.site_flops) == cg_counts(st.n_iters, nf), (x0, st.n_iters)


def test_cg_not_converged_counters(fields):
    f = fields[q.F64]
    st = q.InvCG(f["wilson"], 5)(f["g"].spinor(), rhs(f["g"], 301)[0], 1e-14)
    assert st.status == q.ENOCONV and st.n_iters == 5
    # the loop runs k = 0 .. MaxIters, each iteration with aypx (invcg.h:281-417)
    assert st.mv_apps == 2 + 2 * 6 and st.site_flops == 120 + 240 * 6


def test_cg_multi_counters(fields):
    """Lockstep CG over several right-hand sides counts per right-hand side as CG."""
    f = fields[q.F64]
    g = f["g"]
    bs = [rhs(g, 310 + k)[0] for k in range(3)]
    xs = [g.spinor() for _ in bs]
    sts = q.InvCG(f["wilson"], 1000).solve_multi(xs, bs, 1e-9)
    for st in sts:
        assert st.status == q.OK
        assert (st.mv_apps, st.site_flops) == cg_counts(st.n_iters, 1)


@pytest.mark.parametrize("prec", [q.F64, q.F32], ids=["f64", "f32"])
def test_bicgstab_counters(fields, prec):
    f = fields[prec]
    g = f["g"]
    b = rhs(g, 320)
    st = q.InvBiCGStab(f["clover"], 1000)(g.spinor(), b, 1e-9 if prec == q.F64 else 1e-6)
    assert st.status == q.OK and st.n_iters > 3
    assert (st.mv_apps, st.site_flops) == bicgstab_counts(st.n_iters, 1)


def test_mr_counters(fields):
    f = fields[q.F64]
    g = f["g"]
    st = q.InvMR(f["clover"], 500)(g.spinor(), rhs(g, 330), 1e-6)
    assert st.status == q.OK and st.n_iters > 3
    assert st.mv_apps == 1 + st.n_iters
    assert st.site_flops == 120 + 384 * st.n_iters


def test_minvcg_counters(fields):
    """Shift s converges in the multishift solve at the iteration n_s at which
    the one-shift solve with that shift stops (the unshifted recurrences are
    shared), so site_flops = 189 + 24 S + 190 n + 120 sum_s n_s, n = max n_s."""
    f = fields[q.F64]
    g = f["g"]
    b = rhs(g, 340)[0]
    shifts = [1e-3, 1e-2, 0.1, 1.0]
    tol = [1e-9] * len(shifts)
    ns = []
    for s in shifts:
        m = q.MInvCG(f["clover"], 2000, 1)
        n, _, mv = m([[g.spinor()]], b, [s], [1e-9])
        assert mv == 2 + 2 * n
        assert m.site_flops == 189 + 24 + 310 * n
        ns.append(n)
    m = q.MInvCG(f["clover"], 2000, 8)
    n, _, mv = m([[g.spinor()] for _ in shifts], b, shifts, tol)
    assert n == max(ns)
    assert mv == 2 + 2 * n
    assert m.site_flops == 189 + 24 * len(shifts) + 190 * n + 120 * sum(ns)


def test_richardson_counters(fields):
    """FP64 Richardson (M, not M^+M) around FP32 BiCGStab: each inner solve j
    of N_j counted iterations adds 2 N_j - 1 applications, the outer loop one
    per outer iteration plus the initial and final ones, so
    mv_apps = 2 + 2 sum N_j and site_flops = 189 + 96 n + 960 sum N_j - 888 n."""
    f, fi = fields[q.F64], fields[q.F32]
    g = f["g"]
    solver = q.InvRichardsonMultiPrec(f["clover"], q.InvBiCGStab(fi["clover"], 500), 1e-3, 50)
    st = solver(g.spinor(), rhs(g, 350), 1e-10)
    assert st.status == q.OK and st.n_iters >= 2
    assert (st.mv_apps - 2) % 2 == 0
    sum_n = (st.mv_apps - 2) // 2
    n = st.n_iters
    assert st.site_flops == 189 + 96 * n + 960 * sum_n - 888 * n


def test_batch_cg_counters(fields):
    """The batched CG reports per system the counters of a single CG solve."""
    f = fields[q.F64]
    g = f["g"]
    bt = q.Batch([f["clover"], f["clover"]])
    bs = [rhs(g, 360)[0], rhs(g, 361)[0]]
    sts = bt.cg([g.spinor(), g.spinor()], bs, 1e-9)
    for st in sts:
        assert st.status == q.OK
        assert (st.mv_apps, st.site_flops) == cg_counts(st.n_iters, 1)
