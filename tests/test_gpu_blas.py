"""Fused BLAS on the GPU against closed forms in numpy on the same data
(tests-gtest/full_blas_test.cc checks the same functions against QDP++).
Reductions accumulate in FP64; FP64 results must agree to 1e-13 relative."""
import ctypes as C

import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

# one 4096-site layout block, and a parity of 8,400 sites over three blocks
# (the last one partial)
LATS = [[4, 4, 8, 8], [10, 12, 14, 10]]
PRECS = [q.F64, q.F32, q.F16, q.BF16]


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


@pytest.fixture(scope="module", params=[(p, L) for L in LATS for p in PRECS],
                ids=[f"{n}-{'x'.join(map(str, L))}" for L in LATS for n in ("f64", "f32", "f16", "bf16")])
def env(request):
    prec, L = request.param
    g = q.Geometry(L, prec, prec in (q.F16, q.BF16))
    tol = {q.F64: 1e-13, q.F32: 3e-7, q.F16: 2e-3, q.BF16: 1.6e-2}[prec]
    return g, tol


def vec(g, seed):
    d = g.spinor().upload(oracle.gen_spinor(g.latt, seed, 0))
    return d, d.download()


def cx(v):
    return v.view(np.complex128)


def test_blas_ops(env):
    g, tol = env
    lib = q.lib()
    x, xh = vec(g, 1)
    y, yh = vec(g, 2)
    z, zh = vec(g, 3)
    w, wh = vec(g, 4)
    out = g.spinor()
    r = C.c_double()
    # norms and inner products
    assert abs(q.norm2(g, x) - xh @ xh) / (xh @ xh) < max(tol, 1e-14)
    ip = q.inner_product(g, x, y)
    ref = np.vdot(cx(xh), cx(yh))
    assert abs(ip - ref) / abs(ref) < max(tol * 10, 1e-12)
    o3 = (C.c_double * 3)()
    q.call("qphix_inner_product_norm", g.h, x.p, y.p, o3)
    assert abs(o3[2] - xh @ xh) / (xh @ xh) < max(tol, 1e-14)
    # xmy_norm2: out = x - y
    q.call("qphix_xmy_norm2", g.h, out.p, x.p, y.p, C.byref(r))
    assert rel(out.download(), xh - yh) < tol
    assert abs(r.value - (xh - yh) @ (xh - yh)) / r.value < max(tol, 1e-13)

    def fresh(h):
        return g.spinor().upload(h)

    cases = [
        ("qphix_axy", (C.c_double(1.7), x.p), lambda: 1.7 * xh),
        ("qphix_aypx", (C.c_double(0.3), x.p), lambda: 0.3 * yh + xh),
        ("qphix_axpy", (C.c_double(-0.4), x.p), lambda: -0.4 * xh + yh),
        ("qphix_axpby", (C.c_double(0.5), x.p, C.c_double(-2.0)), lambda: 0.5 * xh - 2.0 * yh),
        ("qphix_ypeqx", (x.p,), lambda: yh + xh),
        ("qphix_ymeqx", (x.p,), lambda: yh - xh),
        ("qphix_bicgstab_xmy", (x.p,), lambda: xh - yh),
    ]
    for name, args, f in cases:
        t = fresh(yh)
        q.call(name, g.h, *args, t.p)
        assert rel(t.download(), f()) < tol * 4, name
    t = fresh(yh)
    q.call("qphix_ax", g.h, C.c_double(2.5), t.p)
    assert rel(t.download(), 2.5 * yh) < tol
    # complex ops
    a = (C.c_double * 2)(0.3, -0.8)
    ac = complex(0.3, -0.8)
    t = fresh(yh)
    q.call("qphix_caxpy", g.h, a, x.p, t.p)
    assert rel(cx(t.download()), ac * cx(xh) + cx(yh)) < tol * 4
    t = fresh(yh)
    q.call("qphix_bicgstab_s_update", g.h, a, t.p, x.p)
    assert rel(cx(t.download()), cx(yh) - ac * cx(xh)) < tol * 4
    beta = (C.c_double * 2)(0.7, 0.2)
    omega = (C.c_double * 2)(-0.1, 0.4)
    t = fresh(yh)  # p
    q.call("qphix_bicgstab_p_update", g.h, x.p, t.p, z.p, beta, omega)
    ref = cx(xh) + complex(0.7, 0.2) * (cx(yh) - complex(-0.1, 0.4) * cx(zh))
    assert rel(cx(t.download()), ref) < tol * 4
    tx, tr = fresh(wh), fresh(yh)
    q.call("qphix_bicgstab_rxupdate", g.h, tx.p, tr.p, z.p, x.p, omega, a, C.byref(r))
    rr = cx(yh) - complex(-0.1, 0.4) * cx(zh)
    xx = cx(wh) + complex(-0.1, 0.4) * cx(yh) + ac * cx(xh)
    assert rel(cx(tx.download()), xx) < tol * 4
    assert rel(cx(tr.download()), rr) < tol * 4
    assert abs(r.value - np.vdot(rr, rr).real) / r.value < max(tol * 4, 1e-13)
    # CG update: r -= a mmp ; x += a p
    tr, tx = fresh(yh), fresh(wh)
    q.call("qphix_rmammp_norm2_rxpap", g.h, tr.p, C.c_double(0.6), x.p, tx.p, z.p, C.byref(r))
    assert rel(tr.download(), yh - 0.6 * xh) < tol * 4
    assert rel(tx.download(), wh + 0.6 * zh) < tol * 4
    # Richardson update
    tx, tr = fresh(wh), fresh(yh)
    q.call("qphix_richardson_rxupdate_norm_r", g.h, tx.p, tr.p, x.p, z.p, C.byref(r))
    assert rel(tx.download(), wh + xh) < tol * 4
    assert rel(tr.download(), yh - zh) < tol * 4
    # axpy_norm2 / xmy2_norm2
    t = fresh(yh)
    q.call("qphix_axpy_norm2", g.h, C.c_double(0.2), x.p, t.p, C.byref(r))
    assert rel(t.download(), 0.2 * xh + yh) < tol * 4
    t = fresh(yh)
    q.call("qphix_xmy2_norm2", g.h, x.p, t.p, C.byref(r))
    assert rel(t.download(), yh - xh) < tol * 4
    # MR update x += a r ; r -= a Mr
    tx, tr = fresh(wh), fresh(yh)
    q.call("qphix_mr_rxupdate", g.h, tx.p, tr.p, x.p, a)
    assert rel(cx(tx.download()), cx(wh) + ac * cx(yh)) < tol * 4
    assert rel(cx(tr.download()), cx(yh) - ac * cx(xh)) < tol * 4
    # twisted mass: (alpha + i mu g5)
    apimu = (C.c_double * 2)(4.1, 0.05)
    t = g.spinor()
    q.call("qphix_twisted_mass", g.h, apimu, x.p, t.p)
    ref = np.empty_like(xh)
    oracle.lib().orc_twisted_apply(g.vh, 4.1, 0.05, oracle.ptr(xh), oracle.ptr(ref))
    assert rel(t.download(), ref) < tol * 4
    # copy / zero
    q.copy(g, out, x)
    assert rel(out.download(), xh) < 1e-16 + tol * (g.prec in (q.F16, q.BF16))
    q.zero(g, out)
    assert not out.download().any()


def test_two_flav_twisted_mass(env):
    g, tol = env
    x, xh = vec(g, 11)
    y, yh = vec(g, 12)
    ou, od = g.spinor(), g.spinor()
    apimu = (C.c_double * 2)(0.9, -0.2)
    q.call("qphix_two_flav_twisted_mass", g.h, apimu, C.c_double(0.3),
           (C.c_void_p * 2)(x.p, y.p), (C.c_void_p * 2)(ou.p, od.p))
    g5 = np.tile(np.repeat([1.0, 1.0, -1.0, -1.0], 3), g.vh)
    a = 0.9 + 1j * (-0.2) * g5
    ref_u = a * cx(xh) + 0.3 * cx(yh)
    ref_d = np.conj(a) * cx(yh) + 0.3 * cx(xh)
    assert rel(cx(ou.download()), ref_u) < tol * 4
    assert rel(cx(od.download()), ref_d) < tol * 4


def test_reduction_deterministic_and_large():
    """FP64 norm of a 16^3x32 field: bitwise repeatable and equal to a Kahan-free
    numpy sum to 1e-13."""
    L2 = [16, 16, 16, 32]
    g = q.Geometry(L2, q.F64, False)
    x = q.gen_spinor(g, 5, 0)
    vals = [q.norm2(g, x) for _ in range(3)]
    assert vals[0] == vals[1] == vals[2]
    xh = x.download()
    assert abs(vals[0] - np.sum(xh * xh)) / vals[0] < 1e-13


def test_stream_calibrate_runs():
    """The streaming calibration kernel runs for every built traffic shape and
    reports a positive mean launch time."""
    g = q.Geometry([8, 8, 8, 16], q.F32, True)
    for wi, wo in ((48, 6), (60, 6), (120, 12)):
        ms = C.c_double()
        q.call("qphix_stream_calibrate", g.h, wi, wo, 3, C.byref(ms))
        assert 0 < ms.value < 100.0, (wi, wo, ms.value)
