"""Multishift CG with its FP64 scalars and zeta / beta recurrences in device
memory (mode 1, eager launches; mode 2, each iteration replayed from a HIP
graph) against the loop with host scalars (mode 0, the reference's control
flow, minvcg.h:279-384): solutions, iteration counts, final residuals and
counters bit-identical, for one and two flavours, FP64 and FP32, converging
and stopped at MaxIters; the solutions checked against the oracle's shifted
systems (M^+M + sigma) x = b."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L8 = [8, 8, 8, 8]
VH = 2048
MASS, MU, EPS = 0.1, 0.07, 0.03
SHIFTS = [1e-4, 1e-3, 1e-2, 0.1, 1.0]


@pytest.fixture(scope="module")
def fields():
    u = oracle.gen_gauge(L8, 6, 0.4)
    clov = oracle.clover_build(L8, u, 1.0, MASS)
    clov = [clov[:VH * 72], clov[VH * 72:]]
    inv = [oracle.clover_invert(c) for c in clov]
    hp = oracle.hop_params(-1.0, 1.0, 1.0)
    orc = {"clover": oracle.Operator(oracle.CLOVER, L8, u, hp, clov=clov, invclov=inv),
           "ndtmwilson": oracle.Operator(oracle.NDTMWILSON, L8, u, hp, mass=MASS, mu=MU, eps=EPS)}
    out = {"orc": orc}
    for prec, c12 in ((q.F64, False), (q.F32, True)):
        g = q.Geometry(L8, prec, c12)
        ud = q.pack_gauge(g, u)
        cl = [g.clover().upload(c) for c in clov]
        iv = [g.clover().upload(c) for c in inv]
        out[prec] = dict(
            g=g, keep=(ud, cl, iv),
            clover=q.Operator(g, q.OP_CLOVER, ud, clov=cl, invclov=iv, t_boundary=-1.0),
            ndtmwilson=q.Operator(g, q.OP_NDTMWILSON, ud, mass=MASS, twisted_mass=MU, epsilon=EPS,
                                  t_boundary=-1.0))
    return out


def run(M, g, nf, bs, mode, max_iters, tol):
    xs = [[g.spinor() for _ in range(nf)] for _ in SHIFTS]
    s = q.MInvCG(M, max_iters, 8, mode=mode)
    status = q.OK
    try:
        nit, fin, mv = s(xs, bs, SHIFTS, [tol] * len(SHIFTS))
    except q.QphixError as e:
        status, nit, fin, mv = e.code, None, None, None
    xh = [np.concatenate([x.download(1) for x in xl]) for xl in xs]
    return status, nit, fin, mv, (s.site_flops if status == q.OK else None), xh


@pytest.mark.parametrize("kind", ["clover", "ndtmwilson"])
@pytest.mark.parametrize("prec", [q.F64, q.F32], ids=["f64", "f32"])
def test_minvcg_device_scalars_bit_identical(fields, kind, prec):
    d = fields[prec]
    g, M = d["g"], d[kind]
    nf = 2 if kind == "ndtmwilson" else 1
    bh = [oracle.gen_spinor(L8, 150 + f, 1) for f in range(nf)]
    bs = [g.spinor().upload(h, 1) for h in bh]
    tol = 1e-10 if prec == q.F64 else 1e-5
    res = [run(M, g, nf, bs, mode, 2000, tol) for mode in (0, 1, 2, 2)]
    ref = res[0]
    assert ref[0] == q.OK and ref[1] > 0
    for r in res[1:]:
        assert r[:5] == ref[:5]
        for a, b in zip(r[5], ref[5]):
            assert np.array_equal(a, b)
    Mo = fields["orc"][kind]
    B = np.concatenate([g.spinor().upload(h, 1).download(1) for h in bh])
    bound = 2e-10 if prec == q.F64 else 5e-5
    for k, sg in enumerate(SHIFTS):
        x = ref[5][k]
        r = B - Mo(Mo(x, 1, 1), -1, 1) - sg * x
        assert np.linalg.norm(r) / np.linalg.norm(B) < bound, (k, sg)


def test_minvcg_device_scalars_at_max_iters(fields):
    """Stopped at MaxIters: ENOCONV in every mode and the same partial solutions."""
    d = fields[q.F64]
    g, M = d["g"], d["clover"]
    bs = [g.spinor().upload(oracle.gen_spinor(L8, 160, 1), 1)]
    res = [run(M, g, 1, bs, mode, 6, 1e-12) for mode in (0, 1, 2)]
    for r in res:
        assert r[0] == q.ENOCONV
        for a, b in zip(r[5], res[0][5]):
            assert np.array_equal(a, b)
