"""The allocation-free clover build (qphix_clover_build_noalloc) against the
allocating one (qphix_clover_build): every output kind, both checkerboards,
FP64 and FP32 (two-row and full gauge), bit for bit, on two calls in a row of
which the second is on a smaller lattice. The larger lattice has more sites
per parity than the static scratch holds (32768), so it is built in several
chunks, the last one partial."""
import ctypes as C

import numpy as np
import pytest

import qphix_amd as q

pytestmark = pytest.mark.gpu

KINDS = [q.CLOV_A, q.CLOV_AINV, q.CLOV_TMINV, q.CLOV_ND_UP, q.CLOV_ND_DN, q.CLOV_ND_OFF]
FULL = (q.CLOV_TMINV, q.CLOV_ND_UP, q.CLOV_ND_DN)


def _raw(field, nbytes):
    """The field's device bytes, copied to the host with hipMemcpy."""
    hip = C.CDLL("libamdhip64.so")
    buf = np.empty(nbytes, dtype=np.uint8)
    st = hip.hipMemcpy(buf.ctypes.data_as(C.c_void_p), field.p, C.c_size_t(nbytes), 2)  # device to host
    assert st == 0, st
    return buf


def _build(fn, g, u, what, cb, out):
    mu, eps = 0.013, 0.004
    q.call(fn, g.h, u[0].p, u[1].p, C.c_double(1.0), C.c_double(0.9), C.c_double(4.03), cb, what,
           C.c_double(mu), C.c_double(eps), out.p)


@pytest.mark.parametrize("prec,c12", [(q.F64, False), (q.F32, True), (q.F32, False)])
def test_noalloc_clover_build_matches_allocating_build(prec, c12):
    # 16^3x20: 40960 sites per parity (one full and one partial scratch
    # chunk), then 6^3x8: 864 sites (one partial chunk)
    for L in ([16, 16, 16, 20], [6, 6, 6, 8]):
        g = q.Geometry(L, prec, c12)
        u = q.gen_gauge(g, 11 + L[0], 0.4)
        q.sync()
        for what in KINDS:
            full = what in FULL
            nbytes = g.bytes[3] if full else g.bytes[2]
            for cb in (0, 1):
                ref = g.full_clover() if full else g.clover()
                got = g.full_clover() if full else g.clover()
                _build("qphix_clover_build", g, u, what, cb, ref)
                _build("qphix_clover_build_noalloc", g, u, what, cb, got)
                q.sync()
                a, b = _raw(ref, nbytes), _raw(got, nbytes)
                assert np.array_equal(a, b), (L, prec, c12, what, cb)
                assert np.any(a != 0)


def test_noalloc_clover_build_rejects_bad_arguments():
    g = q.Geometry([4, 4, 4, 4], q.F64, False)
    u = q.gen_gauge(g, 3, 0.4)
    out = g.clover()
    lib = q.lib()
    args = (C.c_double(1.0), C.c_double(1.0), C.c_double(4.1))
    assert lib.qphix_clover_build_noalloc(g.h, u[0].p, u[1].p, *args, 2, q.CLOV_A, C.c_double(0),
                                          C.c_double(0), out.p) == q.EDOMAIN
    assert lib.qphix_clover_build_noalloc(g.h, u[0].p, u[1].p, *args, 0, 9, C.c_double(0),
                                          C.c_double(0), out.p) == q.EDOMAIN
    assert lib.qphix_clover_build_noalloc(g.h, u[0].p, u[1].p, *args, 0, q.CLOV_A, C.c_double(0),
                                          C.c_double(0), None) == q.EINVAL
