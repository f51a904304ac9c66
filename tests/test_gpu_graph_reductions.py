"""Reductions into device memory (qphix_*_dev): the same values as the host
forms (the same kernels, bit for bit), and capturable in a HIP graph: a graph
of two reductions is captured once on a caller's stream and replayed after
eager updates of its input, every replay checked against the eager host-form
result. The HIP runtime is driven through ctypes (the library's own)."""
import ctypes as C

import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L = [8, 8, 8, 8]


def hip():
    h = C.CDLL("libamdhip64.so")
    vp, pp = C.c_void_p, C.POINTER(C.c_void_p)
    h.hipMalloc.argtypes = [pp, C.c_size_t]
    h.hipFree.argtypes = [vp]
    h.hipMemcpy.argtypes = [vp, vp, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [pp]
    h.hipStreamDestroy.argtypes = [vp]
    h.hipStreamSynchronize.argtypes = [vp]
    h.hipStreamBeginCapture.argtypes = [vp, C.c_int]
    h.hipStreamEndCapture.argtypes = [vp, pp]
    h.hipGraphInstantiate.argtypes = [pp, vp, vp, C.c_char_p, C.c_size_t]
    h.hipGraphLaunch.argtypes = [vp, vp]
    h.hipGraphExecDestroy.argtypes = [vp]
    h.hipGraphDestroy.argtypes = [vp]
    return h


def ok(rc):
    assert rc == 0, rc


class DevBuf:
    def __init__(self, h, n):
        self.h, self.n = h, n
        self.p = C.c_void_p()
        ok(h.hipMalloc(C.byref(self.p), 8 * n))

    def read(self):
        out = np.zeros(self.n)
        ok(self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, 8 * self.n, 2))  # device to host
        return out

    def at(self, k):
        return self.p.value + 8 * k

    def free(self):
        self.h.hipFree(self.p)


@pytest.mark.parametrize("prec,c12", [(q.F64, False), (q.F32, True)], ids=["f64", "f32"])
def test_dev_forms_match_host_forms(prec, c12):
    h = hip()
    g = q.Geometry(L, prec, c12)
    x = g.spinor().upload(oracle.gen_spinor(L, 1, 0))
    y = g.spinor().upload(oracle.gen_spinor(L, 2, 0))
    d = DevBuf(h, 8)
    try:
        q.reduce_dev("norm2_spinor", g, d.at(0), x)
        q.reduce_dev("inner_product", g, d.at(1), x, y)
        q.reduce_dev("inner_product_norm", g, d.at(3), x, y)
        q.sync()
        v = d.read()
        assert v[0] == q.norm2(g, x)
        ip = q.inner_product(g, x, y)
        assert (v[1], v[2]) == (ip.real, ip.imag)
        o = (C.c_double * 3)()
        q.call("qphix_inner_product_norm", g.h, x.p, y.p, o)
        assert list(v[3:6]) == list(o)
        # the updating forms: identical inputs, identical outputs and norms
        y1, y2 = g.spinor(), g.spinor()
        q.copy(g, y1, y)
        q.copy(g, y2, y)
        n1 = C.c_double()
        q.call("qphix_axpy_norm2", g.h, C.c_double(0.75), x.p, y1.p, C.byref(n1))
        q.reduce_dev("axpy_norm2", g, d.at(6), 0.75, x, y2)
        q.sync()
        assert d.read()[6] == n1.value
        assert q.xmy_norm2(g, g.spinor(), y1, y2) == 0.0
        r1, r2 = g.spinor(), g.spinor()
        n1 = q.xmy_norm2(g, r1, x, y)
        q.reduce_dev("xmy_norm2", g, d.at(7), r2, x, y)
        q.sync()
        assert d.read()[7] == n1
        assert q.xmy_norm2(g, g.spinor(), r1, r2) == 0.0
    finally:
        d.free()


def test_graph_capture_and_replay():
    h = hip()
    g = q.Geometry(L, q.F64)
    x = g.spinor().upload(oracle.gen_spinor(L, 3, 0))
    y = g.spinor().upload(oracle.gen_spinor(L, 4, 0))
    d = DevBuf(h, 4)
    stream, graph, exe = C.c_void_p(), C.c_void_p(), C.c_void_p()
    ok(h.hipStreamCreate(C.byref(stream)))
    try:
        q.set_stream(stream)
        q.norm2(g, x)  # the library is initialised before the capture
        ok(h.hipStreamBeginCapture(stream, 1))  # thread-local capture mode
        q.reduce_dev("norm2_spinor", g, d.at(0), x)
        q.reduce_dev("inner_product_norm", g, d.at(1), x, y)
        ok(h.hipStreamEndCapture(stream, C.byref(graph)))
        ok(h.hipGraphInstantiate(C.byref(exe), graph, None, None, 0))
        for rep in range(4):
            q.call("qphix_ax", g.h, C.c_double(1.5), x.p)  # eager work between replays
            ok(h.hipGraphLaunch(exe, stream))
            ok(h.hipStreamSynchronize(stream))
            v = d.read()
            o = (C.c_double * 3)()
            q.call("qphix_inner_product_norm", g.h, x.p, y.p, o)
            assert v[0] == q.norm2(g, x), rep
            assert list(v[1:4]) == list(o), rep
    finally:
        q.set_stream(None)
        if exe.value:
            h.hipGraphExecDestroy(exe)
        if graph.value:
            h.hipGraphDestroy(graph)
        h.hipStreamDestroy(stream)
        d.free()
