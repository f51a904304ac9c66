"""CPU-side checks of the C-ABI boundary: the library loads, exports every
entry point include/qphix_amd.h declares, and the host-side validation paths
reject bad input with the documented status codes without touching a GPU."""
import ctypes as C

import pytest

import qphix_amd as q


def test_library_exports_every_declared_symbol():
    lib = q.lib()
    names = q.exported_symbols()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_geometry_validation_without_gpu():
    lib = q.lib()
    h = C.c_void_p()
    # odd extent -> EDOMAIN (geometry.h:151-154 throws std::domain_error)
    assert lib.qphix_geom_create((C.c_int * 4)(8, 8, 7, 8), q.F64, 0, C.byref(h)) == q.EDOMAIN
    assert b"even" in lib.qphix_last_error()
    # half precision needs two-row compression
    assert lib.qphix_geom_create((C.c_int * 4)(8, 8, 8, 8), q.F16, 0, C.byref(h)) == q.EDOMAIN
    assert lib.qphix_geom_create((C.c_int * 4)(8, 8, 8, 8), q.BF16, 0, C.byref(h)) == q.EDOMAIN
    assert lib.qphix_geom_create((C.c_int * 4)(8, 8, 8, 8), 7, 0, C.byref(h)) == q.EDOMAIN
    assert lib.qphix_geom_create(None, q.F64, 0, C.byref(h)) == q.EINVAL
    # parities beyond the kernels' 32-bit site index
    assert lib.qphix_geom_create((C.c_int * 4)(256, 256, 256, 256), q.F32, 1, C.byref(h)) == q.EDOMAIN
    assert b"32-bit" in lib.qphix_last_error()


@pytest.mark.parametrize("prec,c12,expect", [
    (q.F64, 0, (192, 1152, 576, 1152)),
    (q.F32, 0, (96, 576, 288, 576)),
    (q.F32, 1, (96, 384, 288, 576)),
    (q.F16, 1, (52, 196, 148, 292)),
    (q.BF16, 1, (52, 196, 148, 292)),
])
def test_field_sizes(prec, c12, expect):
    # bytes per site of spinor / gauge (8 links) / clover / full clover, on a
    # parity of 8192 sites (two whole 4096-site layout blocks)
    g = q.Geometry([16, 16, 16, 4], prec, c12)
    per_site = [b / 8192 for b in g.bytes]
    assert per_site == list(expect)
    # a parity is padded to whole layout blocks: 8^4 (2048 sites) takes one
    g = q.Geometry([8, 8, 8, 8], prec, c12)
    assert [b / 4096 for b in g.bytes] == list(expect)


def test_strerror():
    lib = q.lib()
    assert lib.qphix_strerror(q.ENOCONV) == b"solver did not converge"


def test_dropin_program_links():
    """The drop-in C++ test program links against the library with g++ alone."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "cpp", "dropin_test")
    if not os.path.exists(exe):
        pytest.skip("drop-in test program not built")
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libqphix_amd.so" in out and "not found" not in out.split("libqphix_amd.so")[1].split("\n")[0]


def test_tune_rhs_validation():
    """qphix_tune_rhs rejects register tiles other than 1, 2, 4, 8 and launch
    sizes that are not multiples of the tile (host-side checks, no GPU)."""
    lib = q.lib()
    assert lib.qphix_tune_rhs(3, 6) == q.EDOMAIN
    assert lib.qphix_tune_rhs(4, 6) == q.EDOMAIN
    assert lib.qphix_tune_rhs(8, 40) == q.EDOMAIN
    assert lib.qphix_tune_rhs(2, 1) == q.EDOMAIN
    assert lib.qphix_tune_rhs(0, 0) == q.OK
    assert lib.qphix_tune_config(2, 100, 0, 1) == q.EDOMAIN  # block not a multiple of 64


def test_stream_calibrate_validation():
    """Only the built traffic shapes are accepted (host-side check, no GPU)."""
    lib = q.lib()
    g = q.Geometry([8, 8, 8, 8], q.F32, 1)
    d = C.c_double()
    assert lib.qphix_stream_calibrate(g.h, 47, 6, 5, C.byref(d)) == q.EDOMAIN
    assert lib.qphix_stream_calibrate(g.h, 48, 6, 0, C.byref(d)) == q.EDOMAIN
    assert lib.qphix_stream_calibrate(g.h, 48, 6, 5, None) == q.EINVAL
