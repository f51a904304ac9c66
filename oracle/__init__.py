"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, always to produce reference
results, never as the product path.  Arrays
are float64 numpy arrays in the canonical layout documented in
``oracle/qphix_oracle.h``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

D = C.POINTER(C.c_double)
I4 = C.c_int * 4


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _setup(_lib)
    return _lib


class HopParams(C.Structure):
    _fields_ = [("t_boundary", C.c_double), ("aniso_s", C.c_double),
                ("aniso_t", C.c_double), ("use_tbc", C.c_int * 4),
                ("tbc", (C.c_double * 2) * 4)]


class OrcOp(C.Structure):
    _fields_ = [("kind", C.c_int), ("L", C.c_int * 4), ("hop", HopParams),
                ("u", D), ("mass", C.c_double), ("mu", C.c_double),
                ("eps", C.c_double), ("rho", C.c_double),
                ("clov", D * 2), ("invclov", D * 2), ("fcl", (D * 2) * 2),
                ("ndfcl", (D * 2) * 2), ("ndcoff", D * 2)]


class SolveStats(C.Structure):
    _fields_ = [("n_iters", C.c_int), ("rsd_sq_final", C.c_double),
                ("mv_apps", C.c_ulong), ("status", C.c_int)]


WILSON, CLOVER, TMWILSON, TMCLOVER, NDTMWILSON, NDTMCLOVER = range(6)


def _setup(l):
    l.orc_vh.restype = C.c_long
    l.orc_site_index.restype = C.c_long
    l.orc_uniform.restype = C.c_double
    l.orc_uniform.argtypes = [C.c_uint64, C.c_uint64]
    l.orc_gauss_pair.argtypes = [C.c_uint64, C.c_uint64, D]
    l.orc_gen_gauge.argtypes = [I4, C.c_uint64, C.c_double, D]
    l.orc_gen_spinor.argtypes = [I4, C.c_uint64, C.c_int, D]
    l.orc_su3_exp_i.argtypes = [D, C.c_double, D]
    l.orc_reunit.argtypes = [D]
    l.orc_hop.argtypes = [I4, D, D, D, C.c_int, C.c_int, C.POINTER(HopParams)]
    l.orc_achimbdpsi.argtypes = [I4, D, D, D, D, C.c_double, C.c_double,
                                 C.c_int, C.c_int, C.POINTER(HopParams)]
    l.orc_gamma.argtypes = [C.c_int, D]
    l.orc_field_strength.argtypes = [I4, D, D]
    l.orc_clover_build.argtypes = [I4, D, C.c_double, C.c_double, C.c_double, D]
    l.orc_clover_unpack_block.argtypes = [D, C.c_int, D]
    l.orc_clover_pack_block.argtypes = [D, C.c_int, D]
    l.orc_cmat_inverse.argtypes = [C.c_int, D, D]
    l.orc_clover_invert.argtypes = [C.c_long, D, D]
    l.orc_full_clover_tm_inverse.argtypes = [C.c_long, D, C.c_double, D]
    l.orc_ndclover_inverse.argtypes = [C.c_long, D, C.c_double, C.c_double, D, D, D]
    l.orc_clover_apply.argtypes = [C.c_long, D, D, D]
    l.orc_full_clover_apply.argtypes = [C.c_long, D, D, D]
    l.orc_twisted_apply.argtypes = [C.c_long, C.c_double, C.c_double, D, D]
    l.orc_op_apply.argtypes = [C.POINTER(OrcOp), D, D, C.c_int, C.c_int]
    l.orc_norm2.restype = C.c_double
    l.orc_norm2.argtypes = [C.c_long, D]
    l.orc_cdot.argtypes = [C.c_long, D, D, D]
    l.orc_cg.argtypes = [C.POINTER(OrcOp), D, D, C.c_double, C.c_int, C.c_int,
                         C.c_int, C.POINTER(SolveStats)]
    l.orc_bicgstab.argtypes = l.orc_cg.argtypes
    l.orc_minvcg.argtypes = [C.POINTER(OrcOp), D, D, C.c_int, D, D, C.c_int,
                             C.c_int, C.c_int, C.POINTER(C.c_int), D]


def ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(D)


def vh(L):
    return L[0] * L[1] * L[2] * L[3] // 2


def hop_params(t_boundary=1.0, aniso_s=1.0, aniso_t=1.0, tbc=None):
    p = HopParams()
    p.t_boundary, p.aniso_s, p.aniso_t = t_boundary, aniso_s, aniso_t
    if tbc is not None:
        for d in range(4):
            if tbc[d] is not None:
                p.use_tbc[d] = 1
                p.tbc[d][0], p.tbc[d][1] = tbc[d].real, tbc[d].imag
    return p


def gen_gauge(L, seed, eps=0.4):
    u = np.empty(2 * vh(L) * 4 * 18)
    lib().orc_gen_gauge(I4(*L), seed, eps, ptr(u))
    return u


def gen_spinor(L, seed, cb):
    p = np.empty(vh(L) * 24)
    lib().orc_gen_spinor(I4(*L), seed, cb, ptr(p))
    return p


def hop(L, u, psi, isign, cb, p=None):
    out = np.empty(vh(L) * 24)
    lib().orc_hop(I4(*L), ptr(u), ptr(psi), ptr(out), isign, cb,
                  C.byref(p if p is not None else hop_params()))
    return out


def hop_masked(L, u, psi, isign, cb, mask, p=None):
    """The hop terms selected by mask (bit 2mu backward, 2mu+1 forward)."""
    out = np.empty(vh(L) * 24)
    lib().orc_hop_masked(I4(*L), ptr(u), ptr(psi), ptr(out), isign, cb,
                         C.byref(p if p is not None else hop_params()), C.c_uint(mask))
    return out


def dslash_dir(L, u, psi, cb, dir, p=None):
    """Dslash::dslashDir: one hop term, isign = +1 (dslash_body.h:1074-1137)."""
    out = np.empty(vh(L) * 24)
    lib().orc_dslash_dir(I4(*L), ptr(u), ptr(psi), ptr(out), cb, dir,
                         C.byref(p if p is not None else hop_params()))
    return out


def achimbdpsi(L, u, psi, chi, alpha, beta, isign, cb, p=None):
    out = np.empty(vh(L) * 24)
    lib().orc_achimbdpsi(I4(*L), ptr(u), ptr(psi), ptr(chi), ptr(out), alpha,
                         beta, isign, cb,
                         C.byref(p if p is not None else hop_params()))
    return out


def clover_build(L, u, csw, mass, csw_t=None, diag_mass=None):
    """A = (4+mass) + clover term, both parities, cb-major (72 per site).
    csw weighs the planes without t and, unless csw_t is given, those with t;
    diag_mass replaces 4+mass when given (orc_clover_build)."""
    clov = np.empty(2 * vh(L) * 72)
    ct = csw if csw_t is None else csw_t
    dm = 4.0 + mass if diag_mass is None else diag_mass
    lib().orc_clover_build(I4(*L), ptr(u), csw, ct, dm, ptr(clov))
    return clov


def clover_invert(clov):
    n = clov.size // 72
    inv = np.empty_like(clov)
    lib().orc_clover_invert(n, ptr(clov), ptr(inv))
    return inv


def full_clover_tm_inverse(clov, mu):
    n = clov.size // 72
    out = np.empty(n * 144)
    lib().orc_full_clover_tm_inverse(n, ptr(clov), mu, ptr(out))
    return out


def ndclover_inverse(clov, mu, eps):
    n = clov.size // 72
    fu, fd, co = np.empty(n * 144), np.empty(n * 144), np.empty(n * 72)
    lib().orc_ndclover_inverse(n, ptr(clov), mu, eps, ptr(fu), ptr(fd), ptr(co))
    return fu, fd, co


def clover_apply(clov, x):
    out = np.empty_like(x)
    lib().orc_clover_apply(x.size // 24, ptr(clov), ptr(x), ptr(out))
    return out


def full_clover_apply(fcl, x):
    out = np.empty_like(x)
    lib().orc_full_clover_apply(x.size // 24, ptr(fcl), ptr(x), ptr(out))
    return out


def clover_dense(site72, b):
    m = np.empty(72)
    lib().orc_clover_unpack_block(ptr(np.ascontiguousarray(site72)), b, ptr(m))
    return m.view(np.complex128).reshape(6, 6)


class Operator:
    """Holds numpy fields alive for an orc_op."""

    def __init__(self, kind, L, u, hop=None, mass=0.0, mu=0.0, eps=0.0, rho=0.0,
                 clov=None, invclov=None, fcl=None, ndfcl=None, ndcoff=None):
        self.keep = [u]
        op = OrcOp()
        op.kind = kind
        op.L[:] = list(L)
        op.hop = hop if hop is not None else hop_params()
        op.u = ptr(u)
        op.mass, op.mu, op.eps, op.rho = mass, mu, eps, rho

        def setp(arr):
            a = np.ascontiguousarray(arr)
            self.keep.append(a)
            return ptr(a)
        for cb in range(2):
            if clov is not None:
                op.clov[cb] = setp(clov[cb])
            if invclov is not None:
                op.invclov[cb] = setp(invclov[cb])
            if fcl is not None:
                op.fcl[cb][0] = setp(fcl[cb][0])
                op.fcl[cb][1] = setp(fcl[cb][1])
            if ndfcl is not None:
                op.ndfcl[cb][0] = setp(ndfcl[cb][0])
                op.ndfcl[cb][1] = setp(ndfcl[cb][1])
            if ndcoff is not None:
                op.ndcoff[cb] = setp(ndcoff[cb])
        self.op = op
        self.L = L
        self.nflav = 2 if kind in (NDTMWILSON, NDTMCLOVER) else 1

    def __call__(self, x, isign, target_cb=1):
        res = np.empty_like(x)
        lib().orc_op_apply(C.byref(self.op), ptr(res), ptr(np.ascontiguousarray(x)),
                           isign, target_cb)
        return res

    def cg(self, rhs, rsd, max_iters, isign=1, cb=1, x0=None):
        x = np.zeros_like(rhs) if x0 is None else x0.copy()
        st = SolveStats()
        lib().orc_cg(C.byref(self.op), ptr(x), ptr(rhs), rsd, max_iters, isign,
                     cb, C.byref(st))
        return x, st

    def bicgstab(self, rhs, rsd, max_iters, isign=1, cb=1, x0=None):
        x = np.zeros_like(rhs) if x0 is None else x0.copy()
        st = SolveStats()
        lib().orc_bicgstab(C.byref(self.op), ptr(x), ptr(rhs), rsd, max_iters,
                           isign, cb, C.byref(st))
        return x, st

    def minvcg(self, rhs, shifts, rsd, max_iters, isign=1, cb=1):
        ns = len(shifts)
        x = np.zeros((ns, rhs.size))
        sh = np.ascontiguousarray(shifts, dtype=np.float64)
        rt = np.ascontiguousarray(rsd, dtype=np.float64)
        fin = np.zeros(ns)
        nit = C.c_int(0)
        lib().orc_minvcg(C.byref(self.op), ptr(x), ptr(rhs), ns, ptr(sh), ptr(rt),
                         max_iters, isign, cb, C.byref(nit), ptr(fin))
        return x, nit.value, fin
