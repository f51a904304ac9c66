"""Python binding of the gfx950 QPhiX hot path (libqphix_amd.so, C ABI in
include/qphix_amd.h).

It mirrors the reference's C++ objects (Geometry, Dslash / ClovDslash /
TMDslash / TMClovDslash, EvenOdd*Operator, InvCG / InvBiCGStab / MInvCG /
InvRichardsonMultiPrec) closely enough that the tests read like the
reference's own. Device fields are opaque pointers owned by `Field`. There is
no CPU fallback: loading fails loudly when the HIP library is missing.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqphix_amd.so")
INCLUDE = os.path.join(os.path.dirname(_HERE), "include", "qphix_amd.h")

F64, F32, F16, BF16 = 0, 1, 2, 3
OP_WILSON, OP_CLOVER, OP_TMWILSON, OP_TMCLOVER, OP_NDTMWILSON, OP_NDTMCLOVER = range(6)
CLOV_A, CLOV_AINV, CLOV_TMINV, CLOV_ND_UP, CLOV_ND_DN, CLOV_ND_OFF = range(6)
RESID_ABSOLUTE, RESID_RELATIVE = 0, 1
OK, EDOMAIN, ENOMEM, EHIP, EBREAKDOWN, ENOCONV, EINVAL, ENODEV = range(8)

_lib = None
VP = C.c_void_p
D = C.POINTER(C.c_double)


class QphixError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"qphix status {code}: {msg}")
        self.code = code


def build(jobs=8):
    """Compile the HIP library in-tree (hipcc, gfx950)."""
    subprocess.run(["make", "-s", "-C", os.path.join(_HERE, "csrc"), f"-j{jobs}"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run qphix_amd.build() (hipcc)")
        _lib = C.CDLL(LIB_PATH)
        _lib.qphix_last_error.restype = C.c_char_p
        _lib.qphix_strerror.restype = C.c_char_p
        _lib.qphix_gen_gauge.argtypes = [VP, VP, VP, C.c_uint64, C.c_double]
        _lib.qphix_gen_spinor.argtypes = [VP, VP, C.c_uint64, C.c_int]
        _lib.qphix_op_nflav.restype = C.c_int
        _lib.qphix_tune_summary.restype = C.c_char_p
    return _lib


def check(status):
    if status != 0:
        raise QphixError(status, lib().qphix_last_error().decode())
    return status


def call(name, *args):
    return check(getattr(lib(), name)(*args))


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(D)


def darr(vals):
    a = (C.c_double * len(vals))(*vals)
    return a


def exported_symbols():
    """Names of every entry point declared in include/qphix_amd.h."""
    import re
    txt = open(INCLUDE).read()
    return sorted(set(re.findall(r"\b(qphix_[a-z0-9_]+)\s*\(", txt)))


# ---------------------------------------------------------------------------
class HopParams(C.Structure):
    _fields_ = [("t_boundary", C.c_double), ("aniso_s", C.c_double), ("aniso_t", C.c_double),
                ("use_tbc", C.c_int * 4), ("tbc_phases", (C.c_double * 2) * 4),
                ("mass", C.c_double), ("twisted_mass", C.c_double),
                ("prec_mass_rho", C.c_double)]


class OpParams(C.Structure):
    _fields_ = [("kind", C.c_int), ("mass", C.c_double), ("twisted_mass", C.c_double),
                ("epsilon", C.c_double), ("prec_mass_rho", C.c_double), ("hop", HopParams),
                ("u", VP * 2), ("clov", VP * 2), ("invclov", VP * 2),
                ("fullclov", (VP * 2) * 2), ("clov_odiag", VP * 2)]


class SolveStats(C.Structure):
    _fields_ = [("n_iters", C.c_int), ("rsd_sq_final", C.c_double),
                ("site_flops", C.c_ulong), ("mv_apps", C.c_ulong), ("status", C.c_int)]


def hop_params(t_boundary=1.0, aniso_s=1.0, aniso_t=1.0, tbc=None, mass=0.0,
               twisted_mass=0.0, prec_mass_rho=0.0):
    p = HopParams()
    p.t_boundary, p.aniso_s, p.aniso_t = t_boundary, aniso_s, aniso_t
    if tbc is not None:
        for d in range(4):
            if tbc[d] is not None:
                p.use_tbc[d] = 1
                p.tbc_phases[d][0], p.tbc_phases[d][1] = tbc[d].real, tbc[d].imag
    p.mass, p.twisted_mass, p.prec_mass_rho = mass, twisted_mass, prec_mass_rho
    return p


# ---------------------------------------------------------------------------
class Geometry:
    """Geometry<FT, V, S, compress12> (include/qphix/geometry.h:115-441)."""

    def __init__(self, latt, prec=F64, compress12=False):
        self.latt = list(latt)
        self.prec = prec
        self.compress12 = bool(compress12)
        h = VP()
        call("qphix_geom_create", (C.c_int * 4)(*latt), prec, int(compress12), C.byref(h))
        self.h = h
        vh = C.c_long()
        b = (C.c_size_t * 4)()
        call("qphix_geom_info", h, None, None, None, C.byref(vh), b)
        self.vh = vh.value
        self.bytes = list(b)

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().qphix_geom_destroy(self.h)
        except Exception:
            pass

    def spinor(self):
        return Field(self, "spinor")

    def gauge(self):
        return Field(self, "gauge")

    def clover(self):
        return Field(self, "clover")

    def full_clover(self):
        return Field(self, "full_clover")


class Field:
    _alloc = {"spinor": "qphix_alloc_spinor", "gauge": "qphix_alloc_gauge",
              "clover": "qphix_alloc_clover", "full_clover": "qphix_alloc_full_clover"}

    def __init__(self, geom, kind):
        self.geom = geom
        self.kind = kind
        p = VP()
        call(self._alloc[kind], geom.h, C.byref(p))
        self.p = p

    @property
    def ptr(self):
        return self.p

    def __del__(self):
        try:
            if getattr(self, "p", None):
                lib().qphix_free(self.p)
        except Exception:
            pass

    # spinor helpers
    def upload(self, arr, cb=0):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if self.kind == "spinor":
            call("qphix_pack_spinor", self.geom.h, self.p, dptr(a), cb)
        elif self.kind == "clover":
            call("qphix_pack_clover", self.geom.h, self.p, dptr(a), cb)
        elif self.kind == "full_clover":
            call("qphix_pack_full_clover", self.geom.h, self.p, dptr(a), cb)
        else:
            raise ValueError("use pack_gauge for gauge fields")
        return self

    def download(self, cb=0):
        if self.kind == "spinor":
            out = np.empty(self.geom.vh * 24)
            call("qphix_unpack_spinor", self.geom.h, dptr(out), self.p, cb)
        elif self.kind == "clover":
            out = np.empty(self.geom.vh * 72)
            call("qphix_unpack_clover", self.geom.h, dptr(out), self.p, cb)
        else:
            raise ValueError("download supports spinor and clover fields")
        return out


def pack_gauge(geom, u_canonical):
    u = [geom.gauge(), geom.gauge()]
    a = np.ascontiguousarray(u_canonical, dtype=np.float64)
    call("qphix_pack_gauge", geom.h, u[0].p, u[1].p, dptr(a))
    return u


def gen_gauge(geom, seed, eps=0.4):
    u = [geom.gauge(), geom.gauge()]
    call("qphix_gen_gauge", geom.h, u[0].p, u[1].p, seed, eps)
    return u


def gen_spinor(geom, seed, cb):
    s = geom.spinor()
    call("qphix_gen_spinor", geom.h, s.p, seed, cb)
    return s


def clover_build(geom, u, csw, mass, cb, what=CLOV_A, mu=0.0, eps=0.0, csw_t=None,
                 diag_mass=None):
    """csw is the spatial-plane coefficient (and the temporal one unless csw_t
    is given); the diagonal is 4 + mass unless diag_mass is given."""
    full = what in (CLOV_TMINV, CLOV_ND_UP, CLOV_ND_DN)
    f = geom.full_clover() if full else geom.clover()
    ct = csw if csw_t is None else csw_t
    dm = 4.0 + mass if diag_mass is None else diag_mass
    call("qphix_clover_build", geom.h, u[0].p, u[1].p, C.c_double(csw), C.c_double(ct),
         C.c_double(dm), cb, what, C.c_double(mu), C.c_double(eps), f.p)
    return f


def convert_spinor(g_out, out, scale, g_in, inp):
    call("qphix_convert_spinor", g_out.h, out.p, C.c_double(scale), g_in.h, inp.p)


def convert_gauge(g_out, g_in, u_in):
    out = [g_out.gauge(), g_out.gauge()]
    for cb in range(2):
        call("qphix_convert_gauge", g_out.h, out[cb].p, g_in.h, u_in[cb].p)
    return out


def convert_clover(g_out, g_in, c_in, full=False):
    out = g_out.full_clover() if full else g_out.clover()
    call("qphix_convert_clover", g_out.h, out.p, g_in.h, c_in.p, int(full))
    return out


def tune_summary():
    return lib().qphix_tune_summary().decode()


def sync():
    call("qphix_sync")


class Timer:
    """HIP events on the library stream."""

    def start(self):
        call("qphix_timer_start")

    def stop(self):
        ms = C.c_double()
        call("qphix_timer_stop", C.byref(ms))
        return ms.value


# ---------------------------------------------------------------------------
def _p(x):
    return None if x is None else x.p


def _pa(xs):
    return (VP * len(xs))(*[_p(x) for x in xs])


class Dslash:
    """Dslash / ClovDslash / TMDslash / TMClovDslash (one object type here:
    the methods select the kernel family)."""

    def __init__(self, geom, t_boundary=1.0, aniso_s=1.0, aniso_t=1.0, tbc=None,
                 mass=0.0, twisted_mass=0.0, prec_mass_rho=0.0):
        self.geom = geom
        self.params = hop_params(t_boundary, aniso_s, aniso_t, tbc, mass, twisted_mass,
                                 prec_mass_rho)
        h = VP()
        call("qphix_dslash_create", geom.h, C.byref(self.params), C.byref(h))
        self.h = h

    def __del__(self):
        try:
            lib().qphix_dslash_destroy(self.h)
        except Exception:
            pass

    def dslash(self, res, psi, u, isign, cb):
        call("qphix_wilson_dslash", self.h, res.p, psi.p, u.p, isign, cb)

    def achimbdpsi(self, res, psi, chi, u, alpha, beta, isign, cb):
        call("qphix_wilson_achimbdpsi", self.h, res.p, psi.p, chi.p, u.p, C.c_double(alpha),
             C.c_double(beta), isign, cb)

    def clov_dslash(self, res, psi, u, invclov, isign, cb):
        call("qphix_clov_dslash", self.h, res.p, psi.p, u.p, invclov.p, isign, cb)

    def clov_achimbdpsi(self, res, psi, chi, u, clov, beta, isign, cb):
        call("qphix_clov_achimbdpsi", self.h, res.p, psi.p, chi.p, u.p, clov.p,
             C.c_double(beta), isign, cb)

    def clov_mult(self, res, chi, clov):
        call("qphix_clov_mult", self.h, res.p, chi.p, clov.p)

    def tm_dslash(self, res, psi, u, isign, cb):
        call("qphix_tm_dslash", self.h, res.p, psi.p, u.p, isign, cb)

    def tm_achimbdpsi(self, res, psi, chi, u, alpha, beta, isign, cb, musign=1.0):
        call("qphix_tm_achimbdpsi", self.h, res.p, psi.p, chi.p, u.p, C.c_double(alpha),
             C.c_double(beta), isign, cb, C.c_double(musign))

    def tm_two_flav_achimbdpsi(self, res, psi, chi, u, alpha, beta, eps, isign, cb):
        call("qphix_tm_two_flav_achimbdpsi", self.h, _pa(res), _pa(psi), _pa(chi), u.p,
             C.c_double(alpha), C.c_double(beta), C.c_double(eps), isign, cb)

    def tmclov_dslash(self, res, psi, u, invclov, isign, cb, fl=0):
        call("qphix_tmclov_dslash", self.h, res.p, psi.p, u.p, _pa(invclov), isign, cb, fl)

    def tmclov_achimbdpsi(self, res, psi, chi, u, clov, beta, isign, cb, fl=0):
        call("qphix_tmclov_achimbdpsi", self.h, res.p, psi.p, chi.p, u.p, clov.p,
             C.c_double(beta), isign, cb, fl)

    def tmclov_two_flav_achimbdpsi(self, res, psi, chi, u, clov, beta, eps, isign, cb):
        call("qphix_tmclov_two_flav_achimbdpsi", self.h, _pa(res), _pa(psi), _pa(chi), u.p,
             clov.p, C.c_double(beta), C.c_double(eps), isign, cb)

    def dslash_dir(self, res, psi, u, cb, dir):
        call("qphix_wilson_dslash_dir", self.h, res.p, psi.p, u.p, cb, dir)

    def dslash_multi(self, res, psi, u, isign, cb):
        call("qphix_wilson_dslash_multi", self.h, _pa(res), _pa(psi), len(psi), u.p, isign, cb)

    def clov_dslash_multi(self, res, psi, u, invclov, isign, cb):
        call("qphix_clov_dslash_multi", self.h, _pa(res), _pa(psi), len(psi), u.p, invclov.p,
             isign, cb)

    def clov_achimbdpsi_multi(self, res, psi, chi, u, clov, beta, isign, cb):
        call("qphix_clov_achimbdpsi_multi", self.h, _pa(res), _pa(psi), _pa(chi), len(psi), u.p,
             clov.p, C.c_double(beta), isign, cb)

    def two_flav_inverse_clover_term(self, res, psi, fcl, cl_offdiag, isign):
        call("qphix_tmclov_two_flav_inverse_clover_term", self.h, _pa(res), _pa(psi), _pa(fcl),
             cl_offdiag.p, isign)


class Operator:
    """EvenOdd*Operator (linearOp.h:8-147). Fields are per-checkerboard lists."""

    def __init__(self, geom, kind, u, mass=0.0, twisted_mass=0.0, epsilon=0.0,
                 prec_mass_rho=0.0, t_boundary=1.0, aniso_s=1.0, aniso_t=1.0, tbc=None,
                 clov=None, invclov=None, fullclov=None, clov_odiag=None):
        self.geom = geom
        self.keep = [u, clov, invclov, fullclov, clov_odiag]
        p = OpParams()
        p.kind, p.mass, p.twisted_mass, p.epsilon = kind, mass, twisted_mass, epsilon
        p.prec_mass_rho = prec_mass_rho
        p.hop = hop_params(t_boundary, aniso_s, aniso_t, tbc)
        for cb in range(2):
            p.u[cb] = u[cb].p
            if clov is not None and clov[cb] is not None:
                p.clov[cb] = clov[cb].p
            if invclov is not None and invclov[cb] is not None:
                p.invclov[cb] = invclov[cb].p
            if fullclov is not None and fullclov[cb] is not None:
                p.fullclov[cb][0] = fullclov[cb][0].p
                p.fullclov[cb][1] = fullclov[cb][1].p
            if clov_odiag is not None and clov_odiag[cb] is not None:
                p.clov_odiag[cb] = clov_odiag[cb].p
        self.params = p
        h = VP()
        call("qphix_op_create", geom.h, C.byref(p), C.byref(h))
        self.h = h
        self.nflav = lib().qphix_op_nflav(h)

    def __del__(self):
        try:
            lib().qphix_op_destroy(self.h)
        except Exception:
            pass

    def __call__(self, res, inp, isign, target_cb=1):
        res = res if isinstance(res, (list, tuple)) else [res]
        inp = inp if isinstance(inp, (list, tuple)) else [inp]
        call("qphix_op_apply", self.h, _pa(res), _pa(inp), isign, target_cb)

    def apply_multi(self, res, inp, isign, target_cb=1):
        call("qphix_op_apply_multi", self.h, _pa(res), _pa(inp), len(inp), isign, target_cb)

    # multigrid-facing pieces (linearOp.h:26-35, clover.h:174-251); full
    # spinors are [cb0, cb1] pairs
    def m_offdiag(self, res, inp, isign, target_cb):
        call("qphix_op_m_offdiag", self.h, _pa([res]), _pa([inp]), isign, target_cb)

    def m_diag_inv(self, res, inp, isign, cb=0):
        call("qphix_op_m_diag_inv", self.h, _pa([res]), _pa([inp]), isign, cb)

    def m_diag(self, res, inp, isign, cb=-1):
        call("qphix_op_m_diag", self.h, _pa(res), _pa(inp), isign, cb)

    def m_unprec(self, res, inp, isign):
        call("qphix_op_m_unprec", self.h, _pa(res), _pa(inp), isign)

    def dslash_dir(self, res, inp, dir):
        call("qphix_op_dslash_dir", self.h, _pa(res), _pa(inp), dir)


class Solver:
    def __init__(self, h, M, keep=()):
        self.h, self.M, self.keep = h, M, keep

    def __del__(self):
        try:
            lib().qphix_solver_destroy(self.h)
        except Exception:
            pass

    def set_mode(self, mode):
        """0 host scalars, 1 device scalars, 2 device scalars with the
        iteration replayed from a HIP graph (BiCGStab and multishift CG)."""
        call("qphix_solver_set_mode", self.h, mode)
        return self

    def __call__(self, x, rhs, rsd_target, isign=1, target_cb=1, resid=RESID_RELATIVE,
                 verbose=False):
        x = x if isinstance(x, (list, tuple)) else [x]
        rhs = rhs if isinstance(rhs, (list, tuple)) else [rhs]
        st = SolveStats()
        call("qphix_solve", self.h, _pa(x), _pa(rhs), C.c_double(rsd_target), isign,
             int(verbose), target_cb, resid, C.byref(st))
        return st

    def solve_multi(self, x, rhs, rsd_target, isign=1, target_cb=1, resid=RESID_RELATIVE,
                    verbose=False):
        n = len(rhs)
        st = (SolveStats * n)()
        call("qphix_solve_multi", self.h, _pa(x), _pa(rhs), n, C.c_double(rsd_target), isign,
             int(verbose), target_cb, resid, st)
        return list(st)


def InvCG(M, max_iters):
    h = VP()
    call("qphix_cg_create", M.h, max_iters, C.byref(h))
    return Solver(h, M)


def InvMR(M, max_iters, omega=1.0):
    h = VP()
    call("qphix_mr_create", M.h, max_iters, C.c_double(omega), 0, C.byref(h))
    return Solver(h, M)


def InvMRSmoother(M, max_iters, omega=1.0):
    h = VP()
    call("qphix_mr_create", M.h, max_iters, C.c_double(omega), 1, C.byref(h))
    return Solver(h, M)


def unprec_solve(eo_solver, x, rhs, rsd_target, isign=1, solve_cb=1, resid=RESID_RELATIVE,
                 verbose=False):
    """x, rhs: [cb0, cb1] field pairs (UnprecSolverWrapper + AbstractUnprecSolver)."""
    st = SolveStats()
    call("qphix_unprec_solve", eo_solver.h, _pa(x), _pa(rhs), C.c_double(rsd_target), isign,
         int(verbose), solve_cb, resid, C.byref(st))
    return st


def InvBlockCG(M, max_iters):
    """Block CG on M^+M for up to 32 right-hand sides (use solve_multi)."""
    h = VP()
    call("qphix_block_cg_create", M.h, max_iters, C.byref(h))
    return Solver(h, M)


def InvBiCGStab(M, max_iters):
    h = VP()
    call("qphix_bicgstab_create", M.h, max_iters, C.byref(h))
    return Solver(h, M)


def InvRichardsonMultiPrec(M_outer, inner, delta, max_iters, mdagm=False):
    h = VP()
    call("qphix_richardson_create", M_outer.h, inner.h, C.c_double(delta), max_iters,
         int(mdagm), C.byref(h))
    return Solver(h, M_outer, keep=(inner,))


class MInvCG:
    """mode: 0 host scalars, 1 FP64 scalars in device memory (eager
    launches), 2 the same with each iteration replayed from a HIP graph
    (the library's default)."""

    def __init__(self, M, max_iters, max_shifts, mode=None):
        self.M = M
        h = VP()
        call("qphix_minvcg_create", M.h, max_iters, max_shifts, C.byref(h))
        self.h = h
        if mode is not None:
            call("qphix_minvcg_set_mode", h, mode)

    def __del__(self):
        try:
            lib().qphix_solver_destroy(self.h)
        except Exception:
            pass

    def __call__(self, x, rhs, shifts, rsd_target, isign=1, target_cb=1, verbose=False):
        """x: list over shifts of lists over flavours; rhs: list over flavours."""
        ns = len(shifts)
        xs = [f for xl in x for f in (xl if isinstance(xl, (list, tuple)) else [xl])]
        rhs = rhs if isinstance(rhs, (list, tuple)) else [rhs]
        nit = C.c_int()
        fin = (C.c_double * ns)()
        mv = C.c_ulong()
        fl = C.c_ulong()
        call("qphix_minvcg_solve", self.h, _pa(xs), _pa(rhs), ns, darr(shifts), darr(rsd_target),
             isign, int(verbose), target_cb, C.byref(nit), fin, C.byref(fl), C.byref(mv))
        self.site_flops = fl.value
        return nit.value, list(fin), mv.value


class Batch:
    """Several single-flavour systems solved together (batched launches,
    device scalars, optional HIP-graph replay of the CG iteration)."""

    def __init__(self, ops):
        self.ops = list(ops)
        h = VP()
        call("qphix_batch_create", (VP * len(ops))(*[o.h for o in ops]), len(ops), C.byref(h))
        self.h = h

    def __del__(self):
        try:
            lib().qphix_batch_destroy(self.h)
        except Exception:
            pass

    def cg(self, x, rhs, rsd, max_iters=1000, isign=1, target_cb=1, graph_steps=0):
        n = len(self.ops)
        st = (SolveStats * n)()
        tg = rsd if isinstance(rsd, (list, tuple)) else [rsd] * n
        call("qphix_batch_cg", self.h, _pa(x), _pa(rhs), darr(tg), max_iters, isign, target_cb,
             graph_steps, st)
        return list(st)

    def richardson(self, outer_ops, x, rhs, rsd, delta=1e-2, max_outer=50, max_inner=2000,
                   isign=1, target_cb=1, graph_steps=0):
        n = len(self.ops)
        st = (SolveStats * n)()
        call("qphix_batch_richardson", self.h, (VP * n)(*[o.h for o in outer_ops]), _pa(x), _pa(rhs),
             C.c_double(rsd), C.c_double(delta), max_outer, max_inner, isign, target_cb,
             graph_steps, st)
        return list(st)


# ---------------------------------------------------------------------------
# BLAS (blas_new_c.h), single flavour
def norm2(geom, x):
    r = C.c_double()
    call("qphix_norm2_spinor", geom.h, x.p, C.byref(r))
    return r.value


def xmy_norm2(geom, res, x, y):
    """res = x - y; returns |res|^2 (xmyNorm2Spinor)."""
    r = C.c_double()
    call("qphix_xmy_norm2", geom.h, res.p, x.p, y.p, C.byref(r))
    return r.value


def set_stream(stream):
    """Order the library's work on a HIP stream (a hipStream_t as int or
    c_void_p); None returns to the library's own stream."""
    call("qphix_set_stream", VP(stream) if isinstance(stream, int) else stream)


def reduce_dev(name, geom, d_out, *args):
    """The device-memory form of a reduction, qphix_<name>_dev: results land in
    d_out (a device address) without synchronisation."""
    conv = [a.p if isinstance(a, Field) else (C.c_double(a) if isinstance(a, float) else a)
            for a in args]
    call(f"qphix_{name}_dev", geom.h, *conv, VP(d_out) if isinstance(d_out, int) else d_out)


def block_gram(geom, a, b):
    """G[i, j] = <a_i, b_j> (complex numpy matrix)."""
    o = np.zeros(len(a) * len(b) * 2)
    call("qphix_block_gram", geom.h, _pa(a), len(a), _pa(b), len(b), dptr(o))
    return (o[0::2] + 1j * o[1::2]).reshape(len(a), len(b))


def block_combine(geom, a, mat, y, o):
    """o_j = y_j + sum_i a_i mat[i, j] (y None: no addition)."""
    mat = np.asarray(mat, dtype=complex)
    m = np.empty(mat.size * 2)
    m[0::2], m[1::2] = mat.real.ravel(), mat.imag.ravel()
    call("qphix_block_combine", geom.h, _pa(a), len(a), dptr(m), _pa(y) if y is not None else None,
         _pa(o), len(o))


def norm2_multi(geom, xs):
    o = (C.c_double * len(xs))()
    call("qphix_norm2_multi", geom.h, _pa(xs), len(xs), o)
    return list(o)


def inner_product(geom, l, r):
    o = (C.c_double * 2)()
    call("qphix_inner_product", geom.h, l.p, r.p, o)
    return complex(o[0], o[1])


def copy(geom, res, src):
    call("qphix_copy_spinor", geom.h, res.p, src.p)


def zero(geom, res):
    call("qphix_zero_spinor", geom.h, res.p)


# ---------------------------------------------------------------------------
# Chroma LinOpSysSolver surface (README.markdown:163-256)
INV_CLOVER, INV_CLOVER_ITER_REFINE = 0, 1
SOLVER_CG, SOLVER_BICGSTAB = 0, 1


class InvertParam(C.Structure):
    """The <InvertParam> group of the QPhiX clover inverters."""
    _fields_ = [("inv_type", C.c_int), ("solver_type", C.c_int), ("max_iter", C.c_int),
                ("rsd_target", C.c_double), ("delta", C.c_double), ("mass", C.c_double),
                ("clov_coeff_r", C.c_double), ("clov_coeff_t", C.c_double),
                ("aniso_p", C.c_int), ("t_dir", C.c_int), ("xi_0", C.c_double),
                ("nu", C.c_double), ("antiperiodic_t", C.c_int), ("verbose", C.c_int),
                ("tune", C.c_int), ("rsd_tolerance_factor", C.c_double),
                ("ncores", C.c_int), ("threads_per_core", C.c_int), ("by", C.c_int),
                ("bz", C.c_int), ("sy", C.c_int), ("sz", C.c_int), ("pad_xy", C.c_int),
                ("pad_xyz", C.c_int), ("min_ct", C.c_int), ("prec", C.c_int),
                ("compress12", C.c_int), ("inner_prec", C.c_int),
                ("inner_compress12", C.c_int)]

    @classmethod
    def default(cls, **kw):
        p = cls()
        call("qphix_invert_param_default", C.byref(p))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise AttributeError(k)
            setattr(p, k, v)
        return p

    @classmethod
    def from_xml(cls, xml):
        p = cls()
        call("qphix_invert_param_read_xml", xml.encode() if isinstance(xml, str) else xml,
             C.byref(p))
        return p

    def coeffs(self):
        """(aniso_s, aniso_t, t_boundary, csw_r, csw_t, diag_mass)"""
        out = (C.c_double * 6)()
        call("qphix_invert_param_coeffs", C.byref(self), out)
        return tuple(out)


class SysSolverResults(C.Structure):
    _fields_ = [("n_count", C.c_int), ("resid", C.c_double), ("rel_resid", C.c_double),
                ("mv_apps", C.c_ulong), ("seconds", C.c_double), ("status", C.c_int)]


class SysSolver:
    """The odd-checkerboard clover system M psi = chi described by an
    InvertParam, on a canonical host gauge field."""

    def __init__(self, latt, param, u_canonical):
        if isinstance(param, (str, bytes)):
            param = InvertParam.from_xml(param)
        self.param = param
        self.latt = list(latt)
        a = np.ascontiguousarray(u_canonical, dtype=np.float64)
        h = VP()
        call("qphix_syssolver_create", (C.c_int * 4)(*latt), C.byref(param), dptr(a), C.byref(h))
        self.h = h
        g = VP()
        call("qphix_syssolver_geom", h, C.byref(g))
        vh = C.c_long()
        call("qphix_geom_info", g, None, None, None, C.byref(vh), None)
        self.vh = vh.value

    def __del__(self):
        try:
            if getattr(self, "h", None):
                lib().qphix_syssolver_destroy(self.h)
        except Exception:
            pass

    def __call__(self, psi, chi):
        """psi (updated in place), chi: canonical full-lattice spinors."""
        assert psi.size == chi.size == 2 * self.vh * 24
        c = np.ascontiguousarray(chi, dtype=np.float64)
        r = SysSolverResults()
        call("qphix_syssolver_solve", self.h, dptr(psi), dptr(c), C.byref(r))
        return r

    def solve_dev(self, psi, chi):
        r = SysSolverResults()
        call("qphix_syssolver_solve_dev", self.h, psi.p, chi.p, C.byref(r))
        return r
