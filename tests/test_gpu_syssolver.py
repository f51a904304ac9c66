"""The Chroma LinOpSysSolver surface (README.markdown:163-256) on the GPU:
<InvertParam> groups of QPHIX_CLOVER_INVERTER and
QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER solve the odd-checkerboard clover
system A_oo - 1/4 D_oe A_ee^-1 D_eo (clover.h:162-172) built from a canonical
host gauge field.

The solutions are checked against the oracle's operator assembled from the
coefficients written out here from tests/clover_term_qdp_w.h:156-180 (not from
the library's own mapping), and the mass / anisotropy conventions are pinned
independently by the free field: on unit links with periodic boundaries a
constant spinor is an eigenvector of the even-odd operator with eigenvalue
m (2 + 6 nu/xi_0 + m) / (1 + 3 nu/xi_0 + m)."""
import numpy as np
import pytest

import oracle
import qphix_amd as q

pytestmark = pytest.mark.gpu

L = [8, 8, 8, 8]
VH = 2048


def xml(inv="QPHIX_CLOVER_INVERTER", solver="BICGSTAB", max_iter=1000, rsd=1e-10,
        clover="<Kappa>0.12</Kappa><clovCoeff>1.0</clovCoeff>", extra=""):
    return f"""<?xml version="1.0"?>
<InvertParam>
  <invType>{inv}</invType>
  <SolverType>{solver}</SolverType>
  <MaxIter>{max_iter}</MaxIter>
  <RsdTarget>{rsd}</RsdTarget>
  <CloverParams>{clover}</CloverParams>
  <NCores>16</NCores><ThreadsPerCore>2</ThreadsPerCore><By>8</By><Bz>8</Bz>
  <Sy>1</Sy><Sz>2</Sz><PadXY>1</PadXY><PadXYZ>0</PadXYZ><MinCt>1</MinCt>
  {extra}
</InvertParam>"""


ANISO = ("<Mass>0.1</Mass><clovCoeffR>2.0171</clovCoeffR><clovCoeffT>0.95</clovCoeffT>"
         "<AnisoParam><anisoP>true</anisoP><t_dir>3</t_dir><xi_0>2.9</xi_0><nu>0.94</nu>"
         "</AnisoParam>")


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def gauge():
    return oracle.gen_gauge(L, 71, 0.4)


@pytest.fixture(scope="module")
def chi():
    return np.concatenate([oracle.gen_spinor(L, 72, 0), oracle.gen_spinor(L, 73, 1)])


def oracle_op(u, aniso_s, t_boundary, csw_r, csw_t, diag_mass):
    clov = oracle.clover_build(L, u, csw_r, 0.0, csw_t=csw_t, diag_mass=diag_mass)
    cl = [clov[:VH * 72], clov[VH * 72:]]
    inv = [oracle.clover_invert(c) for c in cl]
    return oracle.Operator(oracle.CLOVER, L, u, oracle.hop_params(t_boundary, aniso_s, 1.0),
                           clov=cl, invclov=inv)


# isotropic, kappa = 0.12: 4 + m = 1/(2 kappa); anisotropic (xi_0 = 2.9, nu = 0.94):
# spatial hops nu/xi_0, spatial clover planes c_R/xi_0, diagonal 1 + 3 nu/xi_0 + m
ISO = dict(aniso_s=1.0, csw_r=1.0, csw_t=1.0, diag_mass=1 / 0.24)
ANI = dict(aniso_s=0.94 / 2.9, csw_r=2.0171 / 2.9, csw_t=0.95, diag_mass=1 + 3 * 0.94 / 2.9 + 0.1)


@pytest.mark.parametrize("case", ["iso_bicgstab_antiperiodic", "aniso_cg_periodic"])
def test_clover_inverter(gauge, chi, case):
    if case.startswith("iso"):
        S = q.SysSolver(L, xml(extra="<AntiPeriodicT>true</AntiPeriodicT>"), gauge)
        Mo = oracle_op(gauge, t_boundary=-1.0, **ISO)
        tol = 2e-10
    else:
        # CG on the normal equations meets RsdTarget on M^+ r; the true residual
        # of M psi = chi may be larger by up to the condition number
        S = q.SysSolver(L, xml(solver="CG", clover=ANISO,
                               extra="<RsdToleranceFactor>100</RsdToleranceFactor>"), gauge)
        Mo = oracle_op(gauge, t_boundary=1.0, **ANI)
        tol = 1e-8
    psi = np.zeros(2 * VH * 24)
    psi[:VH * 24] = 7.0  # the even half is not touched
    r = S(psi, chi)
    assert r.status == q.OK, (r.rel_resid, r.n_count)
    assert np.all(psi[:VH * 24] == 7.0)
    x = psi[VH * 24:]
    res = rel(Mo(x, 1, 1), chi[VH * 24:])
    assert res < tol
    # the library's FP64 true residual agrees with the oracle's
    assert abs(r.rel_resid - res) < 1e-3 * res
    assert r.n_count > 1 and r.mv_apps >= r.n_count and r.seconds > 0


@pytest.mark.parametrize("inner,solver", [("f", "BICGSTAB"), ("h", "CG")])
def test_iter_refine_inverter(gauge, chi, inner, solver):
    extra = (f"<Delta>0.1</Delta><AntiPeriodicT>true</AntiPeriodicT><InnerPrecision>{inner}"
             "</InnerPrecision><InnerCompress12>true</InnerCompress12>")
    # an inner CG refines the normal equations M^+M psi = M^+ chi (see above)
    tol = 2e-10 if solver == "BICGSTAB" else 1e-8
    if solver == "CG":
        extra += "<RsdToleranceFactor>100</RsdToleranceFactor>"
    S = q.SysSolver(L, xml("QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER", solver, extra=extra), gauge)
    psi = np.zeros(2 * VH * 24)
    r = S(psi, chi)
    assert r.status == q.OK
    Mo = oracle_op(gauge, t_boundary=-1.0, **ISO)
    res = rel(Mo(psi[VH * 24:], 1, 1), chi[VH * 24:])
    assert res < tol
    assert abs(r.rel_resid - res) < 1e-3 * res
    # the same solution as the FP64 BiCGStab solve
    ref = np.zeros(2 * VH * 24)
    q.SysSolver(L, xml(extra="<AntiPeriodicT>true</AntiPeriodicT>"), gauge)(ref, chi)
    assert rel(psi[VH * 24:], ref[VH * 24:]) < 50 * tol


def unit_gauge():
    u = np.zeros((2 * VH, 4, 3, 3, 2))
    for c in range(3):
        u[:, :, c, c, 0] = 1.0
    return u.ravel()


@pytest.mark.parametrize("case", ["iso_kappa", "aniso_mass", "aniso_refine_half"])
def test_free_field_eigenvalue(case):
    """Unit links, periodic: constant chi on the odd sites gives psi = chi / lambda,
    lambda = m (2 + 6 nu/xi_0 + m) / (1 + 3 nu/xi_0 + m)."""
    rng = np.random.default_rng(5)
    s24 = rng.standard_normal(24)
    chi = np.concatenate([np.zeros(VH * 24), np.tile(s24, VH)])
    if case == "iso_kappa":
        m, f = 1 / 0.24 - 4, 1.0
        x = xml()
    else:
        m, f = 0.1, 0.94 / 2.9
        x = xml(clover=ANISO) if case == "aniso_mass" else xml(
            "QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER", clover=ANISO,
            extra="<InnerPrecision>h</InnerPrecision><Compress12>true</Compress12>")
    lam = m * (2 + 6 * f + m) / (1 + 3 * f + m)
    S = q.SysSolver(L, x, unit_gauge())
    psi = np.zeros_like(chi)
    r = S(psi, chi)
    assert r.status == q.OK
    # one BiCGStab step solves an eigenvector exactly in FP64; the refined
    # half-precision solve stops at RsdTarget 1e-10
    tol = 2e-10 if case == "aniso_refine_half" else 1e-12
    assert np.abs(psi[VH * 24:] - np.tile(s24 / lam, VH)).max() < tol * np.abs(s24 / lam).max()


def test_status_and_tolerance_factor(gauge, chi):
    # too few iterations: the call succeeds, the status says the solve failed;
    # BiCGStab reports MaxIter + 1, where its loop counter stops
    # (invbicgstab.h:155-257)
    S = q.SysSolver(L, xml(max_iter=3, rsd=1e-12), gauge)
    psi = np.zeros(2 * VH * 24)
    r = S(psi, chi)
    assert r.status == q.ENOCONV and r.n_count == 4 and r.rel_resid > 1e-6
    # converged to RsdTarget 1e-6 but held to 1e-6 * 1e-6 by the tolerance factor
    S = q.SysSolver(L, xml(rsd=1e-6, extra="<RsdToleranceFactor>1e-6</RsdToleranceFactor>"), gauge)
    r = S(np.zeros(2 * VH * 24), chi)
    assert r.status == q.ENOCONV and r.rel_resid < 1e-5


def test_device_fields_and_tuning(gauge, chi):
    """qphix_syssolver_solve_dev on packed spinors gives the host-array result
    bit for bit, with the kernels autotuned at creation (<Tune>)."""
    S = q.SysSolver(L, xml(extra="<Tune>true</Tune><AntiPeriodicT>true</AntiPeriodicT>"), gauge)
    psi = np.zeros(2 * VH * 24)
    r1 = S(psi, chi)
    g = q.Geometry(L, q.F64)
    c = g.spinor().upload(chi[VH * 24:], 1)
    x = g.spinor()
    r2 = S.solve_dev(x, c)
    assert r1.n_count == r2.n_count and r1.rel_resid == r2.rel_resid
    assert np.array_equal(x.download(1), psi[VH * 24:])


def test_initial_guess_is_used(gauge, chi):
    S = q.SysSolver(L, xml(extra="<AntiPeriodicT>true</AntiPeriodicT>"), gauge)
    psi = np.zeros(2 * VH * 24)
    r1 = S(psi, chi)
    r2 = S(psi, chi)  # starts from the solution
    assert r2.status == q.OK and r2.n_count < r1.n_count


def test_single_precision_outer(gauge, chi):
    """<OuterPrecision>f</OuterPrecision>: the whole solve in FP32 storage (the
    reference built with single-precision QDP++ types); the residual check
    then runs in FP32 storage with FP64 sums, so targets stay above 1e-7."""
    x = xml(rsd=1e-6, extra="<OuterPrecision>f</OuterPrecision><Compress12>true</Compress12>"
                            "<AntiPeriodicT>true</AntiPeriodicT>")
    S = q.SysSolver(L, x, gauge)
    psi = np.zeros(2 * VH * 24)
    r = S(psi, chi)
    assert r.status == q.OK
    Mo = oracle_op(gauge, t_boundary=-1.0, **ISO)
    assert rel(Mo(psi[VH * 24:], 1, 1), chi[VH * 24:]) < 5e-6
