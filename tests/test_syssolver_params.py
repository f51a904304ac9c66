"""The <InvertParam> reader of the Chroma LinOpSysSolver surface
(README.markdown:163-256) and its coefficient mapping
(tests/clover_term_qdp_w.h:156-180), host side only: no GPU is touched."""
import pytest

import qphix_amd as q

TWO_GROUPS = """<?xml version="1.0" encoding="UTF-8"?>
<!-- two solver groups: the first in document order is the one read -->
<Params version="3">
  <Other><InvertParam><MaxIter>3</MaxIter></InvertParam></Other>
  <InvertParam><invType>QPHIX_CLOVER_INVERTER</invType><MaxIter>3</MaxIter>
    <RsdTarget>1e-6</RsdTarget><CloverParams><Mass>0</Mass><clovCoeff>1</clovCoeff></CloverParams>
  </InvertParam>
</Params>
"""

REFINE = """<?xml version="1.0"?>
<Propagator>
 <InvertParam kind='clover'>
  <invType>QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER</invType>
  <SolverType>CG</SolverType>
  <MaxIter>700</MaxIter>
  <RsdTarget>2.5e-9</RsdTarget>
  <Delta>0.05</Delta>
  <CloverParams>
    <Mass>-0.0125</Mass>
    <clovCoeffR>1.75</clovCoeffR>
    <clovCoeffT>0.875</clovCoeffT>
    <AnisoParam>
      <anisoP>true</anisoP>
      <t_dir>3</t_dir>
      <xi_0>3.5</xi_0>
      <nu>0.75</nu>
    </AnisoParam>
  </CloverParams>
  <AntiPeriodicT>true</AntiPeriodicT>
  <Verbose>false</Verbose>
  <NCores>32</NCores>
  <ThreadsPerCore>4</ThreadsPerCore>
  <By>4</By> <Bz>2</Bz> <Sy>2</Sy> <Sz>1</Sz>
  <PadXY>0</PadXY> <PadXYZ>1</PadXYZ> <MinCt>2</MinCt>
  <RsdToleranceFactor>20.0</RsdToleranceFactor>
  <Tune>true</Tune>
  <InnerPrecision>h</InnerPrecision>
  <Compress12>true</Compress12>
 </InvertParam>
</Propagator>
"""


def plain(body, clover="<Kappa>0.125</Kappa><clovCoeff>1.5</clovCoeff>"):
    return ("<InvertParam><invType>QPHIX_CLOVER_INVERTER</invType><MaxIter>100</MaxIter>"
            f"<RsdTarget>1e-8</RsdTarget><CloverParams>{clover}</CloverParams>{body}</InvertParam>")


def test_read_iter_refine_group():
    p = q.InvertParam.from_xml(REFINE)
    assert p.inv_type == q.INV_CLOVER_ITER_REFINE and p.solver_type == q.SOLVER_CG
    assert p.max_iter == 700 and p.rsd_target == 2.5e-9 and p.delta == 0.05
    assert p.mass == -0.0125 and p.clov_coeff_r == 1.75 and p.clov_coeff_t == 0.875
    assert p.aniso_p == 1 and p.t_dir == 3 and p.xi_0 == 3.5 and p.nu == 0.75
    assert p.antiperiodic_t == 1 and p.verbose == 0 and p.tune == 1
    assert (p.ncores, p.threads_per_core, p.by, p.bz, p.sy, p.sz) == (32, 4, 4, 2, 2, 1)
    assert (p.pad_xy, p.pad_xyz, p.min_ct) == (0, 1, 2)
    assert p.rsd_tolerance_factor == 20.0
    assert (p.prec, p.compress12, p.inner_prec, p.inner_compress12) == (q.F64, 1, q.F16, 1)


def test_defaults_of_optional_tags():
    p = q.InvertParam.from_xml(plain(""))
    d = q.InvertParam.default()
    assert p.inv_type == q.INV_CLOVER and p.solver_type == q.SOLVER_BICGSTAB
    assert p.max_iter == 100 and p.rsd_target == 1e-8
    for f in ("delta", "aniso_p", "t_dir", "xi_0", "nu", "antiperiodic_t", "verbose", "tune",
              "rsd_tolerance_factor", "ncores", "prec", "compress12", "inner_prec",
              "inner_compress12"):
        assert getattr(p, f) == getattr(d, f), f
    assert d.rsd_tolerance_factor == 5.0 and d.delta == 0.1 and d.inner_prec == q.F32


@pytest.mark.parametrize("kappa,mass", [(0.125, 0.0), (0.1, 1.0), (0.1240, 1 / 0.248 - 4)])
def test_kappa_is_converted_to_mass(kappa, mass):
    p = q.InvertParam.from_xml(plain("", f"<Kappa>{kappa}</Kappa><clovCoeff>1</clovCoeff>"))
    assert p.mass == pytest.approx(mass, abs=1e-15)


def test_clover_coefficient_tags():
    # clovCoeff sets both planes; clovCoeffR / clovCoeffT set their own
    p = q.InvertParam.from_xml(plain("", "<Mass>0</Mass><clovCoeff>2.0</clovCoeff>"))
    assert (p.clov_coeff_r, p.clov_coeff_t) == (2.0, 2.0)
    p = q.InvertParam.from_xml(plain("", "<Mass>0</Mass><clovCoeff>2.0</clovCoeff>"
                                         "<clovCoeffT>0.5</clovCoeffT>"))
    assert (p.clov_coeff_r, p.clov_coeff_t) == (2.0, 0.5)
    p = q.InvertParam.from_xml(plain("", "<Mass>0</Mass><clovCoeffR>1.25</clovCoeffR>"
                                         "<clovCoeffT>0.5</clovCoeffT>"))
    assert (p.clov_coeff_r, p.clov_coeff_t) == (1.25, 0.5)


def test_xml_syntax_that_is_accepted():
    xml = ("<?xml version='1.0'?>\n<!DOCTYPE Params>\n<Params><InvertParam a=\"1\">"
           "<invType> QPHIX_CLOVER_INVERTER </invType><!-- comment <x> -->"
           "<MaxIter>\n  42\n</MaxIter><RsdTarget><![CDATA[1e-9]]></RsdTarget>"
           "<CloverParams><Mass>&#48;.5</Mass><clovCoeff>1</clovCoeff><AnisoParam>"
           "<anisoP>false</anisoP></AnisoParam></CloverParams><Verbose/></InvertParam>"
           "</Params>\n<!-- trailing -->\n")
    with pytest.raises(q.QphixError) as e:  # an empty <Verbose/> is not a boolean
        q.InvertParam.from_xml(xml)
    assert e.value.code == q.EINVAL and "Verbose" in str(e.value)
    p = q.InvertParam.from_xml(xml.replace("<Verbose/>", "<Verbose>1</Verbose>"))
    assert p.max_iter == 42 and p.rsd_target == 1e-9 and p.mass == 0.5 and p.verbose == 1
    assert p.aniso_p == 0


def test_first_invert_param_in_document_order():
    with pytest.raises(q.QphixError) as e:
        q.InvertParam.from_xml(TWO_GROUPS)
    assert e.value.code == q.EINVAL and "invType" in str(e.value)


@pytest.mark.parametrize("xml,code,words", [
    # a comment opened with '<--' instead of '<!--' is not XML
    (plain("<AntiPeriodicT>false</AntiPeriodicT>\n<-- note -->"), q.EINVAL, ["line 2", "element name"]),
    (plain("<Verbose>true</Verbos>"), q.EINVAL, ["closing tag", "Verbose"]),
    (plain("<Verbose>true</InvertParam>"), q.EINVAL, ["</InvertParam>", "<Verbose>"]),
    (plain("")[:-len("</InvertParam>")], q.EINVAL, ["<InvertParam>", "not closed"]),
    (plain("", "<Kappa>0.12</Kappa><Mass>0.1</Mass><clovCoeff>1</clovCoeff>"), q.EDOMAIN, ["Mass and Kappa"]),
    (plain("", "<clovCoeff>1</clovCoeff>"), q.EINVAL, ["neither Mass nor Kappa"]),
    (plain("", "<Mass>0.1</Mass><clovCoeffR>1</clovCoeffR>"), q.EINVAL, ["clovCoeff"]),
    (plain("", "<Mass>0.1</Mass><clovCoeff>1</clovCoeff><AnisoParam><anisoP>true</anisoP>"
               "<t_dir>2</t_dir><xi_0>2</xi_0><nu>1</nu></AnisoParam>"), q.EDOMAIN, ["t_dir"]),
    (plain("", "<Mass>0.1</Mass><clovCoeff>1</clovCoeff><AnisoParam><anisoP>true</anisoP>"
               "</AnisoParam>"), q.EINVAL, ["xi_0"]),
    (plain("<MaxIter>5</MaxIter>"), q.EINVAL, ["MaxIter", "twice"]),
    (plain("").replace("1e-8", "1e-8x"), q.EINVAL, ["RsdTarget", "not a number"]),
    (plain("").replace("QPHIX_CLOVER_INVERTER", "QPHIX_TWISTED_MASS_INVERTER"), q.EDOMAIN, ["invType"]),
    (plain("<SolverType>GMRES</SolverType>"), q.EDOMAIN, ["SolverType"]),
    (plain("").replace("<MaxIter>100</MaxIter>", ""), q.EINVAL, ["MaxIter"]),
    (plain("<AntiPeriodicT>yes</AntiPeriodicT>"), q.EINVAL, ["boolean"]),
    (plain("<InnerPrecision>q</InnerPrecision>"), q.EDOMAIN, ["InnerPrecision"]),
    (plain("<NCores>0</NCores>"), q.EDOMAIN, ["blocking"]),
    (plain("<Delta>1.5</Delta>").replace("QPHIX_CLOVER_INVERTER",
                                         "QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER"), q.EDOMAIN, ["Delta"]),
    (plain("<InnerPrecision>h</InnerPrecision>").replace(
        "QPHIX_CLOVER_INVERTER", "QPHIX_CLOVER_ITER_REFINE_BICGSTAB_INVERTER"), q.EDOMAIN, ["compression"]),
    (plain("<!-- open"), q.EINVAL, ["comment"]),
    ("<Params></Params>", q.EINVAL, ["no <InvertParam>"]),
    (plain("") + "<Extra/>", q.EINVAL, ["after the root"]),
    (plain("<Verbose>&bogus;</Verbose>"), q.EINVAL, ["entity"]),
])
def test_rejected_input(xml, code, words):
    with pytest.raises(q.QphixError) as e:
        q.InvertParam.from_xml(xml)
    assert e.value.code == code, str(e.value)
    for w in words:
        assert w in str(e.value), (w, str(e.value))


def test_coefficients_follow_the_clover_term_conventions():
    """tests/clover_term_qdp_w.h:156-180 with Chroma's anisotropy: spatial
    hops nu/xi_0, spatial clover planes c_R/xi_0, diagonal 1 + 3 nu/xi_0 + m."""
    p = q.InvertParam.from_xml(REFINE)
    a_s, a_t, tb, cr, ct, dm = p.coeffs()
    assert a_s == pytest.approx(0.75 / 3.5, rel=1e-15) and a_t == 1.0 and tb == -1.0
    assert cr == pytest.approx(1.75 / 3.5, rel=1e-15) and ct == 0.875
    assert dm == pytest.approx(1 + 3 * 0.75 / 3.5 - 0.0125, rel=1e-15)
    # isotropic: the usual 4 + m = 1/(2 kappa), c_sw on every plane
    p = q.InvertParam.from_xml(plain(""))
    assert p.coeffs() == (1.0, 1.0, 1.0, 1.5, 1.5, 4.0)


def test_create_rejects_bad_params_without_a_gpu():
    lib = q.lib()
    import ctypes as C
    import numpy as np
    p = q.InvertParam.default(t_dir=1)
    u = np.zeros(16)
    h = C.c_void_p()
    assert lib.qphix_syssolver_create((C.c_int * 4)(4, 4, 4, 4), C.byref(p), q.dptr(u), C.byref(h)) == q.EDOMAIN
    assert b"t_dir" in lib.qphix_last_error()
    p = q.InvertParam.default(max_iter=0)
    assert lib.qphix_syssolver_create((C.c_int * 4)(4, 4, 4, 4), C.byref(p), q.dptr(u), C.byref(h)) == q.EDOMAIN
