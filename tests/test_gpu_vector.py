"""GPU parity of the spheroidal/toroidal (sphtor) transforms and their adjoints
against the FP64 oracle (C ABI -> HIP kernels).  Tolerance 1e-12 relative L2 for
FP64 I/O."""
import numpy as np
import pytest

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _st(oracle, lmax, mmax, seed, nf=None):
    S = oracle.random_alm(lmax, mmax, seed, 2.0, nf=nf, lmin=1)
    T = oracle.random_alm(lmax, mmax, seed + 1, 2.0, nf=nf, lmin=1)
    return S, T


@pytest.mark.parametrize("lmax,nlat,nlon,mmax,mres", [
    (3, 8, 10, None, 1),
    (3, 9, 12, None, 1),
    (40, 42, 96, None, 1),
    (63, 64, 128, 50, 1),
    (60, 63, 128, 60, 2),
])
def test_sphtor_match_oracle(shtk, oracle, cuda, lmax, nlat, nlon, mmax, mres):
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    mm = lmax if mmax is None else mmax
    S, T = _st(oracle, lmax, mm, 21)
    for m in range(mm + 1):
        if m % mres:
            S[:, m] = 0
            T[:, m] = 0
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor(cfg, to_dev(S), to_dev(T))
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    rng = np.random.default_rng(4)
    gt = np.asfortranarray(rng.standard_normal((nlat, nlon)))
    gp = np.asfortranarray(rng.standard_normal((nlat, nlon)))
    S_r, T_r = ocfg.analysis_sphtor(gt, gp)
    S2, T2 = shtk.analysis_sphtor(cfg, to_dev(gt), to_dev(gp))
    assert rel_l2(to_host(S2), S_r) < TOL and rel_l2(to_host(T2), T_r) < TOL


def test_sphtor_lmax511_batch(shtk, oracle, cuda):
    """The vector benchmark grid (lmax=511, 512 x 1024), a few fields, l^-2 spectra."""
    lmax, nlat, nlon, nf = 511, 512, 1024, 3
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    S, T = _st(oracle, lmax, lmax, 4, nf=nf)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T))
    e1, e2 = rel_l2(to_host(Vt), Vt_r), rel_l2(to_host(Vp), Vp_r)
    S2, T2 = shtk.analysis_sphtor_batch(cfg, Vt, Vp)
    e3, e4 = rel_l2(to_host(S2), S), rel_l2(to_host(T2), T)
    S_r, T_r = ocfg.analysis_sphtor(Vt_r, Vp_r)
    e5, e6 = rel_l2(to_host(S2), S_r), rel_l2(to_host(T2), T_r)
    print(f"sphtor lmax=511: synth {e1:.2e} {e2:.2e} round trip {e3:.2e} {e4:.2e} analysis {e5:.2e} {e6:.2e}")
    assert max(e1, e2, e5, e6) < TOL and max(e3, e4) < TOL


def test_sphtor_lmax1023(shtk, oracle, cuda):
    lmax, nlat, nlon = 1023, 1024, 2048
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    S, T = _st(oracle, lmax, lmax, 8)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor(cfg, to_dev(S), to_dev(T))
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    S2, T2 = shtk.analysis_sphtor(cfg, Vt, Vp)
    assert rel_l2(to_host(S2), S) < TOL and rel_l2(to_host(T2), T) < TOL


@pytest.mark.parametrize("robert", [False, True])
@pytest.mark.parametrize("norm,cs,rn", [("orthonormal", True, False), ("schmidt", False, True)])
def test_sphtor_conventions_and_robert(shtk, oracle, cuda, robert, norm, cs, rn):
    lmax, nlat, nlon = 20, 22, 48
    kw = dict(nlon=nlon, norm=norm, cs_phase=cs, real_norm=rn, robert_form=robert)
    cfg = shtk.create_gauss_config(lmax, nlat, **kw)
    ocfg = oracle.Config(lmax, nlat, **kw)
    S, T = _st(oracle, lmax, lmax, 31)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor(cfg, to_dev(S), to_dev(T))
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    S2, T2 = shtk.analysis_sphtor(cfg, to_dev(Vt_r), to_dev(Vp_r))
    S_r, T_r = ocfg.analysis_sphtor(Vt_r, Vp_r)
    assert rel_l2(to_host(S2), S_r) < TOL and rel_l2(to_host(T2), T_r) < TOL
    rng = np.random.default_rng(6)
    yt = np.asfortranarray(rng.standard_normal((nlat, nlon)))
    yp = np.asfortranarray(rng.standard_normal((nlat, nlon)))
    Sb_r, Tb_r = ocfg.adjoint_synthesis_sphtor(yt, yp)
    Sb, Tb = shtk._adjoint_synthesis_sphtor(cfg, to_dev(yt), to_dev(yp))
    assert rel_l2(to_host(Sb), Sb_r) < TOL and rel_l2(to_host(Tb), Tb_r) < TOL
    vt_r, vp_r = ocfg.adjoint_analysis_sphtor(S, T)
    vt, vp = shtk._adjoint_analysis_sphtor(cfg, to_dev(S), to_dev(T), project_real=True)
    assert rel_l2(to_host(vt), vt_r.real) < TOL and rel_l2(to_host(vp), vp_r.real) < TOL
    vt, vp = shtk._adjoint_analysis_sphtor(cfg, to_dev(S), to_dev(T))
    assert rel_l2(to_host(vt), vt_r) < TOL and rel_l2(to_host(vp), vp_r) < TOL


def test_segment_lengths(shtk, oracle, cuda, monkeypatch):
    """Every segment length gives the same transform (segment seeds and
    boundaries are exercised at many degrees)."""
    lmax, nlat, nlon = 300, 302, 640
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=12, power=1.0)
    f_ref = ocfg.synthesis(alm)
    a_ref = ocfg.analysis(f_ref)
    S, T = _st(oracle, lmax, lmax, 13)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    for seg in ("8", "24", "128", "256"):
        monkeypatch.setenv("SHTK_SEG", seg)
        cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
        f = shtk.synthesis(cfg, to_dev(alm))
        assert rel_l2(to_host(f), f_ref) < TOL, seg
        assert rel_l2(to_host(shtk.analysis(cfg, f)), a_ref) < TOL, seg
        Vt, Vp = shtk.synthesis_sphtor(cfg, to_dev(S), to_dev(T))
        assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL, seg
        cfg.close()


def test_high_lmax_polar_underflow(shtk, oracle, cuda):
    """lmax = 2047: sectoral seeds underflow FP64 by thousands of binary orders
    near the poles (the scaled-exponent path) and analysis uses several pair
    blocks (accumulating kernel)."""
    lmax, nlat, nlon = 2047, 2048, 4096
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=14, power=1.5)
    f_ref = ocfg.synthesis(alm)
    f = shtk.synthesis(cfg, to_dev(alm))
    assert rel_l2(to_host(f), f_ref) < TOL
    a = shtk.analysis(cfg, f)
    assert rel_l2(to_host(a), alm) < TOL


@pytest.mark.parametrize("robert", [False, True])
def test_sphtor_complex_output(shtk, oracle, cuda, robert):
    """synthesis_sphtor(...; real_output=false): complex grids of the bins m >= 0
    (src/sphtor_transforms.jl:162-241), plain FP64 and split precision."""
    lmax, nlat, nlon, nf = 40, 42, 96, 2
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, robert_form=robert)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, robert_form=robert)
    S, T = _st(oracle, lmax, lmax, 41, nf=nf)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T, real_output=False)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), real_output=False)
    assert to_host(Vt).dtype == np.complex128
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    Vt2, Vp2 = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), real_output=False, precision="split")
    assert rel_l2(to_host(Vt2), Vt_r) < 1e-11 and rel_l2(to_host(Vp2), Vp_r) < 1e-11


def test_split_scalar_complex_output(shtk, oracle, cuda):
    lmax, nlat, nlon, nf = 63, 64, 128, 3
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, 42, 1.0, nf=nf)
    ref = ocfg.synthesis(alm, real_output=False)
    got = to_host(shtk.synthesis_batch(cfg, to_dev(alm), real_output=False, precision="split"))
    assert rel_l2(got, ref) < 1e-11


@pytest.mark.parametrize("robert", [False, True])
def test_sphtor_zonal_lmax1023(shtk, oracle, cuda, robert):
    """Order 0 alone at lmax 1023: dP/dtheta of order 0 must keep its accuracy near
    the poles (the differentiated recurrence of _zonal_dtheta_row!), which the
    mixed random spectra of the other tests would dilute."""
    lmax, nlat, nlon = 1023, 1024, 2048
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, robert_form=robert)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, robert_form=robert)
    S, T = _st(oracle, lmax, lmax, 41)
    S[:, 1:] = 0
    T[:, 1:] = 0
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor(cfg, to_dev(S), to_dev(T))
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    S2, T2 = shtk.analysis_sphtor(cfg, to_dev(Vt_r), to_dev(Vp_r))
    S_r, T_r = ocfg.analysis_sphtor(Vt_r, Vp_r)
    assert rel_l2(to_host(S2), S_r) < TOL and rel_l2(to_host(T2), T_r) < TOL


@pytest.mark.parametrize("grid", ["gauss", "driscoll_healy"])
def test_sphtor_zonal_segments_f32_and_truncation(shtk, oracle, cuda, grid):
    """Order 0 runs in parallel 64-degree segments from plan-time start values:
    a zonal-only field over several segments with FP32 coefficients, truncation
    inside a segment (ltr 200 of lmax 300) and, on the Driscoll-Healy grid, the
    missing south-pole row.  FP32 inputs: 1e-6; FP64: 1e-12."""
    lmax, ltr = 300, 200
    if grid == "gauss":
        nlat, nlon = 302, 608
        cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
        ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    else:
        nlat, nlon = 2 * (lmax + 1), 608
        cfg = shtk.create_regular_config(lmax, nlat, nlon=nlon, include_poles=True, use_dh_weights=True)
        ocfg = oracle.Config(lmax, nlat, nlon=nlon, grid="driscoll_healy")
    S, T = _st(oracle, lmax, lmax, 57)
    S[:, 1:] = 0
    T[:, 1:] = 0
    S32, T32 = S.astype(np.complex64), T.astype(np.complex64)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S32.astype(np.complex128), T32.astype(np.complex128))
    Vt, Vp = shtk.synthesis_sphtor(cfg, to_dev(S32), to_dev(T32))
    assert rel_l2(to_host(Vt), Vt_r) < 1e-6 and rel_l2(to_host(Vp), Vp_r) < 1e-6
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T, ltr=ltr)
    Vt, Vp = shtk.synthesis_sphtor_l(cfg, to_dev(S), to_dev(T), ltr)
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    S_r, T_r = ocfg.analysis_sphtor(Vt_r, Vp_r, ltr=ltr)
    S2, T2 = shtk.analysis_sphtor_l(cfg, to_dev(Vt_r), to_dev(Vp_r), ltr)
    assert rel_l2(to_host(S2), S_r) < TOL and rel_l2(to_host(T2), T_r) < TOL
