"""GPU parity of the scalar transforms (C ABI -> HIP kernels) against the FP64 oracle.

Tolerance: the plain FP64 path must reproduce the oracle to <= 1e-12 relative L2
(the bar for the plain method); float32 I/O is checked at 1e-6.
"""
import numpy as np
import pytest

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.mark.parametrize("lmax,nlat,nlon,mmax,mres", [
    (2, 6, 8, None, 1),        # fixture-sized
    (3, 8, 10, None, 1),       # nlon with a factor 5
    (3, 9, 12, None, 1),       # odd nlat (equator node unpaired), factor 3
    (31, 32, 64, None, 1),
    (63, 64, 128, 40, 1),      # mmax < lmax
    (64, 67, 160, 60, 3),      # mres > 1, odd nlat
    (127, 128, 256, None, 1),  # the reference's CPU-runnable configuration
])
def test_synthesis_analysis_match_oracle(shtk, oracle, cuda, lmax, nlat, nlon, mmax, mres):
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    mm = lmax if mmax is None else mmax
    alm = oracle.random_alm(lmax, mm, seed=11, power=1.0)
    for m in range(mm + 1):
        if m % mres:
            alm[:, m] = 0
    f_ref = ocfg.synthesis(alm)
    f = shtk.synthesis(cfg, to_dev(alm))
    assert rel_l2(to_host(f), f_ref) < TOL
    rng = np.random.default_rng(5)
    g = np.asfortranarray(rng.standard_normal((nlat, nlon)))
    a_ref = ocfg.analysis(g)
    a = shtk.analysis(cfg, to_dev(g))
    assert rel_l2(to_host(a), a_ref) < TOL


def test_lmax1023_plain_fp64(shtk, oracle, cuda):
    """Configuration of the headline metric: lmax=mmax=1023, 1024 x 2048, spectrum (l+1)^-1.5."""
    lmax, nlat, nlon = 1023, 1024, 2048
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=2, power=1.5)
    f_ref = ocfg.synthesis(alm)
    f = shtk.synthesis(cfg, to_dev(alm))
    e_syn = rel_l2(to_host(f), f_ref)
    a_ref = ocfg.analysis(f_ref)
    a = shtk.analysis(cfg, f)
    e_ana = rel_l2(to_host(a), a_ref)
    e_rt = rel_l2(to_host(a), alm)
    print(f"lmax=1023: synthesis {e_syn:.2e} analysis {e_ana:.2e} round trip {e_rt:.2e}")
    assert e_syn < TOL and e_ana < TOL and e_rt < TOL


@pytest.mark.parametrize("norm,cs,rn", [("fourpi", False, False), ("schmidt", True, True),
                                        ("orthonormal", True, True)])
def test_conventions(shtk, oracle, cuda, norm, cs, rn):
    lmax, nlat, nlon = 20, 24, 48
    kw = dict(nlon=nlon, norm=norm, cs_phase=cs, real_norm=rn)
    cfg = shtk.create_gauss_config(lmax, nlat, **kw)
    ocfg = oracle.Config(lmax, nlat, **kw)
    alm = oracle.random_alm(lmax, lmax, seed=3)
    assert rel_l2(to_host(shtk.synthesis(cfg, to_dev(alm))), ocfg.synthesis(alm)) < TOL
    g = np.asfortranarray(np.random.default_rng(1).standard_normal((nlat, nlon)))
    assert rel_l2(to_host(shtk.analysis(cfg, to_dev(g))), ocfg.analysis(g)) < TOL


def test_truncated_ltr(shtk, oracle, cuda):
    lmax, nlat, nlon, ltr = 40, 42, 96, 25
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=4)
    assert rel_l2(to_host(shtk.synthesis_l(cfg, to_dev(alm), ltr)), ocfg.synthesis(alm, ltr=ltr)) < TOL
    g = np.asfortranarray(np.random.default_rng(2).standard_normal((nlat, nlon)))
    assert rel_l2(to_host(shtk.analysis_l(cfg, to_dev(g), ltr)), ocfg.analysis(g, ltr=ltr)) < TOL


def test_batch_and_f32(shtk, oracle, cuda):
    import torch

    lmax, nlat, nlon, nf = 47, 48, 96, 5
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=7, nf=nf)
    f_ref = ocfg.synthesis(alm)
    f = shtk.synthesis_batch(cfg, to_dev(alm))
    assert rel_l2(to_host(f), f_ref) < TOL
    a = shtk.analysis_batch(cfg, f)
    assert rel_l2(to_host(a), ocfg.analysis(f_ref)) < TOL
    # float32 I/O (FP64 inside): error is the input/output rounding only
    f32 = shtk.synthesis_batch(cfg, to_dev(alm).to(torch.complex64))
    assert f32.dtype == torch.float32
    assert rel_l2(to_host(f32), f_ref) < 1e-6
    a32 = shtk.analysis_batch(cfg, f32)
    assert a32.dtype == torch.complex64
    assert rel_l2(to_host(a32), alm) < 1e-6


def test_adjoints_match_oracle(shtk, oracle, cuda):
    lmax, nlat, nlon = 30, 32, 64
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    rng = np.random.default_rng(9)
    ybar = np.asfortranarray(rng.standard_normal((nlat, nlon)))
    ref = ocfg.adjoint_synthesis(ybar)
    got = shtk._adjoint_synthesis(cfg, to_dev(ybar))
    assert rel_l2(to_host(got), ref) < TOL
    abar = oracle.random_alm(lmax, lmax, seed=10)
    ref = ocfg.adjoint_analysis(abar)
    got = shtk._adjoint_analysis(cfg, to_dev(abar))
    assert got.is_complex() and rel_l2(to_host(got), ref) < TOL
    got_re = shtk._adjoint_analysis(cfg, to_dev(abar), project_real=True)
    assert not got_re.is_complex() and rel_l2(to_host(got_re), ref.real) < TOL
    # a complex cotangent of the synthesis: the adjoint is complex-linear
    ycplx = ybar + 1j * np.asfortranarray(rng.standard_normal((nlat, nlon)))
    ref = ocfg.adjoint_synthesis(ycplx.real) + 1j * ocfg.adjoint_synthesis(ycplx.imag)
    assert rel_l2(to_host(shtk._adjoint_synthesis(cfg, to_dev(ycplx))), ref) < TOL


def test_errors(shtk, cuda):
    import torch

    cfg = shtk.create_gauss_config(7, 8, nlon=16)
    with pytest.raises(ValueError):
        shtk.synthesis(cfg, torch.zeros((9, 8), dtype=torch.complex128, device=cuda))
    with pytest.raises(ValueError):
        shtk.analysis(cfg, torch.zeros((8, 15), dtype=torch.float64, device=cuda))
    with pytest.raises(shtk.BackendUnavailableError):
        shtk.analysis(cfg, torch.zeros((8, 16), dtype=torch.float64))
    with pytest.raises(ValueError):
        shtk.create_gauss_config(7, 7)


def test_batch_odd_nlat_all_transforms(shtk, oracle, cuda):
    """A batch on an odd nlat (equator row without a partner) through the
    scalar and vector transforms."""
    lmax, nlat, nlon, nf = 40, 43, 90, 3
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=17, power=1.0, nf=nf)
    f = shtk.synthesis_batch(cfg, to_dev(alm))
    assert rel_l2(to_host(f), ocfg.synthesis(alm)) < 1e-12
    rng = np.random.default_rng(5)
    g = np.asfortranarray(rng.standard_normal((nlat, nlon, nf)))
    assert rel_l2(to_host(shtk.analysis_batch(cfg, to_dev(g))), ocfg.analysis(g)) < 1e-12
    S = oracle.random_alm(lmax, lmax, 18, 2.0, nf=nf, lmin=1)
    T = oracle.random_alm(lmax, lmax, 19, 2.0, nf=nf, lmin=1)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T))
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    assert rel_l2(to_host(Vt), Vt_r) < 1e-12 and rel_l2(to_host(Vp), Vp_r) < 1e-12
    cfg.close()


@pytest.mark.parametrize("chains", ["1", "2", "4"])
@pytest.mark.parametrize("fft_pairs", ["1", "2", "4"])
def test_launch_shapes(shtk, oracle, cuda, monkeypatch, chains, fft_pairs):
    """Every synthesis chain count (pairs per lane) and FFT tile width gives the
    same transform, including a partial last FFT tile (nlat = 302 is no multiple
    of 8) and a truncation degree inside a 16-degree record group."""
    lmax, nlat, nlon, ltr = 300, 302, 640, 290
    monkeypatch.setenv("SHTK_SYNTH_CHAINS", chains)
    monkeypatch.setenv("SHTK_FFT_PAIRS", fft_pairs)
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=14, power=1.0, nf=2)
    f_ref = ocfg.synthesis(alm, ltr=ltr)
    f = shtk.synthesis_batch(cfg, to_dev(alm), ltr=ltr)
    assert rel_l2(to_host(f), f_ref) < TOL
    a_ref = ocfg.analysis(f_ref)
    assert rel_l2(to_host(shtk.analysis_batch(cfg, f)), a_ref) < TOL
    cfg.close()


@pytest.mark.parametrize("units", ["1", "1073741824"])
@pytest.mark.parametrize("lmax,nlat,nlon,mmax,mres,ltr,nf", [
    (300, 302, 640, None, 1, 290, 3),  # truncation inside a segment, 2 pair blocks
    (64, 67, 160, 60, 3, -1, 2),       # mres > 1, odd nlat, mmax < lmax
    (199, 200, 402, None, 1, -1, 5),
])
def test_synthesis_waves_per_block(shtk, oracle, cuda, monkeypatch, units, lmax, nlat, nlon, mmax, mres, ltr, nf):
    """Scalar FP64 synthesis with blocks of one wave (each running all segments of
    its order, field-minor unit order; units = 1) and of four waves adding their
    segment sums in LDS (units above any launch) against the oracle, per field."""
    monkeypatch.setenv("SHTK_SSYNTH_1W_UNITS", units)
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    mm = lmax if mmax is None else mmax
    alm = oracle.random_alm(lmax, mm, seed=lmax + nf, power=1.0, nf=nf)
    for m in range(mm + 1):
        if m % mres:
            alm[:, m] = 0
    f_ref = ocfg.synthesis(alm, ltr=ltr)
    f = to_host(shtk.synthesis_batch(cfg, to_dev(alm), ltr=ltr))
    for k in range(nf):
        assert rel_l2(f[..., k], f_ref[..., k]) < TOL, k
    cfg.close()


@pytest.mark.parametrize("lmax,nlat,nlon,mmax,mres,nf", [
    (3, 8, 10, None, 1, 1),
    (40, 42, 96, None, 1, 3),
    (60, 63, 128, 60, 2, 2),
    (1023, 1024, 2048, None, 1, 1),
])
def test_complex_output_synthesis(shtk, oracle, cuda, lmax, nlat, nlon, mmax, mres, nf):
    """real_output = false / synthesis_cplx: the complex grid of the bins m >= 0
    (src/core_transforms.jl:289-344), FP64 and FP32 I/O."""
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    mm = lmax if mmax is None else mmax
    alm = oracle.random_alm(lmax, mm, 23, 1.0, nf=nf)
    for m in range(mm + 1):
        if m % mres:
            alm[:, m] = 0
    ref = ocfg.synthesis(alm, real_output=False)
    got = to_host(shtk.synthesis_batch(cfg, to_dev(alm), real_output=False))
    assert got.dtype == np.complex128
    assert rel_l2(got, ref) < TOL
    if nf == 1:
        z = to_host(shtk.synthesis_cplx(cfg, to_dev(alm[:, :, 0])))
        assert rel_l2(z, ref[:, :, 0]) < TOL
    got32 = to_host(shtk.synthesis_batch(cfg, to_dev(alm.astype(np.complex64)), real_output=False))
    assert got32.dtype == np.complex64 and rel_l2(got32, ref) < 1e-6
