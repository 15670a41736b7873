"""GPU parity of the transforms at longitude counts with prime factors above 5.

The reference's default ``nlon = 2 lmax + 2`` (src/config.jl:76-80) has such a
factor for most lmax; FFTW accepts every length (src/fftutils.jl:195-238).  The
LDS FFT runs generic prime passes (7 .. 61) or Bluestein's chirp convolution
(larger primes); both are checked here against the FP64 oracle (whose own FFT is
pinned to numpy at these lengths) at 1e-12 relative L2, in every direction and
element kind that goes through the longitude transform.
"""
import numpy as np
import pytest
import torch

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL = 1e-12

# (lmax, nlat, nlon): default nlon 2 lmax + 2 = 14 (2*7), 22 (2*11), 66 (2*3*11),
# 202 (2*101: Bluestein), 608 (2^5*19), 2002 (2*7*11*13), 2042 (2*1021: Bluestein)
DEFAULT_NLON = [(6, 7), (10, 12), (32, 33), (100, 101), (303, 304), (1000, 1001), (1020, 1021)]


@pytest.mark.parametrize("lmax,nlat", DEFAULT_NLON)
def test_scalar_default_nlon(shtk, oracle, cuda, lmax, nlat):
    cfg = shtk.create_gauss_config(lmax, nlat)
    assert cfg.nlon == 2 * lmax + 2
    ocfg = oracle.Config(lmax, nlat, nlon=cfg.nlon)
    print(f"lmax={lmax} nlat={nlat} nlon={cfg.nlon}", flush=True)
    alm = oracle.random_alm(lmax, lmax, seed=lmax, power=1.0)
    f_ref = ocfg.synthesis(alm)
    f = shtk.synthesis(cfg, to_dev(alm))
    e_syn = rel_l2(to_host(f), f_ref)
    g = np.asfortranarray(np.random.default_rng(lmax).standard_normal((nlat, cfg.nlon)))
    a = shtk.analysis(cfg, to_dev(g))
    e_ana = rel_l2(to_host(a), ocfg.analysis(g))
    print(f"  synthesis {e_syn:.2e} analysis {e_ana:.2e}")
    assert e_syn < TOL and e_ana < TOL


@pytest.mark.parametrize("lmax,nlat", [(10, 12), (100, 101), (303, 304)])
def test_sphtor_default_nlon(shtk, oracle, cuda, lmax, nlat):
    cfg = shtk.create_gauss_config(lmax, nlat)
    ocfg = oracle.Config(lmax, nlat, nlon=cfg.nlon)
    S = oracle.random_alm(lmax, lmax, 7, 2.0, lmin=1)
    T = oracle.random_alm(lmax, lmax, 8, 2.0, lmin=1)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor(cfg, to_dev(S), to_dev(T))
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    rng = np.random.default_rng(9)
    gt = np.asfortranarray(rng.standard_normal((nlat, cfg.nlon)))
    gp = np.asfortranarray(rng.standard_normal((nlat, cfg.nlon)))
    S_r, T_r = ocfg.analysis_sphtor(gt, gp)
    S2, T2 = shtk.analysis_sphtor(cfg, to_dev(gt), to_dev(gp))
    assert rel_l2(to_host(S2), S_r) < TOL and rel_l2(to_host(T2), T_r) < TOL


@pytest.mark.parametrize("lmax,nlat,nlon", [(10, 12, 22), (100, 101, 202), (40, 42, 97), (63, 64, 128)])
def test_complex_grids(shtk, oracle, cuda, lmax, nlat, nlon):
    """Complex-valued synthesis (real_output=false) and analysis of complex fields
    (bins m >= 0 of the complex FFT, src/core_transforms.jl:169-207)."""
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=3, power=1.0)
    fc = shtk.synthesis(cfg, to_dev(alm), real_output=False)
    assert rel_l2(to_host(fc), ocfg.synthesis(alm, real_output=False)) < TOL
    rng = np.random.default_rng(nlon)
    g = np.asfortranarray(rng.standard_normal((nlat, nlon)) + 1j * rng.standard_normal((nlat, nlon)))
    a = shtk.analysis(cfg, to_dev(g))
    assert rel_l2(to_host(a), ocfg.analysis_cplx(g)) < TOL
    # FP32 I/O through the same kernels
    a32 = shtk.analysis(cfg, to_dev(g.astype(np.complex64)))
    assert a32.dtype == torch.complex64
    assert rel_l2(to_host(a32), ocfg.analysis_cplx(g.astype(np.complex64).astype(np.complex128))) < 1e-6


def test_complex_analysis_fixture(shtk, oracle, cuda):
    """The SHTns complex fixture's field analyses to the m >= 0 half of its coefficients."""
    import _fixtures as FX
    from test_oracle import cplx_split

    fx = FX.fixture("scalar_complex_full")
    p = FX.payloads(fx)
    cfg = shtk.create_gauss_config(fx["lmax"], fx["nlat"], nlon=fx["nphi"])
    A, _ = cplx_split(fx["lmax"], p["coefficients"])
    a = to_host(shtk.analysis(cfg, to_dev(np.asfortranarray(p["field"]))))
    assert np.abs(a - A).max() < 1e-12


@pytest.mark.parametrize("lmax,nlat", [(100, 101), (10, 12)])
def test_split_batch_default_nlon(shtk, oracle, cuda, lmax, nlat):
    cfg = shtk.create_gauss_config(lmax, nlat)
    ocfg = oracle.Config(lmax, nlat, nlon=cfg.nlon)
    alm = oracle.random_alm(lmax, lmax, seed=5, power=1.0, nf=3)
    f = shtk.synthesis_batch(cfg, to_dev(alm), precision="split")
    fh = to_host(f)
    ref = ocfg.synthesis(alm)
    back = to_host(shtk.analysis_batch(cfg, f, precision="split"))
    for k in range(3):
        assert rel_l2(fh[..., k], ref[..., k]) < 1e-11
        assert rel_l2(back[..., k], alm[..., k]) < 1e-11


def test_fp32_io_bluestein(shtk, oracle, cuda):
    lmax, nlat = 100, 101
    cfg = shtk.create_gauss_config(lmax, nlat)
    ocfg = oracle.Config(lmax, nlat, nlon=cfg.nlon)
    alm = oracle.random_alm(lmax, lmax, seed=6, power=1.0).astype(np.complex64)
    f = shtk.synthesis(cfg, to_dev(alm))
    assert f.dtype == torch.float32
    assert rel_l2(to_host(f), ocfg.synthesis(alm.astype(np.complex128))) < 1e-6
