"""GPU parity of the batched split-precision (Ozaki, int8 matrix core) synthesis
against the FP64 oracle, through the C ABI.

The split contraction is FP64-grade: every field must match the oracle to a
relative L2 error of 1e-11 (the target stated for the split path); the FP64
Legendre values, seeds, exponents and cross-slice sums keep it there."""
import numpy as np
import pytest

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL_SPLIT = 1e-11


def _per_field(out, ref):
    return [rel_l2(out[..., k], ref[..., k]) for k in range(ref.shape[-1])]


@pytest.mark.parametrize("lmax,nlat,nlon,nf,power", [
    (3, 8, 10, 1, 1.0),
    (63, 64, 128, 3, 1.0),
    (100, 128, 256, 17, 1.5),
    (300, 302, 640, 20, 1.0),
])
def test_split_synthesis_matches_oracle(shtk, oracle, cuda, lmax, nlat, nlon, nf, power):
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=5, power=power, nf=nf)
    ref = ocfg.synthesis(alm)
    if ref.ndim == 2:
        ref = ref[..., None]
    got = to_host(shtk.synthesis_batch(cfg, to_dev(alm if alm.ndim == 3 else alm[..., None]),
                                       precision="split"))
    errs = _per_field(got, ref)
    print(f"split lmax={lmax} nf={nf}: max per-field rel L2 {max(errs):.2e}")
    assert max(errs) < TOL_SPLIT, errs


def test_split_synthesis_lmax1023(shtk, oracle, cuda):
    """The headline configuration: lmax 1023, 1024 x 2048, 16 fields (one field
    group), (l+1)^-1.5 spectra."""
    lmax, nlat, nlon, nf = 1023, 1024, 2048, 16
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=2, power=1.5, nf=nf)
    got = to_host(shtk.synthesis_batch(cfg, to_dev(alm), precision="split"))
    errs = []
    for k in (0, 7, 15):
        errs.append(rel_l2(got[..., k], ocfg.synthesis(alm[..., k])))
    print(f"split lmax=1023: per-field rel L2 {['%.2e' % e for e in errs]}")
    assert max(errs) < TOL_SPLIT


def test_split_matches_plain_batch(shtk, oracle, cuda):
    """Split and FP64 batch paths agree; degree truncation and an mres > 1 plan."""
    lmax, nlat, nlon, nf = 60, 63, 128, 5
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, mres=2)
    alm = oracle.random_alm(lmax, lmax, seed=9, power=1.0, nf=nf)
    alm[:, 1::2, :] = 0
    d = to_dev(alm)
    for ltr in (-1, 41):
        a = to_host(shtk.synthesis_batch(cfg, d, precision="fp64", ltr=ltr))
        b = to_host(shtk.synthesis_batch(cfg, d, precision="split", ltr=ltr))
        assert rel_l2(b, a) < TOL_SPLIT, ltr


def test_split_conventions_and_f32(shtk, oracle, cuda):
    lmax, nlat, nlon, nf = 40, 42, 96, 4
    kw = dict(nlon=nlon, norm="schmidt", cs_phase=False, real_norm=True)
    cfg = shtk.create_gauss_config(lmax, nlat, **kw)
    ocfg = oracle.Config(lmax, nlat, **kw)
    alm = oracle.random_alm(lmax, lmax, seed=3, power=1.0, nf=nf)
    ref = ocfg.synthesis(alm)
    got = to_host(shtk.synthesis_batch(cfg, to_dev(alm), precision="split"))
    assert max(_per_field(got, ref)) < TOL_SPLIT
    got32 = to_host(shtk.synthesis_batch(cfg, to_dev(alm.astype(np.complex64)), precision="split"))
    assert got32.dtype == np.float32
    assert max(_per_field(got32.astype(np.float64), ref)) < 1e-6


@pytest.mark.parametrize("lmax,nlat,nlon,nf", [
    (3, 8, 10, 1),
    (63, 64, 128, 3),
    (100, 128, 256, 17),
    (300, 302, 640, 20),
])
def test_split_analysis_matches_oracle(shtk, oracle, cuda, lmax, nlat, nlon, nf):
    """Split analysis of a non-band-limited grid (random samples, so the
    quadrature itself is exercised) against the oracle."""
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    rng = np.random.default_rng(lmax + nf)
    f = np.asfortranarray(rng.standard_normal((nlat, nlon, nf)))
    ref = ocfg.analysis(f)
    got = to_host(shtk.analysis_batch(cfg, to_dev(f), precision="split"))
    errs = _per_field(got, ref)
    print(f"split analysis lmax={lmax} nf={nf}: max per-field rel L2 {max(errs):.2e}")
    assert max(errs) < TOL_SPLIT, errs


def test_split_round_trip_lmax1023(shtk, oracle, cuda):
    """Split synthesis then split analysis at lmax 1023 returns the spectrum
    (band-limited round trip) and matches the oracle analysis of the grid."""
    lmax, nlat, nlon, nf = 1023, 1024, 2048, 16
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=2, power=1.5, nf=nf)
    f = shtk.synthesis_batch(cfg, to_dev(alm), precision="split")
    back = to_host(shtk.analysis_batch(cfg, f, precision="split"))
    errs = _per_field(back, alm)
    print(f"split round trip lmax=1023: max per-field rel L2 {max(errs):.2e}")
    assert max(errs) < TOL_SPLIT
    fh = to_host(f)
    e = rel_l2(back[..., 3], ocfg.analysis(fh[..., 3]))
    assert e < TOL_SPLIT


def test_split_analysis_conventions_ltr_f32(shtk, oracle, cuda):
    lmax, nlat, nlon, nf = 40, 42, 96, 3
    kw = dict(nlon=nlon, norm="schmidt", cs_phase=False, real_norm=True)
    cfg = shtk.create_gauss_config(lmax, nlat, **kw)
    ocfg = oracle.Config(lmax, nlat, **kw)
    rng = np.random.default_rng(8)
    f = np.asfortranarray(rng.standard_normal((nlat, nlon, nf)))
    ref = ocfg.analysis(f)
    got = to_host(shtk.analysis_batch(cfg, to_dev(f), precision="split"))
    assert max(_per_field(got, ref)) < TOL_SPLIT
    got32 = to_host(shtk.analysis_batch(cfg, to_dev(f.astype(np.float32)), precision="split"))
    assert got32.dtype == np.complex64
    ref32 = ocfg.analysis(f.astype(np.float32).astype(np.float64))
    assert max(_per_field(got32.astype(np.complex128), ref32)) < 1e-6
    cfg2 = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    a = to_host(shtk.analysis_batch(cfg2, to_dev(f), precision="fp64", ltr=25))
    b = to_host(shtk.analysis_batch(cfg2, to_dev(f), precision="split", ltr=25))
    assert rel_l2(b, a) < TOL_SPLIT
    assert np.all(b[26:] == 0)


def test_matrix_core_paths_with_two_pair_blocks(shtk, oracle, cuda):
    """nlat = 1536: 768 latitude pairs, so the analyses contract two 512-pair
    blocks per order and add them into outputs cleared first.  Split (1e-11)
    and FP32-MFMA (2e-6) scalar and vector transforms against the plain FP64
    path per field, the split scalar analysis of one field also against the
    oracle."""
    import torch

    lmax, nlat, nlon, nf = 1535, 1536, 3072, 3
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=21, power=1.5, nf=nf)
    da = to_dev(alm)
    f64 = shtk.synthesis_batch(cfg, da)
    a64 = shtk.analysis_batch(cfg, f64)
    S = oracle.random_alm(lmax, lmax, 22, 2.0, nf=nf, lmin=1)
    T = oracle.random_alm(lmax, lmax, 23, 2.0, nf=nf, lmin=1)
    dS, dT = to_dev(S), to_dev(T)
    Vt64, Vp64 = shtk.synthesis_sphtor_batch(cfg, dS, dT)
    S64, T64 = shtk.analysis_sphtor_batch(cfg, Vt64, Vp64)

    def worst(x, ref):
        x, ref = to_host(x), to_host(ref)
        return max(_per_field(x, ref))

    for prec, tol in (("split", TOL_SPLIT), ("fp32", 2e-6)):
        e_syn = worst(shtk.synthesis_batch(cfg, da, precision=prec), f64)
        e_ana = worst(shtk.analysis_batch(cfg, f64, precision=prec), a64)
        Vt, Vp = shtk.synthesis_sphtor_batch(cfg, dS, dT, precision=prec)
        Sx, Tx = shtk.analysis_sphtor_batch(cfg, Vt64, Vp64, precision=prec)
        e_v = max(worst(Vt, Vt64), worst(Vp, Vp64), worst(Sx, S64), worst(Tx, T64))
        print(f"{prec} lmax={lmax}: synthesis {e_syn:.2e} analysis {e_ana:.2e} sphtor {e_v:.2e}", flush=True)
        assert e_syn < tol and e_ana < tol and e_v < tol
    torch.cuda.synchronize()
    f1 = np.asfortranarray(to_host(f64)[..., 1])
    ref = oracle.Config(lmax, nlat, nlon=nlon).analysis(f1)
    got = to_host(shtk.analysis_batch(cfg, f64, precision="split"))[..., 1]
    assert rel_l2(got, ref) < TOL_SPLIT
    cfg.close()
