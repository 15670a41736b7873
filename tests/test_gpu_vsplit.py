"""GPU parity of the split-precision (Ozaki, int8 matrix core) vector transforms
against the FP64 oracle, through the C ABI.

sin(theta) V_theta and sin(theta) V_phi are computed as scalar split syntheses of
degree lmax + 1 and the analysis as two scalar split analyses with the weights
w/sin(theta) plus an FP64 three-term combination.  Every field must match the
oracle to a relative L2 error of 1e-11 (the target stated for the split path)."""
import numpy as np
import pytest

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL_SPLIT = 1e-11


def _per_field(out, ref):
    return [rel_l2(out[..., k], ref[..., k]) for k in range(ref.shape[-1])]


def _st(oracle, lmax, mmax, seed, nf, mres=1):
    S = oracle.random_alm(lmax, mmax, seed, 2.0, nf=nf, lmin=1)
    T = oracle.random_alm(lmax, mmax, seed + 1, 2.0, nf=nf, lmin=1)
    for m in range(mmax + 1):
        if m % mres:
            S[:, m] = 0
            T[:, m] = 0
    return S, T


@pytest.mark.parametrize("lmax,nlat,nlon,mmax,mres,nf", [
    (3, 8, 10, None, 1, 1),
    (40, 42, 96, None, 1, 3),
    (63, 64, 128, 50, 1, 2),
    (60, 63, 128, 60, 2, 2),
    (200, 202, 432, None, 1, 18),
])
def test_split_sphtor_matches_oracle(shtk, oracle, cuda, lmax, nlat, nlon, mmax, mres, nf):
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    mm = lmax if mmax is None else mmax
    S, T = _st(oracle, lmax, mm, 41, nf, mres)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), precision="split")
    e = _per_field(to_host(Vt), Vt_r) + _per_field(to_host(Vp), Vp_r)
    rng = np.random.default_rng(9)
    gt = np.asfortranarray(rng.standard_normal((nlat, nlon, nf)))
    gp = np.asfortranarray(rng.standard_normal((nlat, nlon, nf)))
    S_r, T_r = ocfg.analysis_sphtor(gt, gp)
    S2, T2 = shtk.analysis_sphtor_batch(cfg, to_dev(gt), to_dev(gp), precision="split")
    a = _per_field(to_host(S2), S_r) + _per_field(to_host(T2), T_r)
    print(f"split sphtor lmax={lmax} nf={nf}: synthesis {max(e):.2e} analysis {max(a):.2e}")
    assert max(e) < TOL_SPLIT and max(a) < TOL_SPLIT


def test_split_sphtor_lmax511(shtk, oracle, cuda):
    """The vector benchmark grid (lmax 511, 512 x 1024, l^-2 spectra): split
    synthesis and analysis against the oracle, and the split round trip."""
    lmax, nlat, nlon, nf = 511, 512, 1024, 3
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    S, T = _st(oracle, lmax, lmax, 4, nf)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), precision="split")
    e = _per_field(to_host(Vt), Vt_r) + _per_field(to_host(Vp), Vp_r)
    S_r, T_r = ocfg.analysis_sphtor(Vt_r, Vp_r)
    S2, T2 = shtk.analysis_sphtor_batch(cfg, to_dev(Vt_r), to_dev(Vp_r), precision="split")
    a = _per_field(to_host(S2), S_r) + _per_field(to_host(T2), T_r)
    S3, T3 = shtk.analysis_sphtor_batch(cfg, Vt, Vp, precision="split")
    r = _per_field(to_host(S3), S) + _per_field(to_host(T3), T)
    print(f"split sphtor lmax=511: synthesis {max(e):.2e} analysis {max(a):.2e} round trip {max(r):.2e}")
    assert max(e) < TOL_SPLIT and max(a) < TOL_SPLIT and max(r) < TOL_SPLIT


@pytest.mark.parametrize("robert", [False, True])
def test_split_sphtor_conventions_ltr_f32(shtk, oracle, cuda, robert):
    """Non-canonical convention (Schmidt, no CS phase, real_norm), Robert form,
    truncation ltr and FP32 I/O through the split vector path."""
    lmax, nlat, nlon, nf = 30, 32, 64, 2
    kw = dict(nlon=nlon, norm="schmidt", cs_phase=False, real_norm=True, robert_form=robert)
    cfg = shtk.create_gauss_config(lmax, nlat, **kw)
    ocfg = oracle.Config(lmax, nlat, **kw)
    S, T = _st(oracle, lmax, lmax, 51, nf)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), precision="split")
    assert max(_per_field(to_host(Vt), Vt_r) + _per_field(to_host(Vp), Vp_r)) < TOL_SPLIT
    S_r, T_r = ocfg.analysis_sphtor(Vt_r, Vp_r)
    S2, T2 = shtk.analysis_sphtor_batch(cfg, to_dev(Vt_r), to_dev(Vp_r), precision="split")
    assert max(_per_field(to_host(S2), S_r) + _per_field(to_host(T2), T_r)) < TOL_SPLIT
    ltr = 17
    Vt_l, Vp_l = ocfg.synthesis_sphtor(S, T, ltr=ltr)
    Vt2, Vp2 = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), precision="split", ltr=ltr)
    assert max(_per_field(to_host(Vt2), Vt_l) + _per_field(to_host(Vp2), Vp_l)) < TOL_SPLIT
    S_l, T_l = ocfg.analysis_sphtor(Vt_r, Vp_r, ltr=ltr)
    S4, T4 = shtk.analysis_sphtor_batch(cfg, to_dev(Vt_r), to_dev(Vp_r), precision="split", ltr=ltr)
    assert max(_per_field(to_host(S4), S_l) + _per_field(to_host(T4), T_l)) < TOL_SPLIT
    # FP32 I/O: the FP32 rounding of inputs and outputs bounds the error
    S32, T32 = S.astype(np.complex64), T.astype(np.complex64)
    Vt32, Vp32 = shtk.synthesis_sphtor_batch(cfg, to_dev(S32), to_dev(T32), precision="split")
    assert max(_per_field(to_host(Vt32), Vt_r) + _per_field(to_host(Vp32), Vp_r)) < 1e-6
    S5, T5 = shtk.analysis_sphtor_batch(cfg, to_dev(Vt_r.astype(np.float32)), to_dev(Vp_r.astype(np.float32)),
                                         precision="split")
    assert max(_per_field(to_host(S5), S_r) + _per_field(to_host(T5), T_r)) < 1e-6


def test_split_sphtor_rejects_bad_precision(shtk, cuda):
    cfg = shtk.create_gauss_config(8, 10, nlon=20)
    S = np.zeros((9, 9, 1), dtype=np.complex128, order="F")
    with pytest.raises(ValueError):
        shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(S), precision="fp16")
