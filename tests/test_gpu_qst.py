"""GPU parity of the 3-D (QST) vector transforms and the degree-truncated vector
variants against the SHTns 3.7 fixtures held by the reference and the FP64 oracle.
QST is the scalar transform of Q plus the sphtor transform of (S, T) on the same
grid (src/qst_transforms.jl:102-160, device form :206-230)."""
import numpy as np
import pytest

import _fixtures as FX
from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_SPLIT = 1e-11


def _cfgs(shtk, oracle, fx):
    kw = FX.config_kwargs(fx)
    return shtk.create_gauss_config(fx["lmax"], fx["nlat"], **kw), oracle.Config(fx["lmax"], fx["nlat"], **kw)


def test_qst_fixture_synthesis(shtk, oracle, cuda):
    fx = FX.fixture("qst_full")
    p = FX.payloads(fx)
    cfg, _ = _cfgs(shtk, oracle, fx)
    Vr, Vt, Vp = shtk.synthesis_qst(cfg, to_dev(FX.unpack(fx, p["Q"])), to_dev(FX.unpack(fx, p["S"])),
                                    to_dev(FX.unpack(fx, p["T"])))
    for got, key in ((Vr, "Vr"), (Vt, "Vt"), (Vp, "Vp")):
        assert FX.isapprox(to_host(got), p[key], fx["atol"], fx["rtol"])[0], key


@pytest.mark.parametrize("fid", ["analysis_qst_full", "analysis_qst_l"])
def test_qst_fixture_analysis(shtk, oracle, cuda, fid):
    fx = FX.fixture(fid)
    p = FX.payloads(fx)
    cfg, _ = _cfgs(shtk, oracle, fx)
    if p["Vr"].ndim == 3:
        out = shtk.analysis_qst_batch(cfg, to_dev(p["Vr"]), to_dev(p["Vt"]), to_dev(p["Vp"]), ltr=fx["ltr"])
    else:
        out = shtk.analysis_qst_l(cfg, to_dev(p["Vr"]), to_dev(p["Vt"]), to_dev(p["Vp"]), fx["ltr"])
    for got, key in zip(out, ("Q", "S", "T")):
        got, ref = to_host(got), p[key]
        if got.ndim == 2:
            got, ref = got[:, :, None], ref[:, None]
        for k in range(ref.shape[1]):
            assert FX.isapprox(FX.pack(fx, got[:, :, k]), ref[:, k], fx["atol"], fx["rtol"])[0], key


@pytest.mark.parametrize("precision,tol", [("fp64", TOL), ("split", TOL_SPLIT)])
def test_qst_batch_matches_oracle(shtk, oracle, cuda, precision, tol):
    lmax, nlat, nlon, nf = 63, 64, 128, 3
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    Q = oracle.random_alm(lmax, lmax, 61, 1.0, nf=nf)
    S = oracle.random_alm(lmax, lmax, 62, 2.0, nf=nf, lmin=1)
    T = oracle.random_alm(lmax, lmax, 63, 2.0, nf=nf, lmin=1)
    Vr, Vt, Vp = shtk.synthesis_qst_batch(cfg, to_dev(Q), to_dev(S), to_dev(T), precision=precision)
    Vr_r = ocfg.synthesis(Q)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    for got, ref in ((Vr, Vr_r), (Vt, Vt_r), (Vp, Vp_r)):
        g = to_host(got)
        assert max(rel_l2(g[..., k], ref[..., k]) for k in range(nf)) < tol
    Q2, S2, T2 = shtk.analysis_qst_batch(cfg, Vr, Vt, Vp, precision=precision)
    for got, ref in ((Q2, Q), (S2, S), (T2, T)):
        g = to_host(got)
        assert max(rel_l2(g[..., k], ref[..., k]) for k in range(nf)) < tol


def test_qst_promotes_to_the_widest_input(shtk, oracle, cuda):
    lmax, nlat = 15, 16
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=32)
    Q = oracle.random_alm(lmax, lmax, 71, 1.0).astype(np.complex64)
    S = oracle.random_alm(lmax, lmax, 72, 2.0, lmin=1)
    T = oracle.random_alm(lmax, lmax, 73, 2.0, lmin=1)
    Vr, Vt, Vp = shtk.synthesis_qst(cfg, to_dev(Q), to_dev(S), to_dev(T))
    assert Vr.dtype == Vt.dtype == Vp.dtype == __import__("torch").float64


def test_sphtor_truncated_variants(shtk, oracle, cuda):
    lmax, nlat, nlon, ltr = 40, 42, 96, 25
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    S = oracle.random_alm(lmax, lmax, 81, 2.0, lmin=1)
    T = oracle.random_alm(lmax, lmax, 82, 2.0, lmin=1)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T, ltr=ltr)
    Vt, Vp = shtk.synthesis_sphtor_l(cfg, to_dev(S), to_dev(T), ltr)
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    S_r, T_r = ocfg.analysis_sphtor(Vt_r, Vp_r, ltr=ltr)
    S2, T2 = shtk.analysis_sphtor_l(cfg, to_dev(Vt_r), to_dev(Vp_r), ltr)
    assert rel_l2(to_host(S2), S_r) < TOL and rel_l2(to_host(T2), T_r) < TOL
    Z = np.zeros_like(S)
    sVt, sVp = ocfg.synthesis_sphtor(S, Z, ltr=ltr)
    a, b = shtk.synthesis_sph_l(cfg, to_dev(S), ltr)
    assert rel_l2(to_host(a), sVt) < TOL and rel_l2(to_host(b), sVp) < TOL
    tVt, tVp = ocfg.synthesis_sphtor(Z, T, ltr=ltr)
    a, b = shtk.synthesis_tor_l(cfg, to_dev(T), ltr)
    assert rel_l2(to_host(a), tVt) < TOL and rel_l2(to_host(b), tVp) < TOL
