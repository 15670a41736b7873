"""GPU parity of the single-order transforms (synthesis_packed_ml, analysis_packed_ml,
synthesis_sphtor_ml, analysis_sphtor_ml and their compositions) against the FP64
oracle and the SHTns 3.7 single-order fixtures (C ABI -> HIP kernels).
Tolerance 1e-12 relative L2 for FP64 I/O, 1e-6 for FP32."""
import numpy as np
import pytest
import torch

import _fixtures as FX
from _util import rel_l2

pytestmark = pytest.mark.gpu


def _dev(a):
    a = np.asarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a.T)).to("cuda:0")
    return t.t() if a.ndim == 2 else t


def _host(t):
    return t.detach().cpu().numpy()


def _cfgs(shtk, oracle, lmax, nlat, grid, **kw):
    if grid == "gauss":
        cfg = shtk.create_gauss_config(lmax, nlat, **kw)
    else:
        cfg = shtk.create_regular_config(lmax, nlat, include_poles=grid == "regular_poles", **kw)
    return cfg, oracle.Config(lmax, nlat, grid=grid, **kw)


@pytest.mark.parametrize("fid", ["scalar_ml", "sphtor_ml", "qst_ml", "analysis_scalar_ml", "analysis_sphtor_ml",
                                 "analysis_qst_ml"])
def test_fixture_single_order(shtk, oracle, cuda, fid):
    fx = FX.fixture(fid)
    p = FX.payloads(fx)
    grid = FX.grid_of(fx)
    cfg, _ = _cfgs(shtk, oracle, fx["lmax"], fx["nlat"], grid, **FX.config_kwargs(fx))
    im, ltr, sc = fx["stored_im"], fx["ltr"], fx["fixed_mode_scale"]
    tol = (fx["atol"], fx["rtol"])
    syn = fx.get("direction") != "analysis"
    col = lambda k: np.asarray(p[k]).ravel()
    if "Q" in p:
        if syn:
            Vr, Vt, Vp = shtk.synthesis_qst_ml(cfg, im, _dev(col("Q")), _dev(col("S")), _dev(col("T")), ltr)
            for got, k in ((Vr, "Vr"), (Vt, "Vt"), (Vp, "Vp")):
                assert FX.isapprox(_host(got), sc * col(k), *tol)[0], k
        Q, S, T = shtk.analysis_qst_ml(cfg, im, *(_dev(sc * col(k)) for k in ("Vr", "Vt", "Vp")), ltr)
        for got, k in ((Q, "Q"), (S, "S"), (T, "T")):
            assert FX.isapprox(_host(got), col(k), *tol)[0], k
        return
    if "coefficients" in p:
        if syn:
            v = shtk.synthesis_packed_ml(cfg, im, _dev(col("coefficients")), ltr)
            assert FX.isapprox(_host(v), sc * col("field"), *tol)[0]
        q = shtk.analysis_packed_ml(cfg, im, _dev(sc * col("field")), ltr)
        assert FX.isapprox(_host(q), col("coefficients"), *tol)[0]
    if "S" in p:
        if syn:
            Vt, Vp = shtk.synthesis_sphtor_ml(cfg, im, _dev(col("S")), _dev(col("T")), ltr)
            assert FX.isapprox(_host(Vt), sc * col("Vt"), *tol)[0] and FX.isapprox(_host(Vp), sc * col("Vp"), *tol)[0]
        S, T = shtk.analysis_sphtor_ml(cfg, im, _dev(sc * col("Vt")), _dev(sc * col("Vp")), ltr)
        assert FX.isapprox(_host(S), col("S"), *tol)[0] and FX.isapprox(_host(T), col("T"), *tol)[0]


CASES = [
    # grid, lmax, nlat, kw, im, ltr
    ("gauss", 40, 42, {}, 0, 40),
    ("gauss", 40, 42, {}, 1, 33),
    ("gauss", 40, 42, {"norm": "schmidt", "real_norm": True}, 7, 40),
    ("regular", 40, 81, {"mres": 2, "mmax": 40}, 3, 40),
    ("regular_poles", 40, 81, {}, 1, 40),
    ("regular_poles", 40, 81, {"robert_form": True}, 1, 40),
    ("regular_poles", 40, 81, {"norm": "fourpi", "cs_phase": False}, 0, 29),
    ("gauss", 40, 42, {"robert_form": True}, 5, 40),
    ("gauss", 1023, 1024, {}, 0, 1023),
    ("gauss", 1023, 1024, {}, 600, 1023),
    ("regular_poles", 511, 1023, {}, 1, 511),
]


@pytest.mark.parametrize("grid,lmax,nlat,kw,im,ltr", CASES)
def test_single_order_match_oracle(shtk, oracle, cuda, grid, lmax, nlat, kw, im, ltr):
    cfg, ocfg = _cfgs(shtk, oracle, lmax, nlat, grid, **kw)
    m = im * cfg.mres
    rng = np.random.default_rng(im + ltr)
    nl = ltr - m + 1
    q = (rng.standard_normal(nl) + 1j * rng.standard_normal(nl)) / (1.0 + np.arange(m, ltr + 1)) ** 1.5
    t = (rng.standard_normal(nl) + 1j * rng.standard_normal(nl)) / (1.0 + np.arange(m, ltr + 1)) ** 1.5
    v = rng.standard_normal(nlat) + 1j * rng.standard_normal(nlat)
    u = rng.standard_normal(nlat) + 1j * rng.standard_normal(nlat)
    assert rel_l2(_host(shtk.synthesis_packed_ml(cfg, im, _dev(q), ltr)), ocfg.synthesis_ml(im, q, ltr)) < 1e-12
    assert rel_l2(_host(shtk.analysis_packed_ml(cfg, im, _dev(v), ltr)), ocfg.analysis_ml(im, v, ltr)) < 1e-12
    Vt, Vp = shtk.synthesis_sphtor_ml(cfg, im, _dev(q), _dev(t), ltr)
    Vt_r, Vp_r = ocfg.synthesis_sphtor_ml(im, q, t, ltr)
    assert rel_l2(_host(Vt), Vt_r) < 1e-12 and rel_l2(_host(Vp), Vp_r) < 1e-12
    if kw.get("robert_form") and grid == "regular_poles" and m == 1:
        with pytest.raises(ValueError):
            shtk.analysis_sphtor_ml(cfg, im, _dev(v), _dev(u), ltr)
        return
    S, T = shtk.analysis_sphtor_ml(cfg, im, _dev(v), _dev(u), ltr)
    S_r, T_r = ocfg.analysis_sphtor_ml(im, v, u, ltr)
    assert rel_l2(_host(S), S_r) < 1e-12 and rel_l2(_host(T), T_r) < 1e-12


def test_single_order_f32_and_batch(shtk, oracle, cuda):
    """FP32 columns, and a batch of columns (n, nf) matching the one-column calls."""
    lmax, nlat, im, ltr = 60, 62, 4, 55
    cfg, ocfg = _cfgs(shtk, oracle, lmax, nlat, "gauss")
    rng = np.random.default_rng(3)
    nl = ltr - im + 1
    Q = rng.standard_normal((nl, 3)) + 1j * rng.standard_normal((nl, 3))
    V = shtk.synthesis_packed_ml(cfg, im, _dev(Q), ltr)
    for k in range(3):
        assert rel_l2(_host(V)[:, k], ocfg.synthesis_ml(im, Q[:, k], ltr)) < 1e-12
    V32 = shtk.synthesis_packed_ml(cfg, im, _dev(Q[:, 0]).to(torch.complex64), ltr)
    assert V32.dtype == torch.complex64
    assert rel_l2(_host(V32), ocfg.synthesis_ml(im, Q[:, 0], ltr)) < 1e-6
    Vt = rng.standard_normal((nlat, 2)) + 1j * rng.standard_normal((nlat, 2))
    Vp = rng.standard_normal((nlat, 2)) + 1j * rng.standard_normal((nlat, 2))
    S, T = shtk.analysis_sphtor_ml(cfg, im, _dev(Vt), _dev(Vp), ltr)
    for k in range(2):
        S_r, T_r = ocfg.analysis_sphtor_ml(im, Vt[:, k], Vp[:, k], ltr)
        assert rel_l2(_host(S)[:, k], S_r) < 1e-12 and rel_l2(_host(T)[:, k], T_r) < 1e-12


def test_single_order_argument_checks(shtk, oracle, cuda):
    cfg, _ = _cfgs(shtk, oracle, 8, 10, "gauss", mres=2, mmax=8)
    q = _dev(np.ones(5, dtype=np.complex128))
    with pytest.raises(ValueError):
        shtk.synthesis_packed_ml(cfg, 5, q, 8)  # im > mmax/mres
    with pytest.raises(ValueError):
        shtk.synthesis_packed_ml(cfg, 3, q, 5)  # ltr < m
    with pytest.raises(ValueError):
        shtk.synthesis_packed_ml(cfg, 1, q, 8)  # wrong column length (needs 7: 8 - 2 + 1)
