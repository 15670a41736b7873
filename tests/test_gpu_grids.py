"""GPU parity on the regular grids of create_regular_config (src/config.jl:1235-1330):
Fejer midpoints (grid "regular") and Clenshaw-Curtis nodes with both poles
(grid "regular_poles"), against the FP64 oracle and the SHTns 3.7 fixtures
(C ABI -> HIP kernels).  Tolerance 1e-12 relative L2 for FP64 I/O."""
import numpy as np
import pytest
import torch

import _fixtures as FX
from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL = 1e-12
GRIDS = ["regular", "regular_poles"]


def _pair(shtk, oracle, lmax, nlat, grid, **kw):
    cfg = shtk.create_regular_config(lmax, nlat, include_poles=grid != "regular",
                                     use_dh_weights=grid == "driscoll_healy", **kw)
    return cfg, oracle.Config(lmax, nlat, grid=grid, **kw)


def _alm(oracle, lmax, mmax, mres, seed, nf=None, lmin=0):
    a = oracle.random_alm(lmax, mmax, seed, 2.0, nf=nf, lmin=lmin)
    for m in range(mmax + 1):
        if m % mres:
            a[:, m] = 0
    return a


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("nlat", [9, 10, 64])
def test_nodes_and_weights(shtk, oracle, cuda, grid, nlat):
    cfg, ocfg = _pair(shtk, oracle, 7, nlat, grid)
    assert np.abs(cfg.x - ocfg.x).max() < 5e-16  # two ulps of cos
    assert np.abs(cfg.w - ocfg.w).max() < 1e-15
    if grid == "regular_poles":
        assert cfg.x[0] == 1.0 and cfg.x[-1] == -1.0


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("lmax,nlat,nlon,mmax,mres", [
    (3, 8, 10, None, 1),
    (40, 81, 96, None, 1),
    (40, 42, 96, 30, 1),
    (60, 121, 128, 60, 2),
    (255, 512, 512, None, 1),
])
def test_scalar_match_oracle(shtk, oracle, cuda, grid, lmax, nlat, nlon, mmax, mres):
    if grid == "regular" and nlat < lmax + 2:
        nlat = lmax + 2
    mm = lmax if mmax is None else mmax
    cfg, ocfg = _pair(shtk, oracle, lmax, nlat, grid, nlon=nlon, mmax=mm, mres=mres)
    alm = _alm(oracle, lmax, mm, mres, 5, nf=2)
    f_r = ocfg.synthesis(alm)
    f = shtk.synthesis_batch(cfg, to_dev(alm))
    assert rel_l2(to_host(f), f_r) < TOL
    rng = np.random.default_rng(8)
    g = np.asfortranarray(rng.standard_normal((nlat, nlon, 2)))
    a_r = ocfg.analysis(g)
    a = shtk.analysis_batch(cfg, to_dev(g))
    assert rel_l2(to_host(a), a_r) < TOL
    # split precision on the same grid
    fs = shtk.synthesis_batch(cfg, to_dev(alm), precision="split")
    as_ = shtk.analysis_batch(cfg, to_dev(g), precision="split")
    assert rel_l2(to_host(fs), f_r) < 1e-11 and rel_l2(to_host(as_), a_r) < 1e-11


@pytest.mark.parametrize("grid", GRIDS)
@pytest.mark.parametrize("lmax,nlat,nlon,mmax,mres,robert", [
    (3, 8, 10, None, 1, False),
    (40, 81, 96, None, 1, False),
    (40, 81, 96, None, 1, True),
    (60, 121, 128, 60, 2, False),
    (255, 512, 512, None, 1, False),
])
def test_sphtor_match_oracle(shtk, oracle, cuda, grid, lmax, nlat, nlon, mmax, mres, robert):
    """On the pole grid the order-1 rows at theta = 0, pi take the analytic
    limits of dP/dtheta and P/sin(theta) in both directions."""
    mm = lmax if mmax is None else mmax
    cfg, ocfg = _pair(shtk, oracle, lmax, nlat, grid, nlon=nlon, mmax=mm, mres=mres, robert_form=robert)
    S = _alm(oracle, lmax, mm, mres, 21, nf=2, lmin=1)
    T = _alm(oracle, lmax, mm, mres, 22, nf=2, lmin=1)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T))
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    if grid == "regular_poles":
        # pole rows on their own: the fix-up is what sets them
        for got, ref in ((Vt, Vt_r), (Vp, Vp_r)):
            h = to_host(got)
            for row in (0, nlat - 1):
                assert np.abs(h[row] - ref[row]).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    rng = np.random.default_rng(4)
    gt = np.asfortranarray(rng.standard_normal((nlat, nlon, 2)))
    gp = np.asfortranarray(rng.standard_normal((nlat, nlon, 2)))
    if robert and grid == "regular_poles":
        with pytest.raises(ValueError):
            shtk.analysis_sphtor_batch(cfg, to_dev(gt), to_dev(gp))
        return
    S_r, T_r = ocfg.analysis_sphtor(gt, gp)
    S2, T2 = shtk.analysis_sphtor_batch(cfg, to_dev(gt), to_dev(gp))
    assert rel_l2(to_host(S2), S_r) < TOL and rel_l2(to_host(T2), T_r) < TOL
    # split precision on both grids (pole rows from the limits on the pole grid)
    Vts, Vps = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), precision="split")
    Ss, Ts = shtk.analysis_sphtor_batch(cfg, to_dev(gt), to_dev(gp), precision="split")
    assert rel_l2(to_host(Vts), Vt_r) < 1e-11 and rel_l2(to_host(Vps), Vp_r) < 1e-11
    assert rel_l2(to_host(Ss), S_r) < 1e-11 and rel_l2(to_host(Ts), T_r) < 1e-11


def test_pole_grid_f32_and_truncation(shtk, oracle, cuda):
    """FP32 I/O (the analysis adds the pole terms to float outputs) and ltr < lmax."""
    lmax, nlat, nlon, ltr = 30, 61, 64, 17
    cfg, ocfg = _pair(shtk, oracle, lmax, nlat, "regular_poles", nlon=nlon)
    rng = np.random.default_rng(12)
    gt = np.asfortranarray(rng.standard_normal((nlat, nlon)))
    gp = np.asfortranarray(rng.standard_normal((nlat, nlon)))
    S_r, T_r = ocfg.analysis_sphtor(gt, gp, ltr=ltr)
    S2, T2 = shtk.analysis_sphtor(cfg, to_dev(gt).float(), to_dev(gp).float(), ltr=ltr)
    assert rel_l2(to_host(S2), S_r) < 1e-6 and rel_l2(to_host(T2), T_r) < 1e-6
    S = _alm(oracle, lmax, lmax, 1, 3, lmin=1)
    T = _alm(oracle, lmax, lmax, 1, 4, lmin=1)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T, ltr=ltr)
    Vt, Vp = shtk.synthesis_sphtor(cfg, to_dev(S), to_dev(T), ltr=ltr)
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL


@pytest.mark.parametrize("grid,robert", [("regular_poles", False), ("driscoll_healy", False),
                                         ("driscoll_healy", True), ("regular", True)])
def test_vector_adjoints_pole_grids(shtk, oracle, cuda, grid, robert):
    """Vector adjoints (the rrule pullbacks) on grids with poles: the order-1 pole
    terms move to the other side of the transform (no weight in the adjoint
    synthesis, the quadrature weight in the adjoint analysis)."""
    lmax, nlon = 40, 96
    nlat = 2 * (lmax + 1) if grid == "driscoll_healy" else 81
    cfg, ocfg = _pair(shtk, oracle, lmax, nlat, grid, nlon=nlon, robert_form=robert)
    S = _alm(oracle, lmax, lmax, 1, 31, lmin=1)
    T = _alm(oracle, lmax, lmax, 1, 32, lmin=1)
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


@pytest.mark.parametrize("norm,rn,robert", [("schmidt", True, False), ("orthonormal", False, True)])
def test_pole_grid_split_vector(shtk, oracle, cuda, norm, rn, robert):
    """Split vector transforms on the pole grid: non-canonical conventions (the pole
    terms are added in the plan's convention) and Robert form (zero pole rows;
    the analysis raises with order 1 stored)."""
    lmax, nlat, nlon = 63, 127, 128
    cfg, ocfg = _pair(shtk, oracle, lmax, nlat, "regular_poles", nlon=nlon, norm=norm, real_norm=rn,
                      robert_form=robert)
    S = _alm(oracle, lmax, lmax, 1, 51, nf=2, lmin=1)
    T = _alm(oracle, lmax, lmax, 1, 52, nf=2, lmin=1)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), precision="split")
    assert rel_l2(to_host(Vt), Vt_r) < 1e-11 and rel_l2(to_host(Vp), Vp_r) < 1e-11
    rng = np.random.default_rng(9)
    gt = np.asfortranarray(rng.standard_normal((nlat, nlon, 2)))
    gp = np.asfortranarray(rng.standard_normal((nlat, nlon, 2)))
    if robert:
        with pytest.raises(ValueError):
            shtk.analysis_sphtor_batch(cfg, to_dev(gt), to_dev(gp), precision="split")
        return
    S_r, T_r = ocfg.analysis_sphtor(gt, gp)
    S2, T2 = shtk.analysis_sphtor_batch(cfg, to_dev(gt), to_dev(gp), precision="split")
    assert rel_l2(to_host(S2), S_r) < 1e-11 and rel_l2(to_host(T2), T_r) < 1e-11


@pytest.mark.parametrize("fid", ["packed_storage", "qst_batch", "sphtor_l", "qst_l"])
def test_fixture_regular_grids(shtk, oracle, cuda, fid):
    """SHTns 3.7 syntheses on the regular grids (and qst_l on Gauss nodes) through the HIP path."""
    fx = FX.fixture(fid)
    p = FX.payloads(fx)
    grid = FX.grid_of(fx)
    kw = FX.config_kwargs(fx)
    if grid == "gauss":
        cfg = shtk.create_gauss_config(fx["lmax"], fx["nlat"], **kw)
    else:
        cfg = shtk.create_regular_config(fx["lmax"], fx["nlat"], include_poles=grid == "regular_poles", **kw)
    ltr = fx["ltr"]
    tol = (fx["atol"], fx["rtol"])
    if "coefficients" in p:
        f = shtk.synthesis(cfg, to_dev(FX.unpack(fx, p["coefficients"])), ltr=ltr)
        assert FX.isapprox(to_host(f), p["field"], *tol)[0]
    if "Q" in p:
        Q, S, T = (to_dev(FX.unpack(fx, p[k])) for k in ("Q", "S", "T"))
        if fid == "qst_batch":
            Vr, Vt, Vp = shtk.synthesis_qst_batch(cfg, Q, S, T)
        else:
            Vr, Vt, Vp = shtk.synthesis_qst_l(cfg, Q, S, T, ltr)
        for got, key in ((Vr, "Vr"), (Vt, "Vt"), (Vp, "Vp")):
            assert FX.isapprox(to_host(got), p[key], *tol)[0], key
    elif "S" in p:
        Vt, Vp = shtk.synthesis_sphtor_l(cfg, to_dev(FX.unpack(fx, p["S"])), to_dev(FX.unpack(fx, p["T"])), ltr)
        assert FX.isapprox(to_host(Vt), p["Vt"], *tol)[0]
        assert FX.isapprox(to_host(Vp), p["Vp"], *tol)[0]


def test_fixture_scalar_complex_full(shtk, oracle, cuda):
    """Complex-valued synthesis on the SHTns complex fixture: the m >= 0 bins of
    the packed coefficients plus the conjugate synthesis of the m < 0 ones."""
    from test_oracle import cplx_split

    fx = FX.fixture("scalar_complex_full")
    p = FX.payloads(fx)
    cfg = shtk.create_gauss_config(fx["lmax"], fx["nlat"], **FX.config_kwargs(fx))
    A, B = cplx_split(fx["lmax"], p["coefficients"])
    f = to_host(shtk.synthesis_cplx(cfg, to_dev(A))) + np.conj(to_host(shtk.synthesis_cplx(cfg, to_dev(B))))
    assert FX.isapprox(f, p["field"], fx["atol"], fx["rtol"])[0]


@pytest.mark.parametrize("grid", GRIDS)
def test_round_trip_exact_grids(shtk, oracle, cuda, grid):
    """nlat = 2 lmax + 1 makes both rules exact for band-limited products:
    analysis(synthesis(a)) = a on the GPU, scalar and vector."""
    lmax = 127
    cfg, _ = _pair(shtk, oracle, lmax, 2 * lmax + 1, grid, nlon=256)
    alm = _alm(oracle, lmax, lmax, 1, 9, nf=2)
    back = to_host(shtk.analysis_batch(cfg, shtk.synthesis_batch(cfg, to_dev(alm))))
    assert rel_l2(back, alm) < 1e-12
    S = _alm(oracle, lmax, lmax, 1, 10, nf=2, lmin=1)
    T = _alm(oracle, lmax, lmax, 1, 11, nf=2, lmin=1)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T))
    S2, T2 = shtk.analysis_sphtor_batch(cfg, Vt, Vp)
    assert rel_l2(to_host(S2), S) < 1e-12 and rel_l2(to_host(T2), T) < 1e-12


# --------------------------------------------------------------------------- Driscoll-Healy
DH_CASES = [(3, 10, None, 1, False), (40, 96, None, 1, False), (40, 96, None, 1, True), (60, 128, 60, 2, False),
            (255, 512, None, 1, False)]


@pytest.mark.parametrize("lmax,nlon,mmax,mres,robert", DH_CASES)
def test_driscoll_healy_match_oracle(shtk, oracle, cuda, lmax, nlon, mmax, mres, robert):
    """Driscoll-Healy grid (nlat = 2 (lmax + 1), north pole a node, south pole not):
    the plan runs the symmetric grid of nlat + 1 nodes whose south-pole row does
    not exist in the arrays.  Scalar and vector, FP64 and split scalar, against
    the oracle; Robert-form analysis is allowed (the pole weight is zero)."""
    nlat = 2 * (lmax + 1)
    mm = lmax if mmax is None else mmax
    cfg, ocfg = _pair(shtk, oracle, lmax, nlat, "driscoll_healy", nlon=nlon, mmax=mm, mres=mres,
                      robert_form=robert)
    assert cfg.x.shape == (nlat,) and np.abs(cfg.x - ocfg.x).max() < 5e-16 and np.abs(cfg.w - ocfg.w).max() < 1e-15
    alm = _alm(oracle, lmax, mm, mres, 5, nf=2)
    f_r = ocfg.synthesis(alm)
    assert rel_l2(to_host(shtk.synthesis_batch(cfg, to_dev(alm))), f_r) < TOL
    assert rel_l2(to_host(shtk.synthesis_batch(cfg, to_dev(alm), precision="split")), f_r) < 1e-11
    rng = np.random.default_rng(8)
    g = np.asfortranarray(rng.standard_normal((nlat, nlon, 2)))
    a_r = ocfg.analysis(g)
    assert rel_l2(to_host(shtk.analysis_batch(cfg, to_dev(g))), a_r) < TOL
    assert rel_l2(to_host(shtk.analysis_batch(cfg, to_dev(g), precision="split")), a_r) < 1e-11
    S = _alm(oracle, lmax, mm, mres, 21, nf=2, lmin=1)
    T = _alm(oracle, lmax, mm, mres, 22, nf=2, lmin=1)
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T))
    assert rel_l2(to_host(Vt), Vt_r) < TOL and rel_l2(to_host(Vp), Vp_r) < TOL
    gt = np.asfortranarray(rng.standard_normal((nlat, nlon, 2)))
    gp = np.asfortranarray(rng.standard_normal((nlat, nlon, 2)))
    S_r, T_r = ocfg.analysis_sphtor(gt, gp)
    S2, T2 = shtk.analysis_sphtor_batch(cfg, to_dev(gt), to_dev(gp))
    assert rel_l2(to_host(S2), S_r) < TOL and rel_l2(to_host(T2), T_r) < TOL
    Vts, Vps = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), precision="split")
    assert rel_l2(to_host(Vts), Vt_r) < 1e-11 and rel_l2(to_host(Vps), Vp_r) < 1e-11
    S3, T3 = shtk.analysis_sphtor_batch(cfg, to_dev(gt), to_dev(gp), precision="split")
    assert rel_l2(to_host(S3), S_r) < 1e-11 and rel_l2(to_host(T3), T_r) < 1e-11
    # single order on the same grid
    im = 1 if mres == 1 else 0
    q = S[im * mres:, im * mres, 0]
    col = shtk.synthesis_packed_ml(cfg, im, torch.from_numpy(np.ascontiguousarray(q)).to("cuda:0"), lmax)
    assert rel_l2(col.cpu().numpy(), ocfg.synthesis_ml(im, q, lmax)) < TOL


def test_driscoll_healy_round_trip(shtk, oracle, cuda):
    lmax = 127
    cfg, _ = _pair(shtk, oracle, lmax, 2 * (lmax + 1), "driscoll_healy", nlon=256)
    alm = _alm(oracle, lmax, lmax, 1, 9, nf=2)
    back = to_host(shtk.analysis_batch(cfg, shtk.synthesis_batch(cfg, to_dev(alm))))
    assert rel_l2(back, alm) < 1e-12
    with pytest.raises(ValueError):
        shtk.create_regular_config(lmax, 2 * lmax, include_poles=True, use_dh_weights=True)


def test_packed_wrappers_fixture(shtk, oracle, cuda):
    """``synthesis_packed`` / ``analysis_packed`` (src/transforms.jl:125-157) on the SHTns
    packed_storage fixture: packed coefficients -> flattened grid equals the fixture field,
    and the packed analysis equals the oracle's analysis packed in LM order."""
    fx = FX.fixture("packed_storage")
    p = FX.payloads(fx)
    cfg = shtk.create_regular_config(fx["lmax"], fx["nlat"], include_poles=True, **FX.config_kwargs(fx))
    q = np.asarray(p["coefficients"], dtype=np.complex128)
    vr = shtk.synthesis_packed(cfg, torch.from_numpy(q).to("cuda:0"))
    field = np.asarray(p["field"])
    assert vr.shape == (cfg.nlat * cfg.nlon,)
    assert FX.isapprox(to_host(vr), field.ravel(order="F"), fx["atol"], fx["rtol"])[0]
    ocfg = oracle.Config(fx["lmax"], fx["nlat"], grid="regular_poles", **FX.config_kwargs(fx))
    qa = shtk.analysis_packed(cfg, torch.from_numpy(field.ravel(order="F").copy()).to("cuda:0"))
    ref = FX.pack(fx, ocfg.analysis(np.asfortranarray(field)))
    assert np.abs(to_host(qa) - ref).max() < 1e-12
    with pytest.raises(ValueError):
        shtk.synthesis_packed(cfg, torch.from_numpy(q[:-1]).to("cuda:0"))
