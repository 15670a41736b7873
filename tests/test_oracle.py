"""Pins the CPU oracle before it is trusted as the GPU checker.

Each check is independent of the oracle's own formulas:
* SHTns 3.7 golden fixtures (upstream C library outputs held by the reference);
* 256-bit mpmath Legendre rows (the reference's BigFloat construction);
* Gauss nodes refined in 60-digit arithmetic and numpy's leggauss;
* numpy.fft (pocketfft) and a brute-force DFT;
* closed forms through l = 2, Parseval, adjoint dot-product identities and
  central finite differences.
"""
import math
import os

import numpy as np
import pytest

import _fixtures as FX

HERE = os.path.dirname(os.path.abspath(__file__))


# --------------------------------------------------------------------------- Gauss nodes
@pytest.mark.parametrize("n", [6, 8, 128, 1024])
def test_gauss_nodes_match_high_precision(oracle, n):
    g = np.load(os.path.join(HERE, "golden", "gauss_mpmath.npz"))
    x, w = oracle.gausslegendre(n)
    assert np.max(np.abs(x - g[f"x{n}"])) < 5e-16
    # the reference's weight formula 2/((1-z^2) P_n'^2) loses digits near the
    # poles: relative error up to a few 1e-13 on the smallest weights
    assert np.max(np.abs(w - g[f"w{n}"])) < 5e-16
    assert np.max(np.abs(w - g[f"w{n}"]) / g[f"w{n}"]) < 5e-12
    assert abs(w.sum() - 2.0) < 1e-13
    assert np.all(x[::-1] == -x) and np.all(w[::-1] == w)


def test_gauss_nodes_leggauss_and_exactness(oracle):
    for n in (5, 16, 33):
        x, w = oracle.gausslegendre(n)
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.allclose(x, xr, atol=1e-15) and np.allclose(w, wr, atol=1e-14)
        for d in range(0, 2 * n):  # exact for polynomials of degree <= 2n-1
            exact = 0.0 if d % 2 else 2.0 / (d + 1)
            assert abs(np.sum(w * x ** d) - exact) < 1e-13


# --------------------------------------------------------------------------- Legendre rows
def test_plm_closed_forms(oracle):
    for x in (0.3, -0.77, 0.999):
        s = math.sqrt(1 - x * x)
        P0, P1, P2 = (oracle.plm_row(x, 2, m) for m in range(3))
        assert np.allclose(P0, [math.sqrt(1 / (4 * math.pi)), math.sqrt(3 / (4 * math.pi)) * x,
                                math.sqrt(5 / (16 * math.pi)) * (3 * x * x - 1)], rtol=3e-15, atol=1e-16)
        assert np.allclose(P1[1:], [-math.sqrt(3 / (8 * math.pi)) * s,
                                    -math.sqrt(15 / (8 * math.pi)) * s * x], rtol=3e-15)
        assert abs(P2[2] - math.sqrt(15 / (32 * math.pi)) * s * s) < 3e-15
        assert P1[0] == 0 and P2[0] == 0 and P2[1] == 0


def test_plm_rows_match_mpmath(oracle):
    g = np.load(os.path.join(HERE, "golden", "plm_mpmath.npz"))
    for k in range(int(g["nrows"][0])):
        lmax, m, s, x = g[f"row{k}_meta"]
        lmax, m = int(lmax), int(m)
        ref = g[f"row{k}"]
        P = oracle.plm_row(float(x), lmax, m)
        scale = max(1.0, np.max(np.abs(ref)))
        # 1e-11 is the reference's bound at its own points (k < 4); the extra
        # lmax=1023 rows include the node nearest the pole, where 1023 FP64 steps
        # of the recurrence at x ~ 1 - 3e-6 drift by ~1.1e-11 relative
        tol = 1e-11 if k < 4 else 3e-11
        assert np.max(np.abs(P - ref)) <= tol * scale, (lmax, m, s)


def test_plm_bounded_and_dtheta(oracle):
    g = np.load(os.path.join(HERE, "golden", "plm_mpmath.npz"))
    for m in (0, 3, 200, 512):
        for x in (0.9999, 0.5, 0.01):
            assert np.max(np.abs(oracle.plm_row(x, 512, m))) < 2.0 * math.sqrt(2 * 512 + 1)
    for j in range(4):
        x = float(g[f"zonal{j}_x"][0])
        ref = g[f"zonal{j}"]
        _, dP, _ = oracle.plm_dtheta_over_sinth_row(x, 32, 0)
        if abs(x) < 1:
            assert np.allclose(dP, ref, rtol=5e-13, atol=1e-300)


def test_dtheta_rows_against_finite_differences(oracle):
    """dP/dtheta from the d_{l,m} identity against a centred difference in theta."""
    lmax = 60
    for m in (1, 5, 30):
        for th in (0.3, 1.2, 2.8):
            x = math.cos(th)
            P, dP, Ps = oracle.plm_dtheta_over_sinth_row(x, lmax, m)
            h = 1e-6
            fd = (oracle.plm_row(math.cos(th + h), lmax, m) - oracle.plm_row(math.cos(th - h), lmax, m)) / (2 * h)
            assert np.allclose(dP, fd, atol=1e-7 * max(1, np.max(np.abs(fd))))
            assert np.allclose(Ps, P / math.sin(th), rtol=1e-14, atol=1e-300)


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [8, 10, 12, 14, 22, 101, 202, 256, 608, 2002, 2042, 2048, 1000])
def test_fft_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert np.allclose(oracle.fft(v), np.fft.fft(v), rtol=0, atol=1e-12 * np.sqrt(n) * np.max(np.abs(v)))
    assert np.allclose(oracle.fft(v, inverse=True), np.fft.ifft(v), atol=1e-14 * np.max(np.abs(v)))


def test_fft_brute_force_dft(oracle):
    n = 12
    v = np.arange(n) * 0.3 - 1j * np.arange(n) ** 2 * 0.01
    jk = np.outer(np.arange(n), np.arange(n))
    assert np.allclose(oracle.fft(v), np.exp(-2j * np.pi * jk / n) @ v, atol=1e-13)


# --------------------------------------------------------------------------- SHTns 3.7 fixtures
def _cfg(oracle, fx):
    return oracle.Config(fx["lmax"], fx["nlat"], grid=FX.grid_of(fx), **FX.config_kwargs(fx))


def test_fixture_scalar_real_full(oracle):
    fx = FX.fixture("scalar_real_full_gauss_f64")
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    ok, _ = FX.isapprox(cfg.synthesis(FX.unpack(fx, p["coefficients"])), p["field"], fx["atol"], fx["rtol"])
    assert ok
    ok, _ = FX.isapprox(FX.pack(fx, cfg.analysis(p["field"])), p["coefficients"], fx["atol"], fx["rtol"])
    assert ok


def test_fixture_scalar_batch_real_norm(oracle):
    fx = FX.fixture("scalar_batch")
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    q = p["coefficients"]
    assert FX.isapprox(cfg.synthesis(FX.unpack(fx, q)), p["field"], fx["atol"], fx["rtol"])[0]
    got = cfg.analysis(p["field"])
    for k in range(q.shape[1]):
        assert FX.isapprox(FX.pack(fx, got[:, :, k]), q[:, k], fx["atol"], fx["rtol"])[0]


def test_fixture_scalar_l_conventions(oracle):
    """fourpi normalization, no Condon-Shortley phase, degree truncation ltr=2."""
    fx = FX.fixture("scalar_l")
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    ltr = fx["ltr"]
    q = p["coefficients"]
    assert FX.isapprox(cfg.synthesis(FX.unpack(fx, q), ltr=ltr), p["field"], fx["atol"], fx["rtol"])[0]
    got = FX.pack(fx, cfg.analysis(p["field"], ltr=ltr))
    assert FX.isapprox(got, FX.truncated(fx, q, ltr), fx["atol"], fx["rtol"])[0]


@pytest.mark.parametrize("fid", ["analysis_scalar_full", "analysis_scalar_l"])
def test_fixture_analysis_scalar(oracle, fid):
    """Non-band-limited deterministic inputs: pins the quadrature itself."""
    fx = FX.fixture(fid)
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    ltr = fx["ltr"]
    f = p["field"]
    q = p["coefficients"]
    got = cfg.analysis(f, ltr=ltr)
    if got.ndim == 2:
        got = got[:, :, None]
        q = q[:, None]
    for k in range(q.shape[1]):
        assert FX.isapprox(FX.pack(fx, got[:, :, k]), q[:, k], fx["atol"], fx["rtol"])[0]


def test_fixture_analysis_inputs_pristine():
    """The analysis payloads hold the deterministic generator's bytes
    (test/parity/shtns37_fixtures.jl:261-330)."""
    fx = FX.fixture("analysis_scalar_full")
    f = FX.payloads(fx)["field"]
    nlat, nphi = fx["nlat"], fx["nphi"]
    for k, seed in enumerate((0.2, 0.3)):
        it, ip = np.meshgrid(np.arange(1, nlat + 1), np.arange(1, nphi + 1), indexing="ij")
        i = (it - 1) * nphi + ip
        assert np.allclose(f[:, :, k], seed + 0.013 * i - 0.0007 * i * i, atol=8 * 2.2e-16, rtol=8 * 2.2e-16)


def test_fixture_sphtor_full_f32(oracle):
    fx = FX.fixture("sphtor_full_f32")
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    Vt, Vp = cfg.synthesis_sphtor(FX.unpack(fx, p["S"]), FX.unpack(fx, p["T"]))
    assert FX.isapprox(Vt, p["Vt"], fx["atol"], fx["rtol"])[0]
    assert FX.isapprox(Vp, p["Vp"], fx["atol"], fx["rtol"])[0]
    S, T = cfg.analysis_sphtor(p["Vt"].astype(np.float64), p["Vp"].astype(np.float64))
    assert FX.isapprox(FX.pack(fx, S), p["S"], fx["atol"], fx["rtol"])[0]
    assert FX.isapprox(FX.pack(fx, T), p["T"], fx["atol"], fx["rtol"])[0]


def test_fixture_sphtor_batch_real_norm(oracle):
    fx = FX.fixture("sphtor_batch")
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    Vt, Vp = cfg.synthesis_sphtor(FX.unpack(fx, p["S"]), FX.unpack(fx, p["T"]))
    assert FX.isapprox(Vt, p["Vt"], fx["atol"], fx["rtol"])[0]
    assert FX.isapprox(Vp, p["Vp"], fx["atol"], fx["rtol"])[0]
    S, T = cfg.analysis_sphtor(p["Vt"], p["Vp"])
    for k in range(2):
        assert FX.isapprox(FX.pack(fx, S[:, :, k]), p["S"][:, k], fx["atol"], fx["rtol"])[0]
        assert FX.isapprox(FX.pack(fx, T[:, :, k]), p["T"][:, k], fx["atol"], fx["rtol"])[0]


@pytest.mark.parametrize("fid", ["analysis_sphtor_full", "analysis_sphtor_l"])
def test_fixture_analysis_sphtor(oracle, fid):
    fx = FX.fixture(fid)
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    S, T = cfg.analysis_sphtor(p["Vt"], p["Vp"], ltr=fx["ltr"])
    Sq, Tq = p["S"], p["T"]
    if S.ndim == 2:
        S, T, Sq, Tq = S[:, :, None], T[:, :, None], Sq[:, None], Tq[:, None]
    for k in range(Sq.shape[1]):
        assert FX.isapprox(FX.pack(fx, S[:, :, k]), Sq[:, k], fx["atol"], fx["rtol"])[0]
        assert FX.isapprox(FX.pack(fx, T[:, :, k]), Tq[:, k], fx["atol"], fx["rtol"])[0]


def test_complex_output_against_real_output(oracle):
    """synthesis with real_output = false keeps the bins m >= 0 alone
    (src/core_transforms.jl:294-344), so for real a_l0 the real grid is
    2 Re(complex grid) minus the m = 0 part, which pins the complex path on the
    fixture-pinned real one."""
    lmax, nlat, nlon = 12, 14, 32
    cfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, 3, 1.0)
    alm[:, 0] = alm[:, 0].real
    zc = cfg.synthesis(alm, real_output=False)
    a0 = np.zeros_like(alm)
    a0[:, 0] = alm[:, 0]
    f = cfg.synthesis(alm)
    f0 = cfg.synthesis(a0)
    assert np.allclose(2 * zc.real - f0, f, atol=1e-12, rtol=0)
    # and one bin alone is a pure phase along longitude
    a1 = np.zeros_like(alm)
    a1[3, 2] = 1.0 + 0.5j
    z1 = cfg.synthesis(a1, real_output=False)
    phi = 2 * np.pi * np.arange(nlon) / nlon
    assert np.allclose(z1, z1[:, :1] * np.exp(2j * phi)[None, :], atol=1e-13)


def test_fixture_qst_full(oracle):
    """QST = scalar Q plus sphtor (S, T) on one grid (src/qst_transforms.jl:102-160): the
    oracle's scalar and vector restatements against the SHTns 3.7 QST fixture."""
    fx = FX.fixture("qst_full")
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    Vr = cfg.synthesis(FX.unpack(fx, p["Q"]))
    Vt, Vp = cfg.synthesis_sphtor(FX.unpack(fx, p["S"]), FX.unpack(fx, p["T"]))
    for got, key in ((Vr, "Vr"), (Vt, "Vt"), (Vp, "Vp")):
        assert FX.isapprox(got, p[key], fx["atol"], fx["rtol"])[0], key


@pytest.mark.parametrize("fid", ["analysis_qst_full", "analysis_qst_l"])
def test_fixture_analysis_qst(oracle, fid):
    fx = FX.fixture(fid)
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    Q = cfg.analysis(p["Vr"], ltr=fx["ltr"])
    S, T = cfg.analysis_sphtor(p["Vt"], p["Vp"], ltr=fx["ltr"])
    for got, key in ((Q, "Q"), (S, "S"), (T, "T")):
        ref = p[key]
        if got.ndim == 2:
            got, ref = got[:, :, None], ref[:, None]
        for k in range(ref.shape[1]):
            assert FX.isapprox(FX.pack(fx, got[:, :, k]), ref[:, k], fx["atol"], fx["rtol"])[0], key


# --------------------------------------------------------------------------- identities
def test_closed_form_synthesis(oracle):
    """Y_1^0 and Y_2^1 on the grid (test/parity/scalar_full.jl:48-80)."""
    cfg = oracle.Config(4, 6, nlon=10)
    th = np.arccos(cfg.x)[:, None]
    ph = (2 * np.pi / 10) * np.arange(10)[None, :]
    a = np.zeros((5, 5), complex)
    a[1, 0] = 1.0
    assert np.allclose(cfg.synthesis(a), math.sqrt(3 / (4 * math.pi)) * np.cos(th) + 0 * ph, atol=3e-12)
    a = np.zeros((5, 5), complex)
    a[2, 1] = 0.5 - 0.25j
    ref = 2 * np.real((0.5 - 0.25j) * -math.sqrt(15 / (8 * math.pi)) * np.sin(th) * np.cos(th) * np.exp(1j * ph))
    assert np.allclose(cfg.synthesis(a), ref, atol=3e-12)


def test_parseval_and_round_trip(oracle):
    lmax = 40
    cfg = oracle.Config(lmax, 42, nlon=96)
    a = oracle.random_alm(lmax, lmax, seed=5)
    f = cfg.synthesis(a)
    grid_energy = np.sum(cfg.w[:, None] * f ** 2) * (2 * np.pi / cfg.nlon)
    spec_energy = np.sum(np.abs(a[:, 0]) ** 2) + 2 * np.sum(np.abs(a[:, 1:]) ** 2)
    assert abs(grid_energy - spec_energy) < 1e-12 * spec_energy
    assert oracle.rel_l2(cfg.analysis(f), a) < 1e-13


def test_vector_identities(oracle):
    """Round trip and the divergence identity: the spheroidal part S gives the
    scalar field -l(l+1) S as the surface divergence (src/sphtor_transforms.jl:115-122)."""
    lmax = 30
    cfg = oracle.Config(lmax, 32, nlon=64)
    S = oracle.random_alm(lmax, lmax, 4, 2.0, lmin=1)
    T = oracle.random_alm(lmax, lmax, 5, 2.0, lmin=1)
    Vt, Vp = cfg.synthesis_sphtor(S, T)
    S2, T2 = cfg.analysis_sphtor(Vt, Vp)
    assert oracle.rel_l2(S2, S) < 1e-13 and oracle.rel_l2(T2, T) < 1e-13
    # toroidal-only field is divergence free: its spheroidal analysis vanishes
    Vt, Vp = cfg.synthesis_sphtor(np.zeros_like(S), T)
    S3, _ = cfg.analysis_sphtor(Vt, Vp)
    assert np.max(np.abs(S3)) < 1e-13 * np.max(np.abs(T))


def _inner_real(a, b):
    return float(np.sum(np.real(np.conj(a) * b)))


def test_adjoint_dot_product_identities(oracle):
    """<synthesis(a), y> = Re<a, adj(y)> for the Zygote-scaled adjoint, which
    counts m>0 coefficients twice (they enter the grid with their conjugates)."""
    lmax = 20
    cfg = oracle.Config(lmax, 22, nlon=48)
    rng = np.random.default_rng(3)
    a = oracle.random_alm(lmax, lmax, seed=8)
    y = np.asfortranarray(rng.standard_normal((22, 48)))
    g = cfg.adjoint_synthesis(y)
    lhs = np.sum(cfg.synthesis(a) * y)
    # adjoint returns 2 dL/dconj(a); directional derivative along a is Re<g, a>
    assert abs(lhs - _inner_real(g, a)) < 1e-11 * abs(lhs)
    # analysis adjoint: <analysis(f), b>_R = <f, adj(b)>
    f = np.asfortranarray(rng.standard_normal((22, 48)))
    b = oracle.random_alm(lmax, lmax, seed=9)
    lhs = _inner_real(cfg.analysis(f), b)
    rhs = float(np.sum(f * cfg.adjoint_analysis(b).real))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)
    # vector pair
    S = oracle.random_alm(lmax, lmax, 4, 1.0, lmin=1)
    T = oracle.random_alm(lmax, lmax, 5, 1.0, lmin=1)
    yt = np.asfortranarray(rng.standard_normal((22, 48)))
    yp = np.asfortranarray(rng.standard_normal((22, 48)))
    Vt, Vp = cfg.synthesis_sphtor(S, T)
    Sb, Tb = cfg.adjoint_synthesis_sphtor(yt, yp)
    lhs = np.sum(Vt * yt) + np.sum(Vp * yp)
    assert abs(lhs - _inner_real(Sb, S) - _inner_real(Tb, T)) < 1e-11 * abs(lhs)
    Sa, Ta = cfg.analysis_sphtor(yt, yp)
    vt, vp = cfg.adjoint_analysis_sphtor(S, T)
    lhs = _inner_real(Sa, S) + _inner_real(Ta, T)
    assert abs(lhs - np.sum(yt * vt.real) - np.sum(yp * vp.real)) < 1e-11 * abs(lhs)
    # the whole complex analysis adjoint against the analysis of complex fields:
    # Re<analysis_cplx(g), b> = Re<g, adj(b)> for complex g
    gc = np.asfortranarray(rng.standard_normal((22, 48)) + 1j * rng.standard_normal((22, 48)))
    lhs = _inner_real(cfg.analysis_cplx(gc), b)
    rhs = _inner_real(gc, cfg.adjoint_analysis(b))
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_adjoint_synthesis_finite_difference(oracle):
    """Gradient of L = sum w_i (f - t)^2 against a central difference (h=1e-6,
    rtol 5e-4, test/serial/test_gradients.jl:114-149)."""
    lmax = 12
    cfg = oracle.Config(lmax, 14, nlon=32)
    a = oracle.random_alm(lmax, lmax, seed=6)
    t = cfg.synthesis(oracle.random_alm(lmax, lmax, seed=7))
    wgrid = cfg.w[:, None] * np.ones((1, cfg.nlon))

    def loss(x):
        return float(np.sum(wgrid * (cfg.synthesis(x) - t) ** 2))

    g = cfg.adjoint_synthesis(2 * wgrid * (cfg.synthesis(a) - t))
    rng = np.random.default_rng(1)
    d = oracle.random_alm(lmax, lmax, seed=12) * rng.standard_normal()
    h = 1e-6
    fd = (loss(a + h * d) - loss(a - h * d)) / (2 * h)
    assert abs(fd - _inner_real(g, d)) <= 5e-4 * abs(fd)


# --------------------------------------------------------------------------- regular grids
@pytest.mark.parametrize("n", [2, 3, 5, 6, 9, 10, 17, 32])
def test_regular_rules_exactness(oracle, n):
    """Fejer-1 and Clenshaw-Curtis with n nodes integrate x^k exactly on [-1, 1]
    for k <= n - 1 (and k = n for odd n), nodes at the stated angles, north first."""
    for rule, th in ((oracle.fejer1, (np.arange(n) + 0.5) * np.pi / n),
                     (oracle.clenshaw_curtis, np.arange(n) * np.pi / max(n - 1, 1))):
        if rule is oracle.fejer1 or n >= 2:
            x, w = rule(n)
            assert np.allclose(x, np.cos(th), atol=1e-15, rtol=0)
            assert np.all(np.diff(x) < 0)
            kmax = n - 1 if n % 2 == 0 else n
            for k in range(kmax + 1):
                exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
                assert abs(np.dot(w, x ** k) - exact) < 1e-14, (rule.__name__, n, k)
    x, w = oracle.clenshaw_curtis(n)
    assert x[0] == 1.0 and x[-1] == -1.0


def test_regular_rules_small_cases(oracle):
    """Hand values: Fejer-1 on 2 nodes is the 2-point midpoint rule in theta with
    weights 1, 1; Clenshaw-Curtis on 3 nodes is Simpson's rule (1/3, 4/3, 1/3)."""
    x, w = oracle.fejer1(2)
    assert np.allclose(w, [1.0, 1.0], atol=1e-15)
    x, w = oracle.clenshaw_curtis(3)
    assert np.allclose(w, [1 / 3, 4 / 3, 1 / 3], atol=1e-15) and np.allclose(x, [1.0, 0.0, -1.0], atol=1e-16)
    x, w = oracle.clenshaw_curtis(2)
    assert np.allclose(w, [1.0, 1.0], atol=1e-15)


@pytest.mark.parametrize("fid", ["packed_storage", "qst_batch", "sphtor_l", "qst_l"])
def test_fixture_regular_grid_synthesis(oracle, fid):
    """SHTns 3.7 syntheses on regular grids: Fejer midpoints (sphtor_l, fourpi,
    no Condon-Shortley phase, ltr = 2) and the pole-inclusive grid (packed_storage,
    qst_batch with real_norm and two fields), whose pole rows take the order-1
    limits of dP/dtheta and P/sin(theta)."""
    fx = FX.fixture(fid)
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    ltr = fx["ltr"]
    tol = (fx["atol"], fx["rtol"])
    if "coefficients" in p:
        assert FX.isapprox(cfg.synthesis(FX.unpack(fx, p["coefficients"]), ltr=ltr), p["field"], *tol)[0]
    if "Q" in p:
        assert FX.isapprox(cfg.synthesis(FX.unpack(fx, p["Q"]), ltr=ltr), p["Vr"], *tol)[0]
    if "S" in p:
        Vt, Vp = cfg.synthesis_sphtor(FX.unpack(fx, p["S"]), FX.unpack(fx, p["T"]), ltr=ltr)
        assert FX.isapprox(Vt, p["Vt"], *tol)[0]
        assert FX.isapprox(Vp, p["Vp"], *tol)[0]


def cplx_split(lmax, q):
    """SHTns complex packed coefficients (index l (l + 1) + m, m = -l..l,
    LM_cplx_index src/complex_packed.jl:63-72) -> (A, B) with
    f = synth_cplx(A) + conj(synth_cplx(B)): A holds m >= 0, and since order -m
    uses the row of |m| with exp(-i |m| phi) (src/complex_packed.jl:276-286),
    i.e. Y_l^{-m} = conj(Y_l^m), B_lm = conj(a_{l,-m}) for m > 0."""
    A = np.zeros((lmax + 1, lmax + 1), dtype=np.complex128, order="F")
    B = np.zeros_like(A)
    for l in range(lmax + 1):
        for m in range(-l, l + 1):
            v = q[l * (l + 1) + m]
            if m >= 0:
                A[l, m] = v
            else:
                B[l, -m] = np.conj(v)
    return A, B


def test_fixture_scalar_complex_full(oracle):
    """Complex-valued synthesis (real_output = false, bins m >= 0) pinned on the
    SHTns complex fixture through the split of the full m = -l..l sum."""
    fx = FX.fixture("scalar_complex_full")
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    A, B = cplx_split(fx["lmax"], p["coefficients"])
    f = cfg.synthesis(A, real_output=False) + np.conj(cfg.synthesis(B, real_output=False))
    assert FX.isapprox(f, p["field"], fx["atol"], fx["rtol"])[0]


@pytest.mark.parametrize("grid", ["regular", "regular_poles"])
def test_regular_grid_round_trips(oracle, grid):
    """With nlat = 2 lmax + 1 both rules integrate every product of band-limited
    rows exactly, so analysis inverts synthesis (scalar and vector, the vector
    one through the pole limits on the pole grid)."""
    lmax = 11
    cfg = oracle.Config(lmax, 2 * lmax + 1, nlon=24, grid=grid)
    alm = oracle.random_alm(lmax, lmax, 2, 1.0)
    alm[:, 0] = alm[:, 0].real
    back = cfg.analysis(cfg.synthesis(alm))
    assert np.abs(back - alm).max() < 1e-13
    S = oracle.random_alm(lmax, lmax, 2, 1.0)
    T = oracle.random_alm(lmax, lmax, 2, 1.0)
    S[0] = T[0] = 0
    S[:, 0] = S[:, 0].real
    T[:, 0] = T[:, 0].real
    Vt, Vp = cfg.synthesis_sphtor(S, T)
    S2, T2 = cfg.analysis_sphtor(Vt, Vp)
    assert np.abs(S2 - S).max() < 1e-12 and np.abs(T2 - T).max() < 1e-12


def test_fixture_scalar_complex_analysis(oracle):
    """Analysis of a complex field (the m-loop over the bins m >= 0 of the complex
    FFT) returns the m >= 0 half of the SHTns complex fixture's coefficients (the
    grid integrates the band-limited products exactly), and it is linear:
    analysis_cplx(fr + i fi) = analysis(fr) + i analysis(fi)."""
    fx = FX.fixture("scalar_complex_full")
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    A, _ = cplx_split(fx["lmax"], p["coefficients"])
    a = cfg.analysis_cplx(p["field"])
    assert np.abs(a - A).max() < 1e-12
    rng = np.random.default_rng(3)
    fr = rng.standard_normal((fx["nlat"], fx["nphi"]))
    fi = rng.standard_normal((fx["nlat"], fx["nphi"]))
    assert np.abs(cfg.analysis_cplx(fr + 1j * fi) - (cfg.analysis(fr) + 1j * cfg.analysis(fi))).max() < 1e-14


# --------------------------------------------------------------------------- single order
@pytest.mark.parametrize("fid", ["scalar_ml", "sphtor_ml", "qst_ml", "analysis_scalar_ml", "analysis_sphtor_ml",
                                 "analysis_qst_ml"])
def test_fixture_single_order(oracle, fid):
    """SHTns 3.7 single-order transforms (stored order im, degrees m..ltr), with the
    fixtures' fixed_mode_scale on the Fourier column as the reference's test applies
    it (test/parity/shtns37_fixtures.jl:88-92, 240-247): regular grid with Schmidt,
    real_norm and mres = 2, the pole grid for sphtor_ml, Gauss for the analyses."""
    fx = FX.fixture(fid)
    p = FX.payloads(fx)
    cfg = _cfg(oracle, fx)
    im, ltr, sc = fx["stored_im"], fx["ltr"], fx["fixed_mode_scale"]
    tol = (fx["atol"], fx["rtol"])
    syn = fx.get("direction") != "analysis"
    col = lambda k: np.asarray(p[k]).ravel()
    if "coefficients" in p or "Q" in p:
        q, v = (col("coefficients"), col("field")) if "coefficients" in p else (col("Q"), col("Vr"))
        if syn:
            assert FX.isapprox(cfg.synthesis_ml(im, q, ltr), sc * v, *tol)[0]
        assert FX.isapprox(cfg.analysis_ml(im, sc * v, ltr), q, *tol)[0]
    if "S" in p:
        if syn:
            Vt, Vp = cfg.synthesis_sphtor_ml(im, col("S"), col("T"), ltr)
            assert FX.isapprox(Vt, sc * col("Vt"), *tol)[0] and FX.isapprox(Vp, sc * col("Vp"), *tol)[0]
        S, T = cfg.analysis_sphtor_ml(im, sc * col("Vt"), sc * col("Vp"), ltr)
        assert FX.isapprox(S, col("S"), *tol)[0] and FX.isapprox(T, col("T"), *tol)[0]


def test_single_order_against_full_transform(oracle):
    """With only order m present the complex grid of the full synthesis is
    F_m(theta) exp(i m phi), and the single-order column is phi_inv_scale (nlon)
    times F_m, i.e. nlon times the grid at phi = 0."""
    lmax, nlat, nlon = 10, 12, 24
    cfg = oracle.Config(lmax, nlat, nlon=nlon, norm="schmidt", real_norm=True)
    for m in (0, 3):
        a = np.zeros((lmax + 1, lmax + 1), dtype=np.complex128, order="F")
        rng = np.random.default_rng(m)
        a[m:, m] = rng.standard_normal(lmax + 1 - m) + 1j * rng.standard_normal(lmax + 1 - m)
        z = cfg.synthesis(a, real_output=False)
        col = cfg.synthesis_ml(m, a[m:, m], lmax)
        assert np.allclose(col / nlon, z[:, 0], atol=1e-13, rtol=0)


@pytest.mark.parametrize("n", [4, 8, 16, 32, 64])
def test_driscoll_healy_rule(oracle, n):
    """Driscoll-Healy weights on theta_j = pi j / n integrate x^k exactly for
    k <= n - 1 and not for k = n; the north pole is a node with weight 0 and the
    south pole is not a node."""
    x, w = oracle.driscoll_healy(n)
    assert x[0] == 1.0 and w[0] == 0.0 and x[-1] > -1.0
    assert np.allclose(x, np.cos(np.pi * np.arange(n) / n), atol=1e-16)
    for k in range(n):
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        assert abs(np.dot(w, x ** k) - exact) < 2e-14
    if n <= 16:  # the degree-n error shrinks fast with n
        assert abs(np.dot(w, x ** n) - 2.0 / (n + 1)) > 1e-12


def test_driscoll_healy_round_trip(oracle):
    """nlat = 2 (lmax + 1) Driscoll-Healy nodes: analysis inverts synthesis, scalar
    and vector (Robert form included: the pole weight is zero)."""
    lmax = 11
    for robert in (False, True):
        cfg = oracle.Config(lmax, 2 * (lmax + 1), nlon=24, grid="driscoll_healy", robert_form=robert)
        alm = oracle.random_alm(lmax, lmax, 2, 1.0)
        alm[:, 0] = alm[:, 0].real
        assert np.abs(cfg.analysis(cfg.synthesis(alm)) - alm).max() < 1e-13
        S = oracle.random_alm(lmax, lmax, 4, 1.0, lmin=1)
        T = oracle.random_alm(lmax, lmax, 5, 1.0, lmin=1)
        S[:, 0] = S[:, 0].real
        T[:, 0] = T[:, 0].real
        S2, T2 = cfg.analysis_sphtor(*cfg.synthesis_sphtor(S, T))
        assert np.abs(S2 - S).max() < 1e-12 and np.abs(T2 - T).max() < 1e-12


@pytest.mark.parametrize("grid", ["regular_poles", "driscoll_healy"])
def test_adjoint_dot_product_pole_grids(oracle, grid):
    """The vector adjoints on grids with nodes at the poles (order-1 pole limits)
    satisfy the same dot-product identities as on Gauss grids."""
    lmax = 11
    nlat = 2 * (lmax + 1) if grid == "driscoll_healy" else 23
    cfg = oracle.Config(lmax, nlat, nlon=24, grid=grid)
    rng = np.random.default_rng(3)
    S = oracle.random_alm(lmax, lmax, 4, 1.0, lmin=1)
    T = oracle.random_alm(lmax, lmax, 5, 1.0, lmin=1)
    yt = np.asfortranarray(rng.standard_normal((nlat, 24)))
    yp = np.asfortranarray(rng.standard_normal((nlat, 24)))
    Vt, Vp = cfg.synthesis_sphtor(S, T)
    Sb, Tb = cfg.adjoint_synthesis_sphtor(yt, yp)
    lhs = np.sum(Vt * yt) + np.sum(Vp * yp)
    assert abs(lhs - _inner_real(Sb, S) - _inner_real(Tb, T)) < 1e-11 * abs(lhs)
    Sa, Ta = cfg.analysis_sphtor(yt, yp)
    vt, vp = cfg.adjoint_analysis_sphtor(S, T)
    lhs = _inner_real(Sa, S) + _inner_real(Ta, T)
    assert abs(lhs - np.sum(yt * vt.real) - np.sum(yp * vp.real)) < 1e-11 * abs(lhs)
