"""Gradients through the reverse-mode rules (the ChainRulesCore rrules of
ext/SHTnsKitAdvancedADExt.jl:55-241, here torch.autograd functions whose
pullbacks call the hand-written adjoint kernels).

The differentiation workload: 16 fields, lmax 511, 512 x 1024 Gauss grid, FP32
coefficients (complex Gaussian, 1/(l+1), seed 6), target grid from a seeded
random spectrum (seed 7), loss = sum_k sum_ij w_i (synthesis(alm_k) - t_k)^2.
The gradient is checked against (a) the oracle's FP64 adjoint applied to the
loss cotangent -- the reference's Zygote gradient is exactly
_adjoint_synthesis of 2 w (f - t), the composition the rrule performs -- and
(b) central finite differences along random directions (FP64 forward passes,
as test/serial/test_gradients.jl:114-183)."""
import numpy as np
import pytest
import torch

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu


def _setup(shtk, oracle, lmax=511, nlat=512, nlon=1024, nf=16):
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=6, power=1.0, nf=nf)
    t = ocfg.synthesis(oracle.random_alm(lmax, lmax, seed=7, power=1.0, nf=nf))
    return cfg, ocfg, alm, t


def _loss(shtk, cfg, a, t_dev, w_dev):
    f = shtk.ad.synthesis(cfg, a)
    return (w_dev[:, None, None] * (f - t_dev) ** 2).sum()


def test_gradient_fp32_batch_vs_adjoint_oracle(shtk, oracle, cuda):
    cfg, ocfg, alm, t = _setup(shtk, oracle)
    w = np.asarray(ocfg.w)
    a = to_dev(alm.astype(np.complex64)).requires_grad_(True)
    t_dev = to_dev(t.astype(np.float32))
    w_dev = torch.as_tensor(w, dtype=torch.float32, device="cuda:0")
    L = _loss(shtk, cfg, a, t_dev, w_dev)
    L.backward()
    g = to_host(a.grad)
    assert g.dtype == np.complex64
    # FP64 reference gradient: adjoint of synthesis applied to dL/df = 2 w (f - t)
    a64 = alm.astype(np.complex64).astype(np.complex128)
    f64 = ocfg.synthesis(a64)
    fbar = 2.0 * w[:, None, None] * (f64 - t)
    gref = ocfg.adjoint_synthesis(fbar)
    err = rel_l2(g.astype(np.complex128), gref)
    print(f"gradient FP32 (16 fields, lmax 511) vs FP64 adjoint oracle: rel L2 {err:.2e}")
    assert err < 2e-5
    # m = 0 imaginary parts do not enter a real synthesis: zero gradient there
    assert np.all(g[:, 0, :].imag == 0)


def test_gradient_central_differences(shtk, oracle, cuda):
    """Directional derivatives of the FP64 loss against the adjoint gradient."""
    cfg, ocfg, alm, t = _setup(shtk, oracle, lmax=127, nlat=128, nlon=256, nf=4)
    w = torch.as_tensor(np.asarray(ocfg.w), dtype=torch.float64, device="cuda:0")
    t_dev = to_dev(t)
    a = to_dev(alm).requires_grad_(True)
    _loss(shtk, cfg, a, t_dev, w).backward()
    g = a.grad.detach()
    rng = np.random.default_rng(11)
    for k in range(3):
        d = oracle.random_alm(127, 127, seed=100 + k, power=1.0, nf=4)
        dd = to_dev(d)
        h = 1e-6
        with torch.no_grad():
            lp = float(_loss(shtk, cfg, a.detach() + h * dd, t_dev, w))
            lm = float(_loss(shtk, cfg, a.detach() - h * dd, t_dev, w))
        fd = (lp - lm) / (2 * h)
        an = float((g.real * dd.real + g.imag * dd.imag).sum())
        print(f"direction {k}: finite difference {fd:.10e} adjoint {an:.10e}")
        assert abs(fd - an) <= 5e-4 * max(abs(fd), abs(an))


def test_analysis_gradient_matches_adjoint(shtk, oracle, cuda):
    """d/df of Re<analysis(f), c> is the real part of the analysis adjoint."""
    lmax, nlat, nlon = 63, 64, 128
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    rng = np.random.default_rng(3)
    f = to_dev(np.asfortranarray(rng.standard_normal((nlat, nlon)))).requires_grad_(True)
    c = oracle.random_alm(lmax, lmax, seed=4, power=0.5)
    cd = to_dev(c)
    out = shtk.ad.analysis(cfg, f)
    (out.real * cd.real + out.imag * cd.imag).sum().backward()
    ref = ocfg.adjoint_analysis(c).real
    assert rel_l2(to_host(f.grad), ref) < 1e-12
