"""GPU parity of the FP32 matrix-core contraction (precision="fp32") against the
FP64 oracle.

The mode keeps FP64 Legendre values, nodes, weights and cross-tile sums and runs
the contraction on v_mfma_f32_16x16x4_f32 (FP32 products, FP32 accumulation over
tiles of 64 products).  Its results are FP32-grade: the tolerance written here is
2e-6 relative L2 per field (FP32 rounding of the two operands is 6e-8 each; the
sums of up to lmax+1 or nlat/2 terms add a few units more).
"""
import numpy as np
import pytest
import torch

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL = 2e-6


def _per_field(got, ref):
    return max(rel_l2(got[..., k], ref[..., k]) for k in range(ref.shape[-1]))


@pytest.mark.parametrize("lmax,nlat,nlon,nf,mmax,mres", [
    (15, 16, 32, 3, None, 1),
    (63, 64, 128, 17, None, 1),      # two field groups
    (127, 130, 256, 4, 100, 1),      # mmax < lmax, nlat > lmax + 1
    (96, 99, 200, 2, 96, 3),         # mres > 1, odd nlat, nlon 2^3 5^2
    (255, 256, 512, 5, None, 1),
    (511, 512, 1024, 2, None, 1),    # two K halves per analysis block
    (1023, 1024, 2048, 2, None, 1),
])
def test_fp32_batch_matches_oracle(shtk, oracle, cuda, lmax, nlat, nlon, nf, mmax, mres):
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    mm = lmax if mmax is None else mmax
    alm = oracle.random_alm(lmax, mm, seed=lmax + nf, power=1.0, nf=nf)
    for m in range(mm + 1):
        if m % mres:
            alm[:, m, :] = 0
    print(f"fp32 lmax={lmax} nlat={nlat} nlon={nlon} nf={nf}", flush=True)
    f = shtk.synthesis_batch(cfg, to_dev(alm), precision="fp32")
    e_syn = _per_field(to_host(f), ocfg.synthesis(alm))
    g = np.asfortranarray(np.random.default_rng(lmax).standard_normal((nlat, nlon, nf)))
    a = shtk.analysis_batch(cfg, to_dev(g), precision="fp32")
    e_ana = _per_field(to_host(a), ocfg.analysis(g))
    back = shtk.analysis_batch(cfg, f, precision="fp32")
    e_rt = _per_field(to_host(back), alm)
    print(f"  synthesis {e_syn:.2e} analysis {e_ana:.2e} round trip {e_rt:.2e}")
    assert e_syn < TOL and e_ana < TOL and e_rt < 2 * TOL


def test_fp32_io_conventions_and_truncation(shtk, oracle, cuda):
    """FP32 arrays, a non-canonical convention and ltr truncation through the same kernels."""
    lmax, nlat, nlon, nf, ltr = 40, 42, 96, 3, 30
    kw = dict(nlon=nlon, norm="schmidt", cs_phase=False, real_norm=True)
    cfg = shtk.create_gauss_config(lmax, nlat, **kw)
    ocfg = oracle.Config(lmax, nlat, **kw)
    alm = oracle.random_alm(lmax, lmax, seed=3, power=1.0, nf=nf)
    f32 = shtk.synthesis_batch(cfg, to_dev(alm.astype(np.complex64)), precision="fp32")
    assert f32.dtype == torch.float32
    assert _per_field(to_host(f32), ocfg.synthesis(alm.astype(np.complex64).astype(np.complex128))) < 5e-6
    g = np.asfortranarray(np.random.default_rng(2).standard_normal((nlat, nlon, nf)))
    a = shtk.analysis_batch(cfg, to_dev(g), precision="fp32", ltr=ltr)
    assert _per_field(to_host(a), ocfg.analysis(g, ltr=ltr)) < TOL
    fl = shtk.synthesis_batch(cfg, to_dev(alm), precision="fp32", ltr=ltr)
    assert _per_field(to_host(fl), ocfg.synthesis(alm, ltr=ltr)) < TOL


def test_fp32_sphtor_batch(shtk, oracle, cuda):
    lmax, nlat, nlon, nf = 127, 128, 256, 3
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    S = oracle.random_alm(lmax, lmax, 4, 2.0, nf=nf, lmin=1)
    T = oracle.random_alm(lmax, lmax, 5, 2.0, nf=nf, lmin=1)
    Vt, Vp = shtk.synthesis_sphtor_batch(cfg, to_dev(S), to_dev(T), precision="fp32")
    Vt_r, Vp_r = ocfg.synthesis_sphtor(S, T)
    assert _per_field(to_host(Vt), Vt_r) < TOL and _per_field(to_host(Vp), Vp_r) < TOL
    S2, T2 = shtk.analysis_sphtor_batch(cfg, to_dev(Vt_r), to_dev(Vp_r), precision="fp32")
    S_r, T_r = ocfg.analysis_sphtor(Vt_r, Vp_r)
    assert _per_field(to_host(S2), S_r) < TOL and _per_field(to_host(T2), T_r) < TOL


def test_fp32_uneven_batch(shtk, oracle, cuda):
    groups = [(255, 3), (127, 2), (63, 4)]
    cfgs = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    ocfgs = [oracle.Config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    alms = [oracle.random_alm(L, L, seed=90 + L, power=1.0, nf=nf) for L, nf in groups]
    b = shtk.UnevenBatch(cfgs, [nf for _, nf in groups], precision="fp32", graph=True)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        d = [to_dev(a) for a in alms]
        for _ in range(2):  # capture, then replay
            fs = b.synthesis(d)
            back = b.analysis(fs)
        torch.cuda.synchronize()
    for oc, a, f, r in zip(ocfgs, alms, fs, back):
        assert _per_field(to_host(f), oc.synthesis(a)) < TOL
        assert _per_field(to_host(r), a) < 2 * TOL
    b.close()
