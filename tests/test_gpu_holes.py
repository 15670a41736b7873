"""The matrix-core analyses store every output coefficient themselves.

With one pair block (nlat <= 1024) the split and FP32 analyses do not clear
their outputs first: each work item also stores the zeros of the coefficients
it owns but does not contract (l < m, l > ltr, orders between multiples of
mres, degree segments where no latitude is live).  These tests hand the C ABI
output arrays filled with NaN, so any coefficient left unwritten fails the
comparison with the oracle (split: 1e-11 relative L2 per field; FP32: 2e-6),
and check that the zeros are exact."""
import ctypes

import numpy as np
import pytest
import torch

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL = {"split": 1e-11, "fp32": 2e-6}


def _nan_spec(cfg, nf):
    # column-major (lmax+1, mmax+1, nf) complex128, every entry NaN
    t = torch.full((nf, cfg.mmax + 1, cfg.lmax + 1), complex(float("nan"), float("nan")),
                   dtype=torch.complex128, device="cuda:0")
    return t.permute(2, 1, 0)


def _per_field(got, ref):
    return max(rel_l2(got[..., k], ref[..., k]) for k in range(ref.shape[-1]))


@pytest.mark.parametrize("prec", ["split", "fp32"])
@pytest.mark.parametrize("lmax,nlat,nlon,mmax,mres,ltr,nf", [
    (60, 63, 128, None, 2, 41, 3),       # odd orders, truncation
    (255, 256, 512, 200, 1, -1, 2),      # high orders whose leading segments are dead
    (511, 512, 1024, None, 1, 300, 17),  # two K halves per block, two field groups
    (96, 99, 200, 96, 3, 70, 1),
])
def test_scalar_analysis_writes_every_coefficient(shtk, oracle, cuda, prec, lmax, nlat, nlon, mmax, mres, ltr, nf):
    from shtnskit_jl_amd import _lib
    from shtnskit_jl_amd.transforms import _PRECISIONS, _ptr, _stream

    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon, mmax=mmax, mres=mres)
    f = np.asfortranarray(np.random.default_rng(lmax + nf).standard_normal((nlat, nlon, nf)))
    fd = to_dev(f)
    out = _nan_spec(cfg, nf)
    print(f"{prec} lmax={lmax} nlat={nlat} mres={mres} ltr={ltr} nf={nf}", flush=True)
    _lib.check(_lib.load().shtk_analysis_ex(ctypes.c_void_p(cfg.plan), _ptr(fd), _ptr(out), nf, _lib.SHTK_F64,
                                            ltr, _PRECISIONS[prec], _stream(fd.device)), prec)
    got = to_host(out)
    assert not np.isnan(got).any()
    ref = ocfg.analysis(f, ltr=ltr)
    assert _per_field(got, ref) < TOL[prec]
    lt = lmax if ltr < 0 else ltr
    assert np.all(got[lt + 1:] == 0)
    mm = cfg.mmax
    for m in range(mm + 1):
        assert np.all(got[:m, m] == 0)
        if m % mres:
            assert np.all(got[:, m] == 0)


@pytest.mark.parametrize("prec", ["split", "fp32"])
def test_vector_analysis_writes_every_coefficient(shtk, oracle, cuda, prec):
    from shtnskit_jl_amd import _lib
    from shtnskit_jl_amd.transforms import _PRECISIONS, _ptr, _stream

    lmax, nlat, nlon, nf, ltr = 127, 128, 256, 3, 90
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    rng = np.random.default_rng(4)
    Vt = np.asfortranarray(rng.standard_normal((nlat, nlon, nf)))
    Vp = np.asfortranarray(rng.standard_normal((nlat, nlon, nf)))
    dt, dp = to_dev(Vt), to_dev(Vp)
    S, T = _nan_spec(cfg, nf), _nan_spec(cfg, nf)
    _lib.check(_lib.load().shtk_analysis_sphtor_ex(ctypes.c_void_p(cfg.plan), _ptr(dt), _ptr(dp), _ptr(S), _ptr(T),
                                                   nf, _lib.SHTK_F64, ltr, _PRECISIONS[prec],
                                                   _stream(dt.device)), prec)
    Sg, Tg = to_host(S), to_host(T)
    assert not np.isnan(Sg).any() and not np.isnan(Tg).any()
    Sr, Tr = ocfg.analysis_sphtor(Vt, Vp, ltr=ltr)
    assert _per_field(Sg, Sr) < TOL[prec] and _per_field(Tg, Tr) < TOL[prec]
    assert np.all(Sg[ltr + 1:] == 0) and np.all(Tg[ltr + 1:] == 0)


def test_uneven_batch_analysis_writes_every_coefficient(shtk, oracle, cuda):
    from shtnskit_jl_amd import _lib
    from shtnskit_jl_amd.transforms import _stream

    groups = [(255, 3), (63, 2), (511, 1)]
    cfgs = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    b = shtk.UnevenBatch(cfgs, [nf for _, nf in groups], precision="split")
    fs = [np.asfortranarray(np.random.default_rng(L).standard_normal((L + 1, 2 * L + 2, nf))) for L, nf in groups]
    ds = [to_dev(f) for f in fs]
    outs = [_nan_spec(c, nf) for c, (_, nf) in zip(cfgs, groups)]
    n = len(groups)
    ins = (ctypes.c_void_p * n)(*[d.data_ptr() for d in ds])
    op = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs])
    _lib.check(_lib.load().shtk_batch_analysis(ctypes.c_void_p(b._b), ins, op, _lib.SHTK_F64, _stream(ds[0].device)),
               "batch analysis")
    for (L, nf), f, o in zip(groups, fs, outs):
        got = to_host(o)
        assert not np.isnan(got).any(), L
        ref = oracle.Config(L, L + 1, nlon=2 * L + 2).analysis(f)
        assert _per_field(got, ref) < TOL["split"], L
    b.close()
