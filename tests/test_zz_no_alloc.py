"""shtk_synthesis_static (the synthesis that allocates nothing, its workspace a
fixed device array per GPU) against shtk_synthesis through the C ABI: two calls
in a row, the second with a smaller input, and the size limit."""
import ctypes

import pytest
import torch

from _util import rel_l2, to_dev

pytestmark = pytest.mark.gpu


def _call(lib, entry, cfg, alm, nf, real_output=1):
    from shtnskit_jl_amd import _lib

    shape = (cfg.nlat, cfg.nlon, nf)
    dt = torch.float64 if real_output else torch.complex128
    out = torch.empty(shape[::-1], dtype=dt, device="cuda:0").permute(2, 1, 0)
    rc = getattr(lib, entry)(ctypes.c_void_p(cfg.plan), ctypes.c_void_p(alm.data_ptr()),
                             ctypes.c_void_p(out.data_ptr()), nf, _lib.SHTK_F64, real_output, -1,
                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc, entry)
    return out


def test_static_synthesis_matches_allocating_entry(shtk, oracle, cuda):
    from shtnskit_jl_amd import _lib

    lib = _lib.load()
    cases = [(1023, 1024, 2048, 1, 1), (255, 256, 512, 3, 1), (63, 64, 128, 2, 0)]  # second call smaller
    for lmax, nlat, nlon, nf, real in cases:
        print(f"static synthesis lmax={lmax} nf={nf} real_output={real}", flush=True)
        cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
        alm = to_dev(oracle.random_alm(lmax, lmax, seed=lmax, power=1.0, nf=nf))
        got = _call(lib, "shtk_synthesis_static", cfg, alm, nf, real)
        ref = _call(lib, "shtk_synthesis", cfg, alm, nf, real)
        torch.cuda.synchronize()
        err = rel_l2(got.cpu().numpy(), ref.cpu().numpy())
        print(f"  rel L2 vs shtk_synthesis {err:.2e} (identical: {torch.equal(got, ref)})", flush=True)
        assert err < 1e-12
        assert rel_l2(got.cpu().numpy(), oracle.Config(lmax, nlat, nlon=nlon).synthesis(
            alm.cpu().numpy(), real_output=bool(real))) < 1e-12


def test_static_synthesis_rejects_oversize(shtk, cuda):
    """A workspace above the fixed maximum is refused on the host, before any launch."""
    from shtnskit_jl_amd import _lib

    lib = _lib.load()
    cfg = shtk.create_gauss_config(1023, 1024, nlon=2048)
    nf = 3  # 3 x 33.6 MB > 64 MiB
    assert lib.shtk_workspace_size(ctypes.c_void_p(cfg.plan), nf, 0) > (64 << 20)
    alm = torch.zeros((nf, 1024, 1024), dtype=torch.complex128, device="cuda:0").permute(2, 1, 0)
    with pytest.raises(ValueError, match="static scratch"):
        _call(lib, "shtk_synthesis_static", cfg, alm, nf)
