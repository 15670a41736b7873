"""Packed SHTns LM order of the host mirror (``layout.py``: LM_index, pack_lm,
unpack_lm; src/layout.jl:143-162, 282-339), pinned by the order in which the
SHTns 3.7 fixtures store their packed coefficients (the fixture reader's
``lm_pairs``: m-major blocks, l = m..lmax inside a block) and by the oracle's
restatement of LM_index.  Pure host logic: runs without a GPU."""
import numpy as np
import pytest
import torch

import _fixtures as FX


@pytest.mark.parametrize("lmax,mmax,mres", [(3, 3, 1), (10, 7, 1), (12, 12, 3), (9, 8, 2), (0, 0, 1)])
def test_lm_index_is_fixture_order(shtk, oracle, lmax, mmax, mres):
    from shtnskit_jl_amd.layout import LM_index

    pairs = FX.lm_pairs(lmax, mmax, mres)
    for k, (l, m) in enumerate(pairs):
        assert LM_index(lmax, mres, l, m) == k
        assert oracle.lm_index(lmax, mres, l, m) == k
    cfg = shtk.create_gauss_config(lmax, lmax + 1, mmax=mmax, mres=mres, nlon=2 * mmax + 2)
    assert cfg.nlm == len(pairs)


def test_lm_index_rejects_bad_pairs():
    from shtnskit_jl_amd.layout import LM_index

    with pytest.raises(ValueError):
        LM_index(5, 1, 2, 3)   # l < m
    with pytest.raises(ValueError):
        LM_index(5, 2, 3, 1)   # m not a multiple of mres
    with pytest.raises(ValueError):
        LM_index(5, 1, 6, 0)   # l > lmax


@pytest.mark.parametrize("fid", ["packed_storage", "scalar_batch", "scalar_l"])
def test_pack_unpack_against_fixture_reader(shtk, fid):
    from shtnskit_jl_amd.layout import pack_lm, unpack_lm

    fx = FX.fixture(fid)
    q = np.asarray(FX.payloads(fx)["coefficients"])
    cfg = shtk.create_gauss_config(fx["lmax"], fx["lmax"] + 1, mmax=fx["mmax"], mres=fx["mres"],
                                   nlon=fx["nphi"])
    A = unpack_lm(cfg, torch.from_numpy(q))
    assert np.array_equal(A.numpy(), FX.unpack(fx, q))
    assert np.array_equal(pack_lm(cfg, A).numpy(), q)
    with pytest.raises(ValueError):
        unpack_lm(cfg, torch.from_numpy(q[:-1]))
