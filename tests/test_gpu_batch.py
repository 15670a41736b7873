"""Uneven batches (several lmax in one batch object): per-field parity with
the FP64 oracle in split precision, identity with the per-configuration FP64
path, hipGraph replays against the eager result, and the autotuner."""
import numpy as np
import pytest
import torch

from _util import rel_l2, to_dev, to_host

pytestmark = pytest.mark.gpu

TOL_SPLIT = 1e-11
GROUPS = [(63, 4), (31, 3), (15, 17), (47, 2)]  # (lmax, fields); nlat = lmax + 1, nlon = 2 nlat


def _cfgs(shtk, oracle):
    cfgs = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in GROUPS]
    ocfgs = [oracle.Config(L, L + 1, nlon=2 * L + 2) for L, _ in GROUPS]
    alms = [oracle.random_alm(L, L, seed=40 + L, power=1.0, nf=nf) for L, nf in GROUPS]
    return cfgs, ocfgs, alms


def test_batch_split_matches_oracle(shtk, oracle, cuda):
    cfgs, ocfgs, alms = _cfgs(shtk, oracle)
    b = shtk.UnevenBatch(cfgs, [nf for _, nf in GROUPS], precision="split")
    fs = b.synthesis([to_dev(a) for a in alms])
    worst = 0.0
    for oc, a, f in zip(ocfgs, alms, fs):
        ref = oc.synthesis(a)
        fh = to_host(f)
        for k in range(a.shape[2]):
            worst = max(worst, rel_l2(fh[..., k], ref[..., k]))
    back = b.analysis(fs)
    for oc, a, f, r in zip(ocfgs, alms, fs, back):
        ref = oc.analysis(to_host(f))
        rh = to_host(r)
        for k in range(a.shape[2]):
            worst = max(worst, rel_l2(rh[..., k], ref[..., k]), rel_l2(rh[..., k], a[..., k]))
    print(f"uneven split batch: worst per-field rel L2 {worst:.2e}")
    assert worst < TOL_SPLIT
    b.close()


def test_batch_fp64_equals_per_config_path(shtk, oracle, cuda):
    cfgs, ocfgs, alms = _cfgs(shtk, oracle)
    b = shtk.UnevenBatch(cfgs, [nf for _, nf in GROUPS], precision="fp64")
    d = [to_dev(a) for a in alms]
    fs = b.synthesis(d)
    for c, a, f in zip(cfgs, d, fs):
        assert torch.equal(f, shtk.synthesis_batch(c, a))
    back = b.analysis(fs)
    for c, f, r in zip(cfgs, fs, back):
        assert torch.equal(r, shtk.analysis_batch(c, f))
    b.close()


@pytest.mark.parametrize("precision", ["split", "fp64"])
def test_batch_graph_replays(shtk, oracle, cuda, precision):
    """Captured once, replayed several times with eager work and
    synchronisations in between; every replay equals the eager result."""
    cfgs, ocfgs, alms = _cfgs(shtk, oracle)
    nfs = [nf for _, nf in GROUPS]
    eager = shtk.UnevenBatch(cfgs, nfs, precision=precision)
    d = [to_dev(a) for a in alms]
    ref_f = [f.clone() for f in eager.synthesis(d)]
    ref_a = [a.clone() for a in eager.analysis(ref_f)]
    cfgs2 = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in GROUPS]
    g = shtk.UnevenBatch(cfgs2, nfs, precision=precision, graph=True)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        for rep in range(4):
            fs = g.synthesis(d)
            back = g.analysis(fs)
            torch.cuda.synchronize()
            for f, r in zip(fs, ref_f):
                assert rel_l2(to_host(f), to_host(r)) < 1e-14, rep
            for a, r in zip(back, ref_a):
                assert rel_l2(to_host(a), to_host(r)) < 1e-14, rep
            # unrelated eager work between replays
            x = torch.randn(1 << 20, device="cuda:0")
            (x * 2).sum().item()
    eager.close()
    g.close()


def test_batch_autotune(shtk, oracle, cuda):
    cfgs, ocfgs, alms = _cfgs(shtk, oracle)
    nfs = [nf for _, nf in GROUPS]
    b = shtk.UnevenBatch(cfgs, nfs, precision="split")
    d = [to_dev(a) for a in alms]
    fs = b.synthesis(d)
    work_a = [a.clone() for a in d]
    work_f = [f.clone() for f in fs]
    choice = b.autotune(work_a, work_f, reps=2)
    print("autotune choice:", choice)
    assert choice["round_trip_ms"] > 0
    assert choice["synth_blocks"] >= 0 and choice["anal_blocks"] >= 0
    fs2 = b.synthesis(d)
    for oc, a, f in zip(ocfgs, alms, fs2):
        assert rel_l2(to_host(f), oc.synthesis(a)) < TOL_SPLIT
    b.close()


def test_batch_rejects_bad_input(shtk, cuda):
    with pytest.raises(ValueError):
        shtk.UnevenBatch([shtk.create_gauss_config(7, 8)], [1, 2])
    cfg = shtk.create_gauss_config(7, 8, nlon=16)
    b = shtk.UnevenBatch([cfg], [2], precision="split")
    with pytest.raises(ValueError):
        b.synthesis([torch.zeros((8, 8, 3), dtype=torch.complex128, device="cuda:0")])
    b.close()
