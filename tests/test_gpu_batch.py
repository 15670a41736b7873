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


# The 64-field uneven batch of the benchmark draws lmax from {63, 127, 255, 511, 1023}
# with nlat = lmax + 1, nphi = 2 nlat; here three fields per size run through the same
# grouped kernels (the lmax 511 and 1023 groups take the two-half analysis blocks).
C3_GROUPS = [(1023, 3), (511, 3), (255, 3), (127, 3), (63, 3)]


def test_batch_benchmark_sizes_split_graph_autotuned(shtk, oracle, cuda):
    """One sampled field per group against the oracle (synthesis and analysis),
    every field by round trip, on the captured graph with autotuned launch shapes."""
    cfgs = [shtk.create_gauss_config(L, L + 1, nlon=2 * (L + 1)) for L, _ in C3_GROUPS]
    ocfgs = [oracle.Config(L, L + 1, nlon=2 * (L + 1)) for L, _ in C3_GROUPS]
    nfs = [nf for _, nf in C3_GROUPS]
    alms = [oracle.random_alm(L, L, seed=300 + L, power=1.0, nf=nf) for L, nf in C3_GROUPS]
    b = shtk.UnevenBatch(cfgs, nfs, precision="split", graph=True)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        d = [to_dev(a) for a in alms]
        f0 = b.synthesis(d)
        choice = b.autotune([a.clone() for a in d], [f.clone() for f in f0], reps=1)
        print("autotune choice:", choice, flush=True)
        for _ in range(2):  # capture, then replay
            fs = b.synthesis(d)
            back = b.analysis(fs)
        torch.cuda.synchronize()
    worst = 0.0
    for (L, nf), oc, a, f, r in zip(C3_GROUPS, ocfgs, alms, fs, back):
        fh, rh = to_host(f), to_host(r)
        e_syn = rel_l2(fh[..., 0], oc.synthesis(a[..., 0]))
        e_ana = rel_l2(rh[..., 0], oc.analysis(fh[..., 0]))
        e_rt = max(rel_l2(rh[..., k], a[..., k]) for k in range(nf))
        print(f"lmax {L}: synthesis {e_syn:.2e} analysis {e_ana:.2e} round trip {e_rt:.2e}", flush=True)
        worst = max(worst, e_syn, e_ana, e_rt)
    assert worst < TOL_SPLIT
    b.close()


def test_batch_graph_survives_workspace_growth(shtk, oracle, cuda):
    """Captured graphs hold workspace addresses: FP64 calls, then FP32 calls (the
    split workspace grows by the FP32 accumulator), then FP64 again, then a larger
    call through a group's plan (its workspace grows) -- every replay must equal
    the eager result, i.e. stale graphs are recaptured."""
    import ctypes

    from shtnskit_jl_amd import _lib

    groups = [(63, 4), (31, 3)]
    cfgs = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    cfgs2 = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    nfs = [nf for _, nf in groups]
    alms = [oracle.random_alm(L, L, seed=60 + L, power=1.0, nf=nf) for L, nf in groups]
    eager = shtk.UnevenBatch(cfgs, nfs, precision="split")
    g = shtk.UnevenBatch(cfgs2, nfs, precision="split", graph=True)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        d64 = [to_dev(a) for a in alms]
        d32 = [a.to(torch.complex64) for a in d64]
        ref64 = [f.clone() for f in eager.synthesis(d64)]
        ref64a = [a.clone() for a in eager.analysis(ref64)]
        ref32 = [f.clone() for f in eager.synthesis(d32)]

        def check64(tag):
            fs = g.synthesis(d64)
            back = g.analysis(fs)
            torch.cuda.synchronize()
            for f, r in zip(fs, ref64):
                assert rel_l2(to_host(f), to_host(r)) < 1e-14, tag
            for a, r in zip(back, ref64a):
                assert rel_l2(to_host(a), to_host(r)) < 1e-14, tag

        check64("first capture")
        check64("replay")
        fs32 = g.synthesis(d32)
        g.analysis(fs32)
        torch.cuda.synchronize()
        for f, r in zip(fs32, ref32):
            assert rel_l2(to_host(f), to_host(r)) < 1e-6
        check64("after the FP32 calls")
        # a larger call through group 0's plan grows that plan's workspace
        plan = ctypes.c_void_p()
        lib = _lib.load()
        _lib.check(lib.shtk_batch_plan(ctypes.c_void_p(g._b), 0, ctypes.byref(plan)), "batch_plan")
        L, _ = groups[0]
        big = to_dev(oracle.random_alm(L, L, seed=5, power=1.0, nf=40))
        fbig = torch.empty((40, 2 * L + 2, L + 1), dtype=torch.float64, device="cuda:0").permute(2, 1, 0)
        _lib.check(lib.shtk_synthesis(plan, ctypes.c_void_p(big.data_ptr()), ctypes.c_void_p(fbig.data_ptr()), 40,
                                      _lib.SHTK_F64, 1, -1, ctypes.c_void_p(s.cuda_stream)), "synthesis")
        check64("after a larger call through the group plan")
    eager.close()
    g.close()


def test_plan_shared_by_two_streams(shtk, oracle, cuda):
    """Calls on one plan from two streams, interleaved without synchronisation:
    the plan orders them on the device (they share its workspace), so every
    result matches the oracle."""
    lmax, nlat, nlon, nf = 255, 256, 512, 2
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    ocfg = oracle.Config(lmax, nlat, nlon=nlon)
    x = oracle.random_alm(lmax, lmax, seed=71, power=1.0, nf=nf)
    y = oracle.random_alm(lmax, lmax, seed=72, power=1.0, nf=nf)
    fx_ref, fy_ref = ocfg.synthesis(x), ocfg.synthesis(y)
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    dx, dy = to_dev(x), to_dev(y)
    torch.cuda.synchronize()
    outs = []
    for _ in range(6):
        with torch.cuda.stream(sa):
            fa = shtk.synthesis_batch(cfg, dx)
            ra = shtk.analysis_batch(cfg, fa)
        with torch.cuda.stream(sb):
            fb = shtk.synthesis_batch(cfg, dy)
            rb = shtk.analysis_batch(cfg, fb)
        outs.append((fa, ra, fb, rb))
    torch.cuda.synchronize()
    for fa, ra, fb, rb in outs:
        assert rel_l2(to_host(fa), fx_ref) < 1e-12 and rel_l2(to_host(fb), fy_ref) < 1e-12
        assert rel_l2(to_host(ra), x) < 1e-12 and rel_l2(to_host(rb), y) < 1e-12
