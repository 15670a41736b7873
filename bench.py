#!/usr/bin/env python3
"""Round-trip spherical harmonic transform benchmark on one MI355X.

Headline workload: plain FP64 round trip -- synthesis then analysis -- on the
Gauss grid lmax = mmax = 1023, nlat = 1024, nlon = 2048, one field,
orthonormal + Condon-Shortley, input alm complex Gaussian (seed 2) scaled by
(l+1)^-1.5 with m = 0 imaginary parts zero.  A step is one round trip and
`value` is round trips per second.

Second measured line (`split_batched`): the same grid, a batch of 16 fields
(seed 2, same spectrum), synthesis_batch then analysis_batch with the
split-precision contraction on the int8 matrix cores; its value is field round
trips per second of the whole batch, with the per-field relative L2 error
against the FP64 path.

Inputs are resident in HBM before every timed region.  `value` comes from K
steps with no instrumentation; per-kernel durations come from a second pass of
the same K steps in which the library records HIP events around every launch
on its launch stream (shtk_profile_*); the roofline entries are for the
dominant kernel of each line.  The CPU baseline is the oracle's restatement of the reference
CPU path (threads over m, per-step recurrence coefficients, no folding) on the
host cores.

HBM traffic of a dominant kernel is not measured live: it comes from two
rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over this script, scaled by
what the counters read for a plain device copy of known size (run with
--calibrate-copy so that copy is in the same trace), and is read back from
profiles/hbm_traffic.json when that file covers the kernel.

Fourth line (`sphtor_batch`): 32 vector fields at lmax 511 (512 x 1024), S and
T with l^-2 spectra, vector round trips in plain FP64 and in split precision.

Fifth line (`adjoint_batch`): loss and gradient of 16 FP32 fields at lmax 511
through the synthesis rrule (forward synthesis, weighted squared error, adjoint).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-split] [--no-uneven]
                       [--no-sphtor] [--no-adjoint] [--calibrate-copy]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "round-trip SHT/s at lmax=1023: plain FP64 vs batched split-FP32 at ≤1e-11 rel err"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix peaks coincide)
PEAK_INT8_TOPS = 5000.0  # MI355X dense int8 matrix
PEAK_FP32_MFMA_TFLOPS = 157.3  # MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0
SPLIT_PRODUCTS = 21  # slice products of the 6-slice Ozaki split (digit orders k + t <= 5)
LMAX, NLAT, NLON = 1023, 1024, 2048
SPLIT_FIELDS = 16


def n_triples(lmax: int, nlat: int) -> float:
    """Folded Legendre triples T = (nlat/2) * (lmax+1)(lmax+2)/2 per field and direction."""
    return (nlat / 2) * (lmax + 1) * (lmax + 2) / 2


def algorithmic_flops(lmax: int, nlat: int) -> float:
    """One direction of the plain transform: 8 flops per folded triple (4 for
    the recurrence step, 4 for the complex-by-real contraction)."""
    return 8.0 * n_triples(lmax, nlat)


def split_int8_ops(lmax: int, nlat: int, nfields: int) -> float:
    """One direction of the split contraction: each of the 21 slice products is
    a real GEMM over every folded triple and real column, 2 ops per multiply-add
    and 2 real columns per field."""
    return SPLIT_PRODUCTS * 2.0 * n_triples(lmax, nlat) * 2 * nfields


def fp32_mfma_flops(lmax: int, nlat: int, nfields: int) -> float:
    """One direction of the FP32 matrix-core contraction: every folded triple
    against 2 real columns per field, 2 flops per multiply-add."""
    return 4.0 * n_triples(lmax, nlat) * nfields


def fft_bytes(nlat: int, nlon: int, mmax: int, nfields: int = 1) -> float:
    """Compulsory HBM bytes of one phi-FFT launch: the real grid and the
    half-spectrum intermediate, read once and written once."""
    return nfields * (8.0 * nlat * nlon + 16.0 * nlat * (mmax + 1))


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
    from the committed rocprofv3 --pmc summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            tab = json.load(fh)
    except (OSError, ValueError):
        return None
    ent = tab.get("kernels", {}).get(kernel)
    return None if ent is None else ent.get("bytes_per_launch")


def _cpu_round_trips(oracle, lmax, nlat, nlon, power, seed, threads, warmup, runs):
    """Median wall time of `runs` round trips after `warmup` untimed ones (plan and
    tables built outside the timing), on `threads` OpenMP threads."""
    oracle.set_threads(threads)
    cfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=seed, power=power)
    for _ in range(warmup):
        cfg.analysis(cfg.synthesis(alm))
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        f = cfg.synthesis(alm)
        cfg.analysis(f)
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def cpu_baseline(lmax, nlat, nlon, single_thread_runs=2):
    """The oracle's restatement of the reference CPU path (threads over m in the
    balanced order, per-step recurrence coefficients, no folding), on the host
    cores this process may use and on one thread.  All-core figures: median of
    10 round trips after 2 warmups; the one-thread figure at lmax 1023 is a
    bounded sample (`single_thread_runs` after 1 warmup, about 10 s each); the
    lmax 127 workload (128 x 256) is timed the same way on both."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker / baseline only

    # the host share the process may use: OMP_NUM_THREADS if set, else its CPU affinity
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    t_all = _cpu_round_trips(oracle, lmax, nlat, nlon, 1.5, 2, threads, 2, 10)
    t_one = _cpu_round_trips(oracle, lmax, nlat, nlon, 1.5, 2, 1, 1, single_thread_runs)
    c1_all = _cpu_round_trips(oracle, 127, 128, 256, 1.0, 1, threads, 2, 10)
    c1_one = _cpu_round_trips(oracle, 127, 128, 256, 1.0, 1, 1, 2, 10)
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        model = "unknown"
    return {"value": 1.0 / t_all, "unit": "SHT/s", "cores": threads, "kind": "port",
            "sample": f"median of 10 round trips after 2 warmups at lmax=1023 (1024x2048, 1 field, seed 2, "
                      f"(l+1)^-1.5); OpenMP threads over m in balanced order, FP64, per-step recurrence "
                      f"coefficients, project-written FFT (FFTW absent); host CPU: {model}",
            "single_thread": {"value": 1.0 / t_one, "unit": "SHT/s", "cores": 1,
                              "sample": f"median of {single_thread_runs} round trips after 1 warmup, lmax=1023"},
            "lmax127": {"value": 1.0 / c1_all, "unit": "SHT/s", "cores": threads,
                        "single_thread_value": 1.0 / c1_one,
                        "sample": "median of 10 round trips after 2 warmups, lmax=127 (128x256, seed 1, "
                                  "(l+1)^-1), all cores and one thread"}}


def spectrum(np, nfields, seed=2):
    """Complex Gaussian alm (lmax+1, mmax+1[, nf]) scaled by (l+1)^-1.5, l < m and
    m = 0 imaginary parts zero."""
    rng = np.random.Generator(np.random.PCG64(seed))
    shape = (LMAX + 1, LMAX + 1) + ((nfields,) if nfields > 1 else ())
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    l = np.arange(LMAX + 1).reshape((-1, 1) + ((1,) if nfields > 1 else ()))
    m = np.arange(LMAX + 1).reshape((1, -1) + ((1,) if nfields > 1 else ()))
    a = np.where(l >= m, a * (l + 1.0) ** -1.5, 0)
    a[:, 0, ...] = a[:, 0, ...].real
    return a


def to_device(torch, np, a, dev):
    """numpy array -> device tensor with Julia column-major memory."""
    perm = tuple(range(a.ndim - 1, -1, -1))
    return torch.from_numpy(np.ascontiguousarray(a.transpose(perm))).to(dev).permute(*perm)


def timed_steps(torch, _lib, cfg, step, steps, stream):
    """K steps bracketed by synchronisations (the throughput), then the same K
    steps with the library's per-kernel HIP events on (the kernel durations of
    the roofline; the events add a little time between launches).  Returns
    (ms without events, ms with events, {kernel: (total_ms, count)}, last output)."""
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        out = step()
    e1.record(stream)
    torch.cuda.synchronize()
    plain = e0.elapsed_time(e1)
    _lib.profile_enable(cfg.plan, True)
    _lib.profile_read(cfg.plan)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(steps):
        out = step()
    e1.record(stream)
    torch.cuda.synchronize()
    with_ev = e0.elapsed_time(e1)
    kern = _lib.profile_read(cfg.plan)
    _lib.profile_enable(cfg.plan, False)
    return plain, with_ev, kern, out


def roofline_entry(kernel, avg_us, work, peak, unit, bound, note):
    """`work` per launch in units such that work / seconds / 1e12 is in `unit`
    (flop or op counts for T*/s, bytes * 1e3 for GB/s)."""
    achieved = work / (avg_us * 1e-6) / 1e12
    return {"kernel": kernel, "bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": measured_traffic(kernel),
            "algorithmic_per_launch": note, "avg_launch_us": round(avg_us, 2)}


def plain_roofline(kernels_us):
    dom = max(kernels_us, key=kernels_us.get)
    us = kernels_us[dom]
    if dom.startswith("leg"):
        fl = algorithmic_flops(LMAX, NLAT)
        ent = roofline_entry(dom, us, fl, PEAK_FP64_TFLOPS, "TFLOP/s", "mfma",
                             f"{fl:.4e} flop = 8 x (nlat/2) x (lmax+1)(lmax+2)/2 (FP64 VALU)")
        # compute-bound; the instructions are FP64 vector FMAs (v_fmac_f64), whose
        # peak equals the FP64 matrix peak on MI355X
        ent["engine"] = "FP64 VALU (peak = FP64 vector = FP64 matrix, 78.6 TFLOP/s)"
        return ent
    by = fft_bytes(NLAT, NLON, LMAX)
    return roofline_entry(dom, us, by * 1e3, PEAK_HBM_GBS, "GB/s", "hbm", f"{by:.4e} bytes")


def measure_uneven(torch, np, shtk, dev, steps):
    """64 fields with lmax drawn uniformly from {63, 127, 255, 511, 1023} (seed 3),
    nlat = lmax + 1, nlon = 2 nlat, (l+1)^-1 spectra: one split-precision batch
    object, launch sequence captured in a hipGraph, launch shapes autotuned.
    Value: field round trips per second of the whole batch."""
    rng = np.random.Generator(np.random.PCG64(3))
    draws = rng.choice([63, 127, 255, 511, 1023], size=64)
    groups = sorted({int(L): int((draws == L).sum()) for L in set(draws.tolist())}.items(), reverse=True)
    cfgs = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    nfs = [nf for _, nf in groups]
    alms = []
    for L, nf in groups:
        a = rng.standard_normal((L + 1, L + 1, nf)) + 1j * rng.standard_normal((L + 1, L + 1, nf))
        l = np.arange(L + 1)[:, None, None]
        m = np.arange(L + 1)[None, :, None]
        a = np.where(l >= m, a / (l + 1.0), 0)
        a[:, 0, :] = a[:, 0, :].real
        alms.append(to_device(torch, np, a, dev))
    fp64 = shtk.UnevenBatch(cfgs, nfs, precision="fp64")
    ref_f = fp64.synthesis(alms)
    ref_a = fp64.analysis(ref_f)
    torch.cuda.synchronize()
    cfgs2 = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    b = shtk.UnevenBatch(cfgs2, nfs, precision="split")
    tune = b.autotune([a.clone() for a in alms], [f.clone() for f in ref_f], reps=2)
    b.close()
    cfgs3 = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    g = shtk.UnevenBatch(cfgs3, nfs, precision="split", graph=True)
    g.set_tiles(tune["synth_blocks"], tune["anal_blocks"])
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        for _ in range(2):
            fs = g.synthesis(alms)
            back = g.analysis(fs)
        torch.cuda.synchronize()
        err = 0.0
        for fa, ra in zip(back, ref_a):
            for k in range(fa.shape[2]):
                err = max(err, float(torch.linalg.vector_norm(fa[..., k] - ra[..., k]) /
                                     torch.linalg.vector_norm(ra[..., k])))
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(steps):
            fs = g.synthesis(alms)
            back = g.analysis(fs)
        e1.record(s)
        torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    torch.cuda.synchronize()
    e0.record()
    for _ in range(steps):
        fp64.analysis(fp64.synthesis(alms))
    e1.record()
    torch.cuda.synchronize()
    fp64_ms = e0.elapsed_time(e1) / steps
    # FP32 matrix-core contraction, captured the same way
    cfgs4 = [shtk.create_gauss_config(L, L + 1, nlon=2 * L + 2) for L, _ in groups]
    g32 = shtk.UnevenBatch(cfgs4, nfs, precision="fp32", graph=True)
    with torch.cuda.stream(s):
        for _ in range(2):
            back32 = g32.analysis(g32.synthesis(alms))
        torch.cuda.synchronize()
        err32 = 0.0
        for fa, ra in zip(back32, ref_a):
            for k in range(fa.shape[2]):
                err32 = max(err32, float(torch.linalg.vector_norm(fa[..., k] - ra[..., k]) /
                                         torch.linalg.vector_norm(ra[..., k])))
        e0.record(s)
        for _ in range(steps):
            g32.analysis(g32.synthesis(alms))
        e1.record(s)
        torch.cuda.synchronize()
    fp32_ms = e0.elapsed_time(e1) / steps
    g32.close()
    g.close()
    fp64.close()
    return {"value": round(1000.0 * 64 / ms, 2), "unit": "field round trips/s", "fields": 64,
            "groups": {str(L): nf for L, nf in groups}, "steps": steps,
            "ms_per_batch_round_trip": round(ms, 4), "fp64_path_ms_per_batch_round_trip": round(fp64_ms, 4),
            "fp32_path_ms_per_batch_round_trip": round(fp32_ms, 4),
            "fp32_path_value": round(1000.0 * 64 / fp32_ms, 2),
            "max_field_rel_l2_vs_fp64_path": err, "fp32_max_field_rel_l2_vs_fp64_path": err32,
            "autotuned": tune,
            "launches": "one hipGraph per direction; Legendre stages as persistent kernels over all groups"}


def sphtor_spectrum(np, lmax, nfields, seed):
    """Complex Gaussian (seed) scaled by l^-2 for l >= 1, zero for l = 0, l < m
    entries and the m = 0 imaginary parts."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = rng.standard_normal((lmax + 1, lmax + 1, nfields)) + 1j * rng.standard_normal((lmax + 1, lmax + 1, nfields))
    l = np.arange(lmax + 1)[:, None, None]
    m = np.arange(lmax + 1)[None, :, None]
    a = np.where((l >= m) & (l >= 1), a / np.maximum(l, 1.0) ** 2, 0)
    a[:, 0, :] = a[:, 0, :].real
    return a


def measure_sphtor(torch, np, shtk, _lib, dev, steps, stream):
    """Vector (sphtor) batch: 32 fields, lmax 511, 512 x 1024, S and T l^-2
    spectra (seeds 4 and 5); synthesis_sphtor_batch then analysis_sphtor_batch
    in plain FP64 and in split precision.  Value: vector field round trips per
    second of the split batch."""
    lmax, nlat, nlon, nf = 511, 512, 1024, 32
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    S = to_device(torch, np, sphtor_spectrum(np, lmax, nf, 4), dev)
    T = to_device(torch, np, sphtor_spectrum(np, lmax, nf, 5), dev)
    torch.cuda.synchronize()
    out = {}
    for prec in ("fp64", "split"):
        def vstep():
            Vt, Vp = shtk.synthesis_sphtor_batch(cfg, S, T, precision=prec)
            return shtk.analysis_sphtor_batch(cfg, Vt, Vp, precision=prec)

        for _ in range(2):
            S2, T2 = vstep()
        torch.cuda.synchronize()
        err = max(max(float(torch.linalg.vector_norm(S2[..., k] - S[..., k]) / torch.linalg.vector_norm(S[..., k])),
                      float(torch.linalg.vector_norm(T2[..., k] - T[..., k]) / torch.linalg.vector_norm(T[..., k])))
                  for k in range(nf))
        no_ev, _, kern, _ = timed_steps(torch, _lib, cfg, vstep, steps, stream)
        out[prec] = (no_ev / steps, err, {k: round(1000.0 * v[0] / v[1], 2) for k, v in kern.items()})
    ms_split, err_split, k_split = out["split"]
    ms_plain, err_plain, k_plain = out["fp64"]
    cfg.close()
    return {"value": round(1000.0 * nf / ms_split, 2), "unit": "vector field round trips/s", "fields": nf,
            "steps": steps, "ms_per_batch_round_trip": round(ms_split, 4),
            "fp64_path_value": round(1000.0 * nf / ms_plain, 2),
            "fp64_path_ms_per_batch_round_trip": round(ms_plain, 4),
            "speedup_split_vs_fp64": round(ms_plain / ms_split, 3),
            "max_field_rel_l2_round_trip": {"split": err_split, "fp64": err_plain},
            "config": {"workload": "vector (sphtor) batch round trip, synthesis_sphtor_batch+analysis_sphtor_batch",
                       "lmax": lmax, "nlat": nlat, "nlon": nlon, "fields": nf,
                       "data": "complex Gaussian S, T (seeds 4, 5), l^-2 spectra"},
            "kernels_avg_us": {"split": k_split, "fp64": k_plain}}


def measure_adjoint(torch, np, shtk, _lib, dev, steps, stream):
    """Loss and gradient through the synthesis rrule: 16 fields, lmax 511,
    512 x 1024, FP32 coefficients (complex Gaussian, seed 6, (l+1)^-1 spectrum),
    target t = synthesis(alm_t) (seed 7), L = sum_k sum_ij w_i (f_k - t_k)^2 with
    f_k = synthesis(alm_k); a step is the forward synthesis, the loss and the
    backward pass (the adjoint synthesis).  Value: fields per second."""
    lmax, nlat, nlon, nf = 511, 512, 1024, 16
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)

    def spec(seed):
        rng = np.random.Generator(np.random.PCG64(seed))
        a = rng.standard_normal((lmax + 1, lmax + 1, nf)) + 1j * rng.standard_normal((lmax + 1, lmax + 1, nf))
        l = np.arange(lmax + 1)[:, None, None]
        m = np.arange(lmax + 1)[None, :, None]
        a = np.where(l >= m, a / (l + 1.0), 0)
        a[:, 0, :] = a[:, 0, :].real
        return to_device(torch, np, a.astype(np.complex64), dev)

    a = spec(6).requires_grad_(True)
    t = shtk.synthesis_batch(cfg, spec(7)).detach()
    w = torch.as_tensor(cfg.w, dtype=torch.float32, device=dev)[:, None, None]
    torch.cuda.synchronize()

    def step():
        a.grad = None
        f = shtk.ad.synthesis(cfg, a)
        loss = (w * (f - t) ** 2).sum()
        loss.backward()
        return loss

    for _ in range(2):
        step()
    torch.cuda.synchronize()
    no_ev, _, kern, _ = timed_steps(torch, _lib, cfg, step, steps, stream)
    ms = no_ev / steps
    cfg.close()
    return {"value": round(1000.0 * nf / ms, 2), "unit": "field loss+gradient/s", "fields": nf, "steps": steps,
            "ms_per_batch": round(ms, 4), "dtype": "f32 I/O, FP64 Legendre and FFT",
            "config": {"workload": "loss and gradient through the synthesis rrule (forward synthesis + adjoint)",
                       "lmax": lmax, "nlat": nlat, "nlon": nlon, "fields": nf,
                       "data": "complex Gaussian alm (seed 6) and target (seed 7), (l+1)^-1 spectra, FP32"},
            "kernels_avg_us": {k: round(1000.0 * v[0] / v[1], 2) for k, v in kern.items()}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--warmup-seconds", type=float, default=0.5,
                    help="the plain warmup repeats the step (past --warmup steps) until this much time has "
                         "passed: a GPU fresh from idle runs its first tens of milliseconds of work slowly")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-split", action="store_true", help="skip the batched split-precision line")
    ap.add_argument("--no-uneven", action="store_true", help="skip the uneven-lmax batch line")
    ap.add_argument("--no-sphtor", action="store_true", help="skip the vector (sphtor) batch line")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the loss+gradient line")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run 5 device copies of 256 MiB first (HBM counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the path shards by field or by m without changing the "
              f"per-GPU kernel); --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import __graft_entry__

    shtk = __graft_entry__.load_package()
    from shtnskit_jl_amd import _lib

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    cfg = shtk.create_gauss_config(LMAX, NLAT, nlon=NLON)

    if args.calibrate_copy:
        src = torch.empty(1 << 25, dtype=torch.float64, device=dev).normal_()
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- plain FP64, one field (the headline value)
    alm = to_device(torch, np, spectrum(np, 1), dev)
    torch.cuda.synchronize()

    def step():
        return shtk.analysis(cfg, shtk.synthesis(cfg, alm))

    t_w = time.perf_counter()
    n_w = 0
    while n_w < args.warmup or time.perf_counter() - t_w < args.warmup_seconds:
        out = step()
        n_w += 1
        if n_w % 50 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    rt_err = float(torch.linalg.vector_norm(out - alm) / torch.linalg.vector_norm(alm))
    no_ev, with_ev, kern, _ = timed_steps(torch, _lib, cfg, step, args.steps, stream)
    ms_per_step = no_ev / args.steps
    kernels_us = {k: 1000.0 * v[0] / v[1] for k, v in kern.items()}

    res = {
        "metric": METRIC,
        "value": round(1000.0 / ms_per_step, 2),
        "unit": "SHT/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_steps_run": n_w,
        "ms_per_step": round(ms_per_step, 5),
        "ms_per_step_with_kernel_events": round(with_ev / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (complex Gaussian alm, seed 2, (l+1)^-1.5 spectrum)",
        "config": {"workload": "plain FP64 scalar round trip (synthesis+analysis), Gauss grid",
                   "lmax": LMAX, "mmax": LMAX, "nlat": NLAT, "nlon": NLON, "fields": 1,
                   "norm": "orthonormal+CS", "parallelism": "single GPU"},
        "roofline": plain_roofline(kernels_us),
        "kernels_avg_us": {k: round(v, 2) for k, v in kernels_us.items()},
        "round_trip_rel_l2": rt_err,
    }

    # ---- batched split precision, 16 fields
    if not args.no_split:
        nf = SPLIT_FIELDS
        balm = to_device(torch, np, spectrum(np, nf), dev)
        torch.cuda.synchronize()

        def sstep():
            return shtk.analysis_batch(cfg, shtk.synthesis_batch(cfg, balm, precision="split"), precision="split")

        for _ in range(args.warmup):
            sout = sstep()
        ref = shtk.analysis_batch(cfg, shtk.synthesis_batch(cfg, balm))  # FP64 batch path
        torch.cuda.synchronize()
        err_fp64 = max(float(torch.linalg.vector_norm(sout[..., k] - ref[..., k]) /
                             torch.linalg.vector_norm(ref[..., k])) for k in range(nf))
        err_in = max(float(torch.linalg.vector_norm(sout[..., k] - balm[..., k]) /
                           torch.linalg.vector_norm(balm[..., k])) for k in range(nf))
        ssteps = max(4, args.steps // 4)
        s_no, s_ev, skern, _ = timed_steps(torch, _lib, cfg, sstep, ssteps, stream)
        sk_us = {k: 1000.0 * v[0] / v[1] for k, v in skern.items()}
        dom = max(sk_us, key=sk_us.get)
        ops = split_int8_ops(LMAX, NLAT, nf)
        if dom.startswith("oz"):
            sroof = roofline_entry(dom, sk_us[dom], ops, PEAK_INT8_TOPS, "TOP/s", "mfma",
                                   f"{ops:.4e} int8 ops = 21 slice products x 2 x (nlat/2)(lmax+1)(lmax+2)/2 "
                                   f"x 2 x {nf} fields")
        else:
            by = fft_bytes(NLAT, NLON, LMAX, nf)
            sroof = roofline_entry(dom, sk_us[dom], by * 1e3, PEAK_HBM_GBS, "GB/s", "hbm", f"{by:.4e} bytes")
        ms_batch = s_no / ssteps
        res["split_batched"] = {
            "value": round(1000.0 * nf / ms_batch, 2),
            "unit": "SHT/s",
            "fields": nf,
            "steps": ssteps,
            "ms_per_batch_round_trip": round(ms_batch, 5),
            "ms_per_batch_with_kernel_events": round(s_ev / ssteps, 5),
            "speedup_vs_plain": round((1000.0 * nf / ms_batch) / res["value"], 3),
            "dtype": "int8 MFMA slices of FP64 Legendre values, FP64 tables and cross-slice sums",
            "max_field_rel_l2_vs_fp64_path": err_fp64,
            "max_field_rel_l2_round_trip": err_in,
            "config": {"workload": "batched split-precision scalar round trip (synthesis_batch+analysis_batch)",
                       "lmax": LMAX, "nlat": NLAT, "nlon": NLON, "fields": nf},
            "roofline": sroof,
            "kernels_avg_us": {k: round(v, 2) for k, v in sk_us.items()},
        }

    # ---- batched FP32 matrix-core contraction, 16 fields
    if not args.no_split:
        nf = SPLIT_FIELDS
        balm = to_device(torch, np, spectrum(np, nf), dev)
        torch.cuda.synchronize()

        def fstep():
            return shtk.analysis_batch(cfg, shtk.synthesis_batch(cfg, balm, precision="fp32"), precision="fp32")

        for _ in range(args.warmup):
            fout = fstep()
        fsyn = shtk.synthesis_batch(cfg, balm, precision="fp32")
        rsyn = shtk.synthesis_batch(cfg, balm)  # FP64 batch path
        torch.cuda.synchronize()
        err_syn = max(float(torch.linalg.vector_norm(fsyn[..., k] - rsyn[..., k]) /
                            torch.linalg.vector_norm(rsyn[..., k])) for k in range(nf))
        err_in = max(float(torch.linalg.vector_norm(fout[..., k] - balm[..., k]) /
                           torch.linalg.vector_norm(balm[..., k])) for k in range(nf))
        fsteps = max(4, args.steps // 4)
        f_no, f_ev, fkern, _ = timed_steps(torch, _lib, cfg, fstep, fsteps, stream)
        fk_us = {k: 1000.0 * v[0] / v[1] for k, v in fkern.items()}
        dom = max(fk_us, key=fk_us.get)
        if dom.startswith("f32"):
            fl = fp32_mfma_flops(LMAX, NLAT, nf)
            froof = roofline_entry(dom, fk_us[dom], fl, PEAK_FP32_MFMA_TFLOPS, "TFLOP/s", "mfma",
                                   f"{fl:.4e} flop = 2 x 2 real columns x (nlat/2)(lmax+1)(lmax+2)/2 x {nf} fields")
        else:
            by = fft_bytes(NLAT, NLON, LMAX, nf)
            froof = roofline_entry(dom, fk_us[dom], by * 1e3, PEAK_HBM_GBS, "GB/s", "hbm", f"{by:.4e} bytes")
        ms_batch = f_no / fsteps
        res["fp32_batched"] = {
            "value": round(1000.0 * nf / ms_batch, 2),
            "unit": "SHT/s",
            "fields": nf,
            "steps": fsteps,
            "ms_per_batch_round_trip": round(ms_batch, 5),
            "speedup_vs_plain": round((1000.0 * nf / ms_batch) / res["value"], 3),
            "dtype": "FP32 MFMA contraction (FP32 tiles of 64 products, FP64 sums across tiles), FP64 Legendre",
            "max_field_rel_l2_synthesis_vs_fp64_path": err_syn,
            "max_field_rel_l2_round_trip": err_in,
            "config": {"workload": "batched FP32-MFMA scalar round trip (synthesis_batch+analysis_batch)",
                       "lmax": LMAX, "nlat": NLAT, "nlon": NLON, "fields": nf},
            "roofline": froof,
            "kernels_avg_us": {k: round(v, 2) for k, v in fk_us.items()},
        }

    if not args.no_split and not args.no_uneven:
        res["uneven_batch"] = measure_uneven(torch, np, shtk, dev, max(3, args.steps // 10))

    if not args.no_sphtor:
        res["sphtor_batch"] = measure_sphtor(torch, np, shtk, _lib, dev, max(4, args.steps // 5), stream)

    if not args.no_adjoint:
        res["adjoint_batch"] = measure_adjoint(torch, np, shtk, _lib, dev, max(4, args.steps // 5), stream)

    res["cpu_baseline"] = None if args.no_cpu else cpu_baseline(LMAX, NLAT, NLON)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
