#!/usr/bin/env python3
"""Round-trip spherical harmonic transform benchmark on one MI355X.

Workload (the configuration the headline metric is quoted on): plain FP64
round trip -- synthesis then analysis -- on the Gauss grid lmax = mmax = 1023,
nlat = 1024, nlon = 2048, one field, orthonormal + Condon-Shortley, input
alm complex Gaussian (seed 2) scaled by (l+1)^-1.5 with m = 0 imaginary parts
zero.  Inputs are resident in HBM before the timed region starts.

A step is one round trip.  `value` is round trips per second.  The roofline
entry is for the dominant kernel, timed with HIP events on the stream the
library launches it on, inside the timed region.  The CPU baseline is the
oracle's restatement of the reference CPU path (threads over m, per-step
recurrence coefficients, no folding) on the host cores.

HBM traffic of the dominant kernel is not measured live: it comes from two
rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over this same script, scaled
by what the counters read for a plain device copy of known size (run with
--calibrate-copy so that copy is in the same trace), and is read back from
profiles/hbm_traffic.json when that file covers the dominant kernel.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--calibrate-copy]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "round-trip SHT/s at lmax=1023: plain FP64 vs batched split-FP32 at ≤1e-11 rel err"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix peaks coincide)
PEAK_HBM_GBS = 8000.0


def algorithmic_flops(lmax: int, nlat: int) -> float:
    """Folded Legendre triples T = (nlat/2) * N_lm; one direction costs 8 flops per
    triple (4 for the recurrence step, 4 for the complex-by-real contraction)."""
    nlm = (lmax + 1) * (lmax + 2) // 2
    return 8.0 * (nlat / 2) * nlm


def fft_bytes(nlat: int, nlon: int, mmax: int) -> float:
    """Compulsory HBM bytes of one phi-FFT launch: the real grid and the
    half-spectrum intermediate, read once and written once."""
    return 8.0 * nlat * nlon + 16.0 * nlat * (mmax + 1)


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


def cpu_baseline(lmax, nlat, nlon, budget_s=15.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle  # checker / baseline only

    # the host share the process may use: OMP_NUM_THREADS if set, else its CPU affinity
    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or len(os.sched_getaffinity(0))
    oracle.set_threads(threads)
    cfg = oracle.Config(lmax, nlat, nlon=nlon)
    alm = oracle.random_alm(lmax, lmax, seed=2, power=1.5)
    times = []
    t0 = time.perf_counter()
    f = cfg.synthesis(alm)
    cfg.analysis(f)
    first = time.perf_counter() - t0
    n = max(2, min(9, int(budget_s / max(first, 1e-3))))
    for _ in range(n):
        t0 = time.perf_counter()
        f = cfg.synthesis(alm)
        cfg.analysis(f)
        times.append(time.perf_counter() - t0)
    med = statistics.median(times)
    return {"value": 1.0 / med, "unit": "SHT/s", "cores": threads, "kind": "port",
            "sample": f"{n} round trips at lmax=1023 (1024x2048, 1 field) after 1 warmup, median; "
                      f"OpenMP threads over m in balanced order, FP64, per-step recurrence coefficients"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
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

    lmax, nlat, nlon = 1023, 1024, 2048
    dev = torch.device("cuda:0")
    cfg = shtk.create_gauss_config(lmax, nlat, nlon=nlon)
    rng = np.random.Generator(np.random.PCG64(2))
    a = rng.standard_normal((lmax + 1, lmax + 1)) + 1j * rng.standard_normal((lmax + 1, lmax + 1))
    l = np.arange(lmax + 1)[:, None]
    m = np.arange(lmax + 1)[None, :]
    a = np.where(l >= m, a * (l + 1.0) ** -1.5, 0)
    a[:, 0] = a[:, 0].real
    alm = torch.from_numpy(np.ascontiguousarray(a.T)).to(dev).T  # column-major (l fastest)
    torch.cuda.synchronize()

    if args.calibrate_copy:
        src = torch.empty(1 << 25, dtype=torch.float64, device=dev).normal_()
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def step():
        return shtk.analysis(cfg, shtk.synthesis(cfg, alm))

    for _ in range(args.warmup):
        out = step()
    torch.cuda.synchronize()
    rt_err = float(torch.linalg.vector_norm(out - alm) / torch.linalg.vector_norm(alm))

    stream = torch.cuda.current_stream(dev)
    _lib.profile_enable(cfg.plan, True)
    _lib.profile_read(cfg.plan)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        out = step()
    e1.record(stream)
    torch.cuda.synchronize()
    total_ms = e0.elapsed_time(e1)
    kern = _lib.profile_read(cfg.plan)
    _lib.profile_enable(cfg.plan, False)

    # the same region without per-kernel events, to expose their overhead
    e0.record(stream)
    for _ in range(args.steps):
        out = step()
    e1.record(stream)
    torch.cuda.synchronize()
    plain_ms = e0.elapsed_time(e1)

    ms_per_step = total_ms / args.steps
    kernels_us = {k: 1000.0 * v[0] / v[1] for k, v in kern.items()}
    dom = max(kern, key=lambda k: kern[k][0])
    dom_us = kernels_us[dom]
    if dom.startswith("leg"):
        flops = algorithmic_flops(lmax, nlat)
        achieved = flops / (dom_us * 1e-6) / 1e12
        roof = {"kernel": dom, "bound": "mfma", "compute": "fp64",
                "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": measured_traffic(dom),
                "algorithmic_per_launch": f"{flops:.4e} flop = 8 x (nlat/2) x (lmax+1)(lmax+2)/2",
                "avg_launch_us": round(dom_us, 2)}
    else:
        byts = fft_bytes(nlat, nlon, lmax)
        achieved = byts / (dom_us * 1e-6) / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(dom),
                "avg_launch_us": round(dom_us, 2)}

    res = {
        "metric": METRIC,
        "value": round(1000.0 / ms_per_step, 2),
        "unit": "SHT/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "ms_per_step_without_kernel_events": round(plain_ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (complex Gaussian alm, seed 2, (l+1)^-1.5 spectrum)",
        "config": {"workload": "plain FP64 scalar round trip (synthesis+analysis), Gauss grid",
                   "lmax": lmax, "mmax": lmax, "nlat": nlat, "nlon": nlon, "fields": 1,
                   "norm": "orthonormal+CS", "parallelism": "single GPU"},
        "roofline": roof,
        "kernels_avg_us": {k: round(v, 2) for k, v in kernels_us.items()},
        "event_overhead_pct": round(100.0 * (total_ms - plain_ms) / plain_ms, 2),
        "round_trip_rel_l2": rt_err,
    }
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(lmax, nlat, nlon)
    else:
        res["cpu_baseline"] = None
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
