#!/usr/bin/env python
"""Benchmark: the kspaceFirstOrder3D step loop on the 256^3 FP32 workload
(heterogeneous c and rho, power-law absorption y = 1.5, additive 1 MHz
tone-burst disc source, full-grid p_max and p_rms, 1000 steps) on one GPU.

python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one time step of the whole grid.  W warm-up steps run on a first
simulation; then a fresh simulation of the same workload runs exactly K
steps from t = 1 with every input already resident in HBM, timed with the
host clock around stream synchronisations (and with device events).  A third
run brackets every launch with device events to get per-kernel durations and
compulsory bytes for the roofline of the dominant kernel.  The CPU baseline is
the NumPy/scipy.fft restatement of the reference algorithm (oracle/) in
single precision on the host cores, timed on a few steps of the same
workload.  Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "k-wave-ii_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

HBM_PEAK_GBS = 8000.0
# whole-step algorithmic bytes per cell-step of this workload (per-axis-pass
# model: 14 FFTs x 3 passes x 8 B, fused elementwise operands, FP64 p_max and
# p_rms read+write): 536 B while the source is on, 488 B after it
B_ALG_SRC_ON, B_ALG_SRC_OFF = 536.0, 488.0
METRIC = "Mcell-steps/s 256³ FP32 vs plain GPU; served p50/p99 latency at ≤1e-4 rel err"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(N, nsteps, cores):
    """Oracle (float32, scipy.fft workers = cores) on the first nsteps steps
    of the same workload; returns Mcell-steps/s from the median step time of
    the steps after the first two."""
    import numpy as np
    from kwhip import workloads
    from oracle import kwave_oracle as ko

    kg, med, src, sens, opts = workloads.cfg2(N, Nt=nsteps)
    kgo = ko.KGrid(N, kg.dx, N, kg.dy, N, kg.dz)
    kgo.set_time(nsteps, kg.dt)
    times = []
    ko.kspace_first_order(kgo, med, src, sens, data_cast="single", pml_size=opts["PMLSize"],
                          pml_alpha=opts["PMLAlpha"], fft_workers=cores, step_times=times)
    tail = sorted(times[2:]) if len(times) > 2 else sorted(times)
    med_t = tail[len(tail) // 2]
    return N ** 3 / med_t / 1e6, med_t, len(times)


def accuracy_check(N=64, Nt=300):
    """cfg2 recipe at N^3 for Nt steps: GPU vs the FP64 and FP32 oracle
    (relative L-inf of p_max and p_rms)."""
    import kwhip
    from kwhip import workloads
    from oracle import kwave_oracle as ko

    kg, med, src, sens, opts = workloads.cfg2(N, Nt=Nt, pml=10, src_radius=max(4, N // 8))
    got = kwhip.kspaceFirstOrder3D(kg, med, src, sens, **opts)
    kgo = ko.KGrid(N, kg.dx, N, kg.dy, N, kg.dz)
    kgo.set_time(Nt, kg.dt)
    res = {}
    for cast in ("off", "single"):
        ref = ko.kspace_first_order(kgo, med, src, sens, data_cast=cast, pml_size=opts["PMLSize"],
                                    pml_alpha=opts["PMLAlpha"])
        res["fp64" if cast == "off" else "fp32"] = max(ko.rel_linf(got[k], ref[k]) for k in ("p_max", "p_rms"))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(independent simulations shard across GPUs without communication)", file=sys.stderr)
        sys.exit(2)
    import kwhip
    from kwhip import workloads

    N, K, W = args.N, args.steps, args.warmup
    Nt = max(1000, K)
    log(f"[bench] building the {N}^3 workload")
    kg, med, src, sens, opts = workloads.cfg2(N, Nt=Nt)
    ncell = N ** 3
    n_src_on = src["p"].shape[1]

    if W > 0:
        sim = kwhip.Simulation(kg, med, src, sens, 3, **opts)
        sim.step(min(W, Nt))
        sim.sync()
        sim.close()
        log(f"[bench] warm-up {W} steps done")

    sim = kwhip.Simulation(kg, med, src, sens, 3, **opts)
    dev_gb = sim.device_bytes / 1e9
    sim.sync()
    t0 = time.perf_counter()
    ev_ms, _ = sim.step_timed(K)
    sim.sync()
    wall = time.perf_counter() - t0
    sim.close()
    seconds = max(wall, ev_ms / 1e3)
    ms_per_step = seconds * 1e3 / K
    mcps = ncell * K / seconds / 1e6
    log(f"[bench] {K} steps in {seconds:.3f} s (events {ev_ms / 1e3:.3f} s): "
        f"{ms_per_step:.3f} ms/step, {mcps:.1f} Mcell-steps/s")

    on = min(K, n_src_on)
    b_alg = (on * B_ALG_SRC_ON + (K - on) * B_ALG_SRC_OFF) / K
    step_frac = b_alg * mcps * 1e6 / (HBM_PEAK_GBS * 1e9)

    roofline = None
    kernels = None
    if not args.no_profile:
        sim = kwhip.Simulation(kg, med, src, sens, 3, **opts)
        _, kern = sim.step_timed(K, per_kernel=True)
        sim.close()
        kernels = {k: {"ms_total": round(v["ms"], 4), "launches": v["launches"],
                       "avg_us": round(1e3 * v["ms"] / v["launches"], 3) if v["launches"] else None,
                       "GBps": round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None}
                   for k, v in kern.items() if v["launches"]}
        dom = max(kern, key=lambda k: kern[k]["ms"])
        d = kern[dom]
        avg_s = d["ms"] * 1e-3 / d["launches"]
        bytes_per_launch = d["bytes"] / d["launches"]
        achieved = bytes_per_launch / avg_s / 1e9
        traffic = None
        tf = ROOT / "profiles" / f"traffic_{dom}_cfg2_{N}.json"
        if tf.exists():
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "bytes_per_launch": bytes_per_launch,
                    "avg_launch_us": round(avg_s * 1e6, 3)}
        log(f"[bench] dominant kernel {dom}: {achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS:.1%} of peak)")

    cpu = None
    if not args.no_cpu:
        cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        log(f"[bench] CPU baseline: oracle float32, {args.cpu_steps} steps, {cores} threads")
        cval, step_s, n = cpu_baseline(N, args.cpu_steps, cores)
        cpu = {"value": round(cval, 3), "unit": "Mcell-steps/s", "cores": cores, "kind": "port",
               "sample": f"{N}^3 workload, first {n} steps (source on), median of steps 3..{n} "
                         f"({step_s:.2f} s/step), NumPy/scipy.fft float32 restatement"}

    acc = None
    if not args.no_accuracy:
        log("[bench] accuracy: 64^3 x 300 steps of the same recipe vs the FP64/FP32 oracle")
        acc = accuracy_check()

    line = {
        "metric": METRIC, "value": round(mcps, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"kspaceFirstOrder3D {N}^3 heterogeneous c/rho, power-law y=1.5, "
                               f"additive tone-burst disc source, full-grid p_max+p_rms, "
                               f"{K} steps from t=1 (source on for {on})",
                   "grid": [N, N, N], "parallelism": "single GPU"},
        "roofline": roofline,
        "step_roofline": {"b_alg_per_cell_step": round(b_alg, 1), "frac": round(step_frac, 4),
                          "model": "per-axis-pass algorithmic bytes / 8 TB/s"},
        "cpu_baseline": cpu,
        "rel_err_64cubed_300steps": acc,
        "device_gb": round(dev_gb, 3),
        "kernels": kernels,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
