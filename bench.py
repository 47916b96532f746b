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


def _tr_sensor(sens, data):
    return {"mask": sens["mask"], "time_reversal_boundary_data": data}


def batched_bench(n_problems, seed0=0):
    """Photoacoustic time reversal of n problems of uneven size (2D 256^2 -
    1024^2 and 3D 64^3 - 192^3): each problem is a forward run recording p
    on one face and a time-reversal run re-injecting it.  The engine runs
    all of them together (a problem's reconstruction is submitted when its
    forward run returns); the plain path runs the same simulations one after
    another with kspaceFirstOrder2D/3D.  Inputs are generated before the
    timed regions; both timed regions include host pre-processing,
    uploads, the simulations and the read-back of every result."""
    import numpy as np

    import kwhip
    from kwhip import workloads

    probs = [workloads.batch_tr_problem(seed0 + i) for i in range(n_problems)]
    cell_steps = float(sum(kg.total_grid_points * (2 * kg.Nt - 1) for kg, *_ in probs))
    fn = {2: kwhip.kspaceFirstOrder2D, 3: kwhip.kspaceFirstOrder3D}

    t0 = time.perf_counter()
    with kwhip.Engine() as eng:
        fwd = [eng.submit(kg, med, src, sens, kg.dim, **opts) for kg, med, src, sens, opts in probs]
        # each reconstruction is submitted as soon as its forward run is done
        trs = [None] * len(probs)
        left = list(range(len(probs)))
        while left:
            done = [i for i in left if eng.poll(fwd[i])["status"] != "running"
                    and eng.poll(fwd[i])["status"] != "queued"]
            for i in done:
                kg, med, src, sens, opts = probs[i]
                data = eng.wait(fwd[i])
                trs[i] = eng.submit(kg, med, {}, _tr_sensor(sens, data), kg.dim, **opts)
                left.remove(i)
            if not done:
                time.sleep(0.002)
        rec_eng = [eng.wait(t) for t in trs]
        st = eng.stats()
    t_eng = time.perf_counter() - t0
    log(f"[bench] batched: {n_problems} problems (forward + time reversal) in {t_eng:.2f} s, "
        f"{cell_steps / t_eng / 1e6:.0f} Mcell-steps/s, {st['launches']} launches")

    t0 = time.perf_counter()
    rec_plain = []
    for kg, med, src, sens, opts in probs:
        data = fn[kg.dim](kg, med, src, sens, **opts)
        rec_plain.append(fn[kg.dim](kg, med, {}, _tr_sensor(sens, data), **opts))
    t_plain = time.perf_counter() - t0
    log(f"[bench] plain: the same simulations one by one in {t_plain:.2f} s, "
        f"{cell_steps / t_plain / 1e6:.0f} Mcell-steps/s")
    diff = max(float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))
               for a, b in zip(rec_eng, rec_plain))
    n2 = sum(1 for kg, *_ in probs if kg.dim == 2)
    return {"value": round(cell_steps / t_eng / 1e6, 1), "unit": "Mcell-steps/s",
            "plain_value": round(cell_steps / t_plain / 1e6, 1),
            "speedup_vs_plain": round(t_plain / t_eng, 2),
            "seconds": round(t_eng, 3), "plain_seconds": round(t_plain, 3),
            "rel_linf_vs_plain": diff,
            "sample": f"{n_problems} problems (seeds {seed0}..{seed0 + n_problems - 1}: {n2} 2D, "
                      f"{n_problems - n2} 3D), forward + time reversal each, "
                      f"{cell_steps / 1e9:.1f} G cell-steps"}


def _served_requests(n_req, steps_lo, steps_hi, seed):
    import numpy as np

    from kwhip import workloads

    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0, size=n_req)  # arrival gaps at 1 request/s
    steps = rng.integers(steps_lo, steps_hi + 1, size=n_req)
    reqs = [workloads.served_problem(seed * 1000 + i, int(steps[i])) for i in range(n_req)]
    return reqs, gaps


def _served_engine(reqs, arrivals):
    """Completion times (seconds after the first arrival's origin) of the
    requests served by the engine.  Each request is submitted by a handler
    thread of its own at its arrival time, as a server's connection
    handlers would (the engine is thread-safe), so the host-side
    pre-processing of one request does not delay the next."""
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    import kwhip

    n_req = len(reqs)
    with kwhip.Engine() as eng, ThreadPoolExecutor(max_workers=8) as handlers:
        t0 = time.perf_counter()

        def handle(i):
            kg, med, src, sens, opts = reqs[i]
            t = eng.submit(kg, med, src, sens, kg.dim, **opts)
            # wall clock minus engine clock at submission
            return time.perf_counter() - t0 - eng.poll(t)["submit_s"], t

        futures = []
        for i in range(n_req):
            delay = t0 + arrivals[i] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            futures.append(handlers.submit(handle, i))
        tickets = [f.result() for f in futures]
        done_at = np.zeros(n_req)
        for i, (offset, t) in enumerate(tickets):
            _, info = eng.wait(t, with_info=True)
            done_at[i] = info["done_s"] + offset
    return done_at


def _served_plain_service(reqs):
    """Service time of each request run alone through kspaceFirstOrder2D/3D
    (host buffers in and out), the plain per-simulation GPU path."""
    import numpy as np

    import kwhip

    fn = {2: kwhip.kspaceFirstOrder2D, 3: kwhip.kspaceFirstOrder3D}
    service = np.zeros(len(reqs))
    for i, (kg, med, src, sens, opts) in enumerate(reqs):
        t1 = time.perf_counter()
        fn[kg.dim](kg, med, src, sens, **opts)
        service[i] = time.perf_counter() - t1
    return service


def _served_summary(reqs, arrivals, done_at, service, rate, steps_lo, steps_hi, seed):
    import numpy as np

    n_req = len(reqs)
    cell_steps = float(sum(kg.total_grid_points * kg.Nt for kg, *_ in reqs))
    lat = done_at - arrivals
    eng_tp = cell_steps / float(np.max(done_at)) / 1e6
    # the plain server runs the same requests one at a time in arrival order
    finish = np.zeros(n_req)
    tfree = 0.0
    for i in range(n_req):
        tfree = max(tfree, arrivals[i]) + service[i]
        finish[i] = tfree
    lat_plain = finish - arrivals
    plain_tp = cell_steps / float(np.max(finish)) / 1e6
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)  # noqa: E731
    log(f"[bench] served: {n_req} requests at {rate}/s: p50 {pct(lat, 50)} s p99 {pct(lat, 99)} s, "
        f"{eng_tp:.0f} Mcell-steps/s; one at a time: p50 {pct(lat_plain, 50)} s "
        f"p99 {pct(lat_plain, 99)} s, {plain_tp:.0f} Mcell-steps/s")
    return {"rate_per_s": rate, "value": round(eng_tp, 1), "unit": "Mcell-steps/s",
            "p50_s": pct(lat, 50), "p99_s": pct(lat, 99),
            "plain_value": round(plain_tp, 1), "plain_p50_s": pct(lat_plain, 50),
            "plain_p99_s": pct(lat_plain, 99),
            "speedup_vs_plain": round(eng_tp / plain_tp, 2),
            "sample": f"{n_req} requests, Poisson at {rate}/s, {steps_lo}-{steps_hi} steps each "
                      f"(seed {seed}), {cell_steps / 1e9:.1f} G cell-steps"}


def served_bench(n_req, rate, steps_lo, steps_hi, seed=7):
    """A Poisson request stream of forward simulations (2D 256^2 - 1024^2,
    3D 96^3 - 256^3, heterogeneous absorbing nonlinear medium, tone-burst
    disc source, whole-grid p_max/p_rms): latency from arrival to result and
    throughput (cell-steps over the time to the last result) for the engine,
    against a plain server that runs the same requests one at a time in
    arrival order (its latencies follow from the measured service time of
    each request and the same arrival times)."""
    import numpy as np

    reqs, gaps = _served_requests(n_req, steps_lo, steps_hi, seed)
    arrivals = np.cumsum(gaps / rate)
    done_at = _served_engine(reqs, arrivals)
    service = _served_plain_service(reqs)
    return _served_summary(reqs, arrivals, done_at, service, rate, steps_lo, steps_hi, seed)


def served_sweep(rates, n_req, steps_lo, steps_hi, seed=7):
    """The served workload at several arrival rates (the same requests in the
    same order; the plain server's service times are measured once)."""
    import numpy as np

    reqs, gaps = _served_requests(n_req, steps_lo, steps_hi, seed)
    service = _served_plain_service(reqs)
    out = []
    for rate in rates:
        arrivals = np.cumsum(gaps / rate)
        done_at = _served_engine(reqs, arrivals)
        out.append(_served_summary(reqs, arrivals, done_at, service, rate, steps_lo, steps_hi,
                                   seed))
    return out


COPY_BYTES = 1 << 30  # calibration copy: 1 GiB read + 1 GiB written per dispatch


def copy_calibration():
    """A plain device-to-device copy of COPY_BYTES (torch's copy kernel), five
    times: the reference dispatch for scaling the FETCH_SIZE / WRITE_SIZE
    counters to bytes."""
    import torch

    x = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: 5 copies of {COPY_BYTES} bytes")


def _counter_means(run_dir, counter, name_part):
    """Mean per-dispatch value of a counter over the dispatches whose kernel
    name contains name_part, from a rocprofv3 run directory (CSV output or
    the rocpd SQLite database)."""
    import csv
    import sqlite3

    per = {}
    for f in Path(run_dir).rglob("*counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            if r.get("Counter_Name") == counter and name_part in r.get("Kernel_Name", ""):
                key = (str(f), r.get("Dispatch_Id"))
                per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    for f in Path(run_dir).rglob("*.db"):
        con = sqlite3.connect(str(f))
        rows = con.execute("select dispatch_id, name, counter_value from pmc_events "
                           "where counter_name = ?", (counter,)).fetchall()
        for dispatch, name, value in rows:
            if name_part in name:
                key = (str(f), dispatch)
                per[key] = per.get(key, 0.0) + float(value)
    if not per:
        raise SystemExit(f"no {counter} values for kernels matching {name_part!r} in {run_dir}")
    return sum(per.values()) / len(per), len(per)


def reduce_traffic(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, kernel, variant, N):
    """HBM bytes per launch of one kernel from FETCH_SIZE and WRITE_SIZE runs
    (KiB per dispatch), each scaled by factor = known bytes / counter bytes of
    the calibration copy; written to profiles/traffic_<kernel>_cfg2_<N>.json,
    which the bench reports as roofline.traffic."""
    # "(" ends the kernel name: the record variant (<variant>_rec) is not it
    f_kib, nf = _counter_means(fetch_dir, "FETCH_SIZE", variant + "(")
    w_kib, nw = _counter_means(write_dir, "WRITE_SIZE", variant + "(")
    # the calibration copy runs as the runtime's copy kernel (copyBuffer)
    cf_kib, _ = _counter_means(cal_fetch_dir, "FETCH_SIZE", "copyBuffer")
    cw_kib, _ = _counter_means(cal_write_dir, "WRITE_SIZE", "copyBuffer")
    f_scale = COPY_BYTES / (cf_kib * 1024.0)
    w_scale = COPY_BYTES / (cw_kib * 1024.0)
    fetch_b = f_kib * 1024.0 * f_scale
    write_b = w_kib * 1024.0 * w_scale
    out = {"kernel": kernel, "variant": variant, "N": N,
           "bytes_per_launch": fetch_b + write_b, "fetch_bytes": fetch_b, "write_bytes": write_b,
           "fetch_kib_raw": f_kib, "write_kib_raw": w_kib, "dispatches": [nf, nw],
           "copy_fetch_kib": cf_kib, "copy_write_kib": cw_kib,
           "fetch_factor": f_scale, "write_factor": w_scale}
    path = ROOT / "profiles" / f"traffic_{kernel}_cfg2_{N}.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


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
    ap.add_argument("--batched", type=int, default=24,
                    help="problems in the batched time-reversal leg (0: skip)")
    ap.add_argument("--served", type=int, default=24,
                    help="requests in the served leg (0: skip)")
    ap.add_argument("--served-rate", type=float, default=10.0)
    ap.add_argument("--served-sweep", type=float, nargs="+", metavar="RATE",
                    help="run only the served workload at these arrival rates (one JSON line)")
    ap.add_argument("--served-steps", type=int, nargs=2, default=[200, 600])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only the calibration copy (under rocprofv3 --pmc)")
    ap.add_argument("--reduce-traffic", nargs=4, metavar="DIR",
                    help="FETCH_SIZE run, WRITE_SIZE run, and the two calibration runs")
    ap.add_argument("--traffic-kernel", default="spectral")
    ap.add_argument("--traffic-variant", default="kv_2_256")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.reduce_traffic:
        reduce_traffic(*args.reduce_traffic, args.traffic_kernel, args.traffic_variant, args.N)
        return
    if args.served_sweep:
        n = args.served if args.served > 0 else 24
        print(json.dumps({"served_sweep": served_sweep(args.served_sweep, n, *args.served_steps)}),
              flush=True)
        return
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

    batched = batched_bench(args.batched) if args.batched > 0 else None
    served = (served_bench(args.served, args.served_rate, *args.served_steps)
              if args.served > 0 else None)

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
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
