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
# model: 14 FFTs x 3 axis passes x 8 B, fused elementwise operands, FP64
# p_max and p_rms read+write): 488 B; while the source is on the step adds
# the 4 B read of the precomputed source field K_M = ifftn(source_kappa .*
# fftn(mask)) (the two FFTs per step it replaces are an algorithmic saving,
# not counted)
B_ALG_SRC_ON, B_ALG_SRC_OFF = 492.0, 488.0
# The same model per kernel kind, bytes per cell over one step (sum of the
# kind's launches): 8 B per cell for every FFT axis pass a kind performs, 4 B
# per FP32 grid operand and 8 B per FP64 one it must read or write.
#   y_fwd    : y passes of 1 + 3 + 2 forward transforms            48
#   y_inv    : y passes of 3 + 3 + 2 inverse transforms            64
#   spectral : z passes of 4 + 6 + 4 transforms                   112
#   x_grad   : 3 inverse + 3 forward x passes; u r/w, 1/rho0_sg    84
#   x_div    : 3 inverse + 2 forward x passes; rho r/w, rho0,
#              sum(rho) write (+ K_M while the source is on)     72 (+4)
#   x_eos    : 2 inverse + 1 forward x passes; sum(rho), c0, tau,
#              eta; p_max and p_rms r/w in FP64                   72
# (452 B source off: the kinds omit the intermediate p write/read and a few
# operands that the 488 B per-step total counts.)
KIND_B_STEP = {"y_fwd": 48.0, "y_inv": 64.0, "spectral": 112.0, "x_grad": 84.0,
               "x_div": 72.0, "x_eos": 72.0}
KIND_B_SRC_ON = {"x_div": 4.0}
METRIC = "Mcell-steps/s 256³ FP32 vs plain GPU; served p50/p99 latency at ≤1e-4 rel err"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_phase(N, nsteps, cores, cast, source_on):
    """Median wall time of one step of the oracle (scipy.fft workers =
    cores) over steps 3..2+nsteps of the 256^3 workload, with the source on
    throughout (source_on) or switched off after step 1 (a one-sample
    signal): the two regimes of the 1000-step run."""
    from kwhip import workloads
    from oracle import kwave_oracle as ko

    Nt = 2 + nsteps
    kg, med, src, sens, opts = workloads.disc_source_3d(N, Nt=Nt)
    if not source_on:
        src = dict(src, p=src["p"][:, :1])
    kgo = ko.KGrid(N, kg.dx, N, kg.dy, N, kg.dz)
    kgo.set_time(Nt, kg.dt)
    times = []
    ko.kspace_first_order(kgo, med, src, sens, data_cast=cast, pml_size=opts["PMLSize"],
                          pml_alpha=opts["PMLAlpha"], fft_workers=cores, step_times=times)
    tail = sorted(times[2:])
    return tail[len(tail) // 2]


def cpu_baseline(N, nsteps, cores, casts=("single",), n_on=534, n_total=1000):
    """CPU restatement of the reference algorithm (oracle/) on the host
    cores: per precision, the step time of each regime (2 warm-up + nsteps
    timed steps each), weighted by the 1000-step run's n_on source-on and
    n_total - n_on source-off steps; Mcell-steps/s of the weighted step."""
    out = {}
    for cast in casts:
        t_on = cpu_phase(N, nsteps, cores, cast, True)
        t_off = cpu_phase(N, nsteps, cores, cast, False)
        t = (n_on * t_on + (n_total - n_on) * t_off) / n_total
        out["fp32" if cast == "single" else "fp64"] = {
            "value": N ** 3 / t / 1e6, "s_per_step_on": t_on, "s_per_step_off": t_off}
    return out


def accuracy_check(N=64, Nt=300):
    """The disc-source recipe at N^3 for Nt steps: GPU vs the FP64 and FP32 oracle
    (relative L-inf of p_max and p_rms)."""
    import kwhip
    from kwhip import workloads
    from oracle import kwave_oracle as ko

    kg, med, src, sens, opts = workloads.disc_source_3d(N, Nt=Nt, pml=10, src_radius=max(4, N // 8))
    got = kwhip.kspaceFirstOrder3D(kg, med, src, sens, **opts)
    kgo = ko.KGrid(N, kg.dx, N, kg.dy, N, kg.dz)
    kgo.set_time(Nt, kg.dt)
    res = {}
    for cast in ("off", "single"):
        ref = ko.kspace_first_order(kgo, med, src, sens, data_cast=cast, pml_size=opts["PMLSize"],
                                    pml_alpha=opts["PMLAlpha"])
        res["fp64" if cast == "off" else "fp32"] = max(ko.rel_linf(got[k], ref[k]) for k in ("p_max", "p_rms"))
    return res


def circle_2d_leg():
    """The 2-D 128^2 lossless workload (Cartesian circle of 50 points,
    906 steps) in full: the GPU against the CPU restatement of the reference
    algorithm in double precision (the reference's default), both timed, and
    the GPU result's relative L-inf error against it."""
    import kwhip
    from kwhip import workloads
    from oracle import kwave_oracle as ko

    kg, med, src, sens, opts = workloads.circle_2d()
    kwhip.kspaceFirstOrder2D(kg, med, src, sens, **opts)  # warm-up (host set-up paths)
    t0 = time.perf_counter()
    got = kwhip.kspaceFirstOrder2D(kg, med, src, sens, **opts)
    t_gpu = time.perf_counter() - t0
    kgo = ko.KGrid(kg.Nx, kg.dx, kg.Ny, kg.dy)
    kgo.set_time(kg.Nt, kg.dt)
    t0 = time.perf_counter()
    ref = ko.kspace_first_order(kgo, med, src, sens, data_cast="off", pml_size=opts["PMLSize"],
                                pml_alpha=opts["PMLAlpha"], cart_interp=opts["CartInterp"])
    t_cpu = time.perf_counter() - t0
    cells = kg.Nx * kg.Ny * kg.Nt
    return {"steps": int(kg.Nt), "gpu_s": round(t_gpu, 3), "cpu_fp64_s": round(t_cpu, 3),
            "gpu_Mcell_steps_s": round(cells / t_gpu / 1e6, 1),
            "cpu_fp64_Mcell_steps_s": round(cells / t_cpu / 1e6, 2),
            "rel_linf_vs_fp64": ko.rel_linf(got, ref),
            "note": "whole call (host set-up, transfers, 906 steps) on both sides"}


def axisym_leg(Nx=512, Ny=256, Nt=2000, cpu_steps=20, check_steps=100):
    """kspaceFirstOrderAS on an Nx x Ny radial grid (the GPU steps the
    4 Ny expanded grid of 'WSWA-FFT'): heterogeneous c/rho, Stokes
    absorption, B/A = 6, a focused p0 on the axis, a binary sensor line
    recording p and p_max.  The whole call (host set-up, transfers, Nt
    steps) on the GPU; the CPU restatement of kspaceFirstOrderAS.m in single
    precision for cpu_steps steps; the relative L-inf difference of p_final
    between the two after check_steps steps."""
    import numpy as np

    import kwhip
    from oracle import kwave_oracle as ko

    dx = 1e-4
    rng = np.random.default_rng(11)
    x = (np.arange(Nx) - Nx // 2) * dx
    r = np.arange(Ny) * dx
    X, R = np.meshgrid(x, r, indexing="ij")
    c = 1500 + 60 * np.tanh((X + 8e-3) / 2e-3) + 5 * rng.standard_normal((Nx, Ny))
    rho = 1000 + 40 * np.tanh((X + 8e-3) / 2e-3)
    med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "BonA": 6.0}
    p0 = 2e5 * np.exp(-((X + 1e-2) ** 2 + R ** 2) / (2e-3) ** 2)
    mask = np.zeros((Nx, Ny), bool)
    mask[Nx // 2 + 40, :Ny - 30] = True
    dt = 0.2 * dx / float(c.max())

    def run(nt, rec):
        kg = kwhip.kWaveGrid(Nx, dx, Ny, dx)
        kg.setTime(nt, dt)
        return kwhip.kspaceFirstOrderAS(kg, med, {"p0": p0}, {"mask": mask, "record": rec},
                                        PMLSize=[20, 20], Smooth=False)

    run(10, ["p"])  # warm-up (host set-up paths)
    t0 = time.perf_counter()
    run(Nt, ["p", "p_max"])
    t_gpu = time.perf_counter() - t0
    got = run(check_steps, ["p_final"])["p_final"]
    kgo = ko.KGrid(Nx, dx, Ny, dx)
    kgo.set_time(check_steps, dt)
    ref = ko.kspace_first_order_as(kgo, med, {"p0": p0}, {"mask": mask, "record": ["p_final"]},
                                   data_cast="single", pml_size=[20, 20], smooth=False)["p_final"]
    kgc = ko.KGrid(Nx, dx, Ny, dx)
    kgc.set_time(cpu_steps, dt)
    t0 = time.perf_counter()
    ko.kspace_first_order_as(kgc, med, {"p0": p0}, {"mask": mask, "record": ["p"]},
                             data_cast="single", pml_size=[20, 20], smooth=False)
    t_cpu = (time.perf_counter() - t0) / cpu_steps
    cells = Nx * Ny
    return {"grid": [Nx, Ny], "expanded": [Nx, 4 * Ny], "steps": Nt,
            "gpu_s": round(t_gpu, 3), "gpu_ms_per_step": round(1e3 * t_gpu / Nt, 4),
            "gpu_Mcell_steps_s": round(cells * Nt / t_gpu / 1e6, 1),
            "cpu_fp32_ms_per_step": round(1e3 * t_cpu, 2),
            "cpu_fp32_Mcell_steps_s": round(cells / t_cpu / 1e6, 2),
            f"rel_linf_p_final_{check_steps}_steps_vs_fp32": ko.rel_linf(got, ref),
            "note": "radial cells; GPU whole call (host set-up, transfers, steps); CPU: "
                    f"the oracle's kspaceFirstOrderAS restatement, {cpu_steps} steps"}


def nonlinear_bowl_leg(N=384, Nt=2000, steps=None):
    """The nonlinear precision-study workload (B/A = 6, power-law absorption,
    1 MPa 0.5 MHz bowl source, 64 x 64 sensor plane recording p, p_max and
    p_rms): throughput of the whole run at N^3 on the GPU (inputs resident),
    and its error against the oracle from the committed 2000-step fixtures
    of the same recipe at smaller sizes (tests/golden/nonlinear_bowl_<n>.npz: single-
    and double-precision oracle runs, p on a subsampled set of plane points
    and steps, p_max and p_rms on the whole plane)."""
    import numpy as np

    import kwhip
    from kwhip import workloads

    steps = Nt if steps is None else steps
    kg, med, src, sens, opts, zf = workloads.nonlinear_bowl_3d(N, Nt=Nt)
    sim = kwhip.Simulation(kg, med, src, sens, 3, **opts)
    sim.sync()
    t0 = time.perf_counter()
    ev_ms, _ = sim.step_timed(steps)
    sim.sync()
    seconds = max(time.perf_counter() - t0, ev_ms / 1e3)
    out = {"grid": [N, N, N], "steps": steps, "ms_per_step": round(seconds * 1e3 / steps, 4),
           "value": round(N ** 3 * steps / seconds / 1e6, 1), "unit": "Mcell-steps/s",
           "device_gb": round(sim.device_bytes / 1e9, 3)}
    if steps == Nt:
        res = sim.read()
        pk = float(np.max(np.abs(res["p"])))
        out["p_peak_pa"] = round(pk, 1)
    sim.close()
    sys.path.insert(0, str(ROOT / "tests" / "golden"))
    import make_longrun as ml
    errs = {}
    for n in (96, 192):
        path = ROOT / "tests" / "golden" / f"nonlinear_bowl_{n}.npz"
        if not path.exists():
            continue
        fx = np.load(path)
        kgs, meds, srcs, senss, optss = ml.bowl_inputs(n)
        got = kwhip.kspaceFirstOrder3D(kgs, meds, srcs, senss, **optss)
        gp = ml.bowl_p_sub(got["p"]).astype(np.float64)
        e = {}
        for tag in ("fp32", "fp64"):
            r = fx[f"{tag}_p"].astype(np.float64)
            e[f"p_linf_vs_{tag}"] = float(np.max(np.abs(gp - r)) / float(fx[f"{tag}_p_absmax"]))
            e[f"p_l2_vs_{tag}"] = float(np.sqrt(np.sum((gp - r) ** 2) / np.sum(r ** 2)))
            for k in ("p_max", "p_rms"):
                g = np.asarray(got[k], dtype=np.float64).ravel()
                rr = fx[f"{tag}_{k}"].astype(np.float64)
                e[f"{k}_linf_vs_{tag}"] = float(np.max(np.abs(g - rr)) / np.max(np.abs(rr)))
        errs[f"{n}^3"] = e
    out["rel_err_2000_steps"] = errs
    return out


def b_alg_cell_step(dim, absorbing, hetero, nonlinear, full_stats):
    """Algorithmic bytes per cell-step of one simulation under the per-axis-
    pass model: 8 B per cell for every FFT axis pass (gradient 1 forward +
    dim inverse transforms, divergence dim + dim, absorbing equation of state
    2 + 2), 8 B for each velocity and density component read and written,
    8 B for the pressure the equation of state writes and the next gradient
    reads, 4 B per FP32 medium grid (1/rho0_sg_d, rho0, c0; with absorption
    tau, eta and sum(rho) written and read; B/A), and 16 B for each FP64
    whole-grid statistic read and written.  2-D lossless homogeneous: 152 B;
    3-D: 296 B."""
    n_fft = (1 + dim) + 2 * dim + (4 if absorbing else 0)
    b = n_fft * dim * 8.0 + 2 * dim * 8.0 + 8.0
    if hetero:
        b += 4.0 * (dim + 2)
    if absorbing:
        b += 8.0 + (8.0 if hetero else 0.0)
    if nonlinear and hetero:
        b += 4.0
    return b + 16.0 * full_stats


def _problem_bytes(kg, med, sens, steps):
    """Algorithmic bytes of `steps` steps of one problem of the batched or
    served workloads (b_alg_cell_step)."""
    import numpy as np
    hetero = np.ndim(med.get("sound_speed", 0.0)) > 0 or np.ndim(med.get("density", 0.0)) > 0
    absorbing = "alpha_coeff" in med
    nonlinear = "BonA" in med and np.ndim(med["BonA"]) > 0
    full = sum(1 for r in sens.get("record", []) if r in ("p_max", "p_min", "p_rms")) \
        if np.size(sens.get("mask", 0)) == kg.total_grid_points else 0
    return kg.total_grid_points * steps * b_alg_cell_step(kg.dim, absorbing, hetero, nonlinear, full)


def _roofline(nbytes, seconds):
    return round(nbytes / seconds / (HBM_PEAK_GBS * 1e9), 4)


def _pretune(probs):
    """Run the tile autotuner once per grid shape and equation of state of a
    workload before its timed regions (the engine's admission only reads the
    tuner's cache, and the plain path would otherwise measure inside its
    timed region); returns the seconds it took."""
    import kwhip

    t0 = time.perf_counter()
    seen = set()
    for kg, med, src, sens, opts in probs:
        key = (kg.dim, kg.Nx, kg.Ny, kg.Nz, "alpha_coeff" in med, "BonA" in med)
        if key in seen:
            continue
        seen.add(key)
        kwhip.launch_tiles(kg, med, src, sens, kg.dim, **opts)
    dt = time.perf_counter() - t0
    log(f"[bench] tile autotuner: {len(seen)} grid shapes in {dt:.2f} s")
    return dt


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
    tune_s = _pretune(probs)
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
    nbytes = sum(_problem_bytes(kg, med, sens, 2 * kg.Nt - 1) for kg, med, src, sens, _ in probs)
    return {"value": round(cell_steps / t_eng / 1e6, 1), "unit": "Mcell-steps/s",
            "plain_value": round(cell_steps / t_plain / 1e6, 1),
            "speedup_vs_plain": round(t_plain / t_eng, 2),
            "seconds": round(t_eng, 3), "plain_seconds": round(t_plain, 3),
            "roofline_frac": _roofline(nbytes, t_eng),
            "roofline_frac_busy": _roofline(nbytes, st["busy_s"]),
            "plain_roofline_frac": _roofline(nbytes, t_plain),
            "alg_bytes": nbytes,
            "tune_s": round(tune_s, 2),
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


def _served_steps(kg, sens):
    """Time steps of a served request (a reconstruction runs Nt - 1)."""
    return kg.Nt - 1 if "time_reversal_boundary_data" in sens else kg.Nt


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
    n_tr = sum(1 for r in reqs if "time_reversal_boundary_data" in r[3])
    cell_steps = float(sum(kg.total_grid_points * _served_steps(kg, sens)
                           for kg, med, src, sens, _ in reqs))
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
    nbytes = sum(_problem_bytes(kg, med, sens, _served_steps(kg, sens))
                 for kg, med, src, sens, _ in reqs)
    pct = lambda a, q: round(float(np.percentile(a, q)), 3)  # noqa: E731
    log(f"[bench] served: {n_req} requests at {rate}/s: p50 {pct(lat, 50)} s p99 {pct(lat, 99)} s, "
        f"{eng_tp:.0f} Mcell-steps/s; one at a time: p50 {pct(lat_plain, 50)} s "
        f"p99 {pct(lat_plain, 99)} s, {plain_tp:.0f} Mcell-steps/s")
    return {"rate_per_s": rate, "value": round(eng_tp, 1), "unit": "Mcell-steps/s",
            "p50_s": pct(lat, 50), "p99_s": pct(lat, 99),
            "plain_value": round(plain_tp, 1), "plain_p50_s": pct(lat_plain, 50),
            "plain_p99_s": pct(lat_plain, 99),
            "speedup_vs_plain": round(eng_tp / plain_tp, 2),
            "roofline_frac": _roofline(nbytes, float(np.max(done_at))),
            "plain_roofline_frac": _roofline(nbytes, float(np.max(finish))),
            "sample": f"{n_req} requests ({n_tr} time-reversal reconstructions), Poisson at "
                      f"{rate}/s, {steps_lo}-{steps_hi} steps each "
                      f"(seed {seed}), {cell_steps / 1e9:.1f} G cell-steps"}


def served_bench(n_req, rate, steps_lo, steps_hi, seed=7):
    """A Poisson request stream of forward simulations (2D 256^2 - 1024^2,
    3D 96^3 - 256^3, heterogeneous absorbing nonlinear medium, tone-burst
    disc source, whole-grid p_max/p_rms) and time-reversal reconstructions
    (the disc re-emitting delayed bursts): latency from arrival to result and
    throughput (cell-steps over the time to the last result) for the engine,
    against a plain server that runs the same requests one at a time in
    arrival order (its latencies follow from the measured service time of
    each request and the same arrival times)."""
    import numpy as np

    reqs, gaps = _served_requests(n_req, steps_lo, steps_hi, seed)
    _pretune(reqs)
    arrivals = np.cumsum(gaps / rate)
    done_at = _served_engine(reqs, arrivals)
    service = _served_plain_service(reqs)
    return _served_summary(reqs, arrivals, done_at, service, rate, steps_lo, steps_hi, seed)


def served_sweep(rates, n_req, steps_lo, steps_hi, seed=7):
    """The served workload at several arrival rates (the same requests in the
    same order; the plain server's service times are measured once)."""
    import numpy as np

    reqs, gaps = _served_requests(n_req, steps_lo, steps_hi, seed)
    _pretune(reqs)
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
    the calibration copy; written to profiles/traffic_<kernel>_disc_source_<N>.json,
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
    path = ROOT / "profiles" / f"traffic_{kernel}_disc_source_{N}.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out))


COUNTER_KERNELS = (("kv_0_256", "y_fwd"), ("kv_1_256", "y_inv"), ("kv_2_256", "spectral"),
                   ("kv_5_256", "x_grad"), ("kv_6_256", "x_div"), ("kv_7_256", "x_eos"))


def _all_counter_means(run_dir, name_part):
    """Mean per-dispatch value of every counter of a rocprofv3 --pmc run (CSV
    output) over the dispatches of the kernels whose name contains
    name_part."""
    import csv

    per = {}
    for f in Path(run_dir).rglob("*counter_collection.csv"):
        for r in csv.DictReader(open(f)):
            if name_part not in r.get("Kernel_Name", ""):
                continue
            key = (r["Counter_Name"], str(f), r.get("Dispatch_Id"))
            per[key] = per.get(key, 0.0) + float(r["Counter_Value"])
    sums, counts = {}, {}
    for (c, _, _), v in per.items():
        sums[c] = sums.get(c, 0.0) + v
        counts[c] = counts.get(c, 0) + 1
    return {c: sums[c] / counts[c] for c in sums}


def reduce_counters(dirs):
    """Markdown table of the instruction, wait and LDS counters of the step
    kernels from several rocprofv3 --pmc runs (one counter set per run), with
    the ratios that locate where a kernel's time goes: the share of its
    wave-cycles spent waiting for anything, for LDS and for vector memory,
    the share issuing VALU and LDS instructions, the LDS bank-conflict rate,
    and the waves in flight per sampled cycle."""
    rows = []
    for variant, kind in COUNTER_KERNELS:
        c = {}
        for d in dirs:
            c.update(_all_counter_means(d, variant + "("))
        if not c:
            continue

        def ratio(a, b):
            return c[a] / c[b] if a in c and b in c and c[b] else None

        rows.append((kind, variant, {
            "waves": c.get("SQ_WAVES"),
            "wait_any": ratio("SQ_WAIT_ANY", "SQ_WAVE_CYCLES"),
            "wait_inst_any": ratio("SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES"),
            "wait_lds": ratio("SQ_WAIT_INST_LDS", "SQ_WAVE_CYCLES"),
            "active_valu": ratio("SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES"),
            "active_lds": ratio("SQ_ACTIVE_INST_LDS", "SQ_WAVE_CYCLES"),
            "active_vmem": ratio("SQ_ACTIVE_INST_VMEM", "SQ_WAVE_CYCLES"),
            "lds_bank_conflict": ratio("SQ_LDS_BANK_CONFLICT", "SQ_LDS_IDX_ACTIVE"),
            # waves in flight per sampled cycle, summed over the SQ instances
            # the counters aggregate (a relative occupancy measure)
            "waves_level": ratio("SQ_LEVEL_WAVES", "SQ_CYCLES"),
            "valu_per_wave": ratio("SQ_INSTS_VALU", "SQ_WAVES"),
            "lds_per_wave": ratio("SQ_INSTS_LDS", "SQ_WAVES"),
            "vmem_rd_per_wave": ratio("SQ_INSTS_VMEM_RD", "SQ_WAVES"),
            "vmem_wr_per_wave": ratio("SQ_INSTS_VMEM_WR", "SQ_WAVES"),
            "raw": c}))
    cols = ["waves", "wait_any", "wait_inst_any", "wait_lds", "active_valu", "active_lds",
            "active_vmem", "lds_bank_conflict", "waves_level", "valu_per_wave", "lds_per_wave",
            "vmem_rd_per_wave", "vmem_wr_per_wave"]
    print("| kernel | " + " | ".join(cols) + " |")
    print("|---" * (len(cols) + 1) + "|")
    for kind, variant, r in rows:
        vals = []
        for k in cols:
            v = r[k]
            vals.append("-" if v is None else (f"{v:.0f}" if abs(v) >= 100 else f"{v:.3g}"))
        print(f"| {variant} ({kind}) | " + " | ".join(vals) + " |")
    print()
    print(json.dumps({variant: r["raw"] for _, variant, r in rows}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="timed oracle steps per regime (source on, off) after 2 warm-up steps")
    ap.add_argument("--cpu-fp64", action="store_true",
                    help="also time the oracle in double precision")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-circle-2d", action="store_true")
    ap.add_argument("--bowl", type=int, default=384,
                    help="grid size of the nonlinear precision-study leg (0: skip)")
    ap.add_argument("--bowl-steps", type=int, default=2000)
    ap.add_argument("--axisym", type=int, nargs=2, default=[0, 0], metavar=("NX", "NY"),
                    help="also run kspaceFirstOrderAS on an NX x NY radial grid (0 0: skip)")
    ap.add_argument("--axisym-steps", type=int, default=2000)
    ap.add_argument("--batched", type=int, default=96,
                    help="problems in the batched time-reversal leg (0: skip)")
    ap.add_argument("--served", type=int, default=24,
                    help="requests in the served leg (0: skip)")
    ap.add_argument("--served-rate", type=float, default=10.0)
    ap.add_argument("--served-sweep", type=float, nargs="+", metavar="RATE",
                    help="run only the served workload at these arrival rates (one JSON line)")
    ap.add_argument("--served-steps", type=int, nargs=2, default=[500, 3000])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only the calibration copy (under rocprofv3 --pmc)")
    ap.add_argument("--reduce-traffic", nargs=4, metavar="DIR",
                    help="FETCH_SIZE run, WRITE_SIZE run, and the two calibration runs")
    ap.add_argument("--reduce-counters", nargs="+", metavar="DIR",
                    help="rocprofv3 --pmc runs of the step kernels (one counter set each)")
    ap.add_argument("--traffic-kernel", default="spectral")
    ap.add_argument("--traffic-variant", default="kv_2_256")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.reduce_counters:
        reduce_counters(args.reduce_counters)
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
    kg, med, src, sens, opts = workloads.disc_source_3d(N, Nt=Nt)
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
        # per kind: algorithmic bytes of the model above over the K steps,
        # and the compulsory bytes the library counts (every array a launch
        # touches, once per launch)
        kernels = {}
        for k, v in kern.items():
            if not v["launches"] or k not in KIND_B_STEP:
                continue
            alg = ncell * (K * KIND_B_STEP[k] + on * KIND_B_SRC_ON.get(k, 0.0))
            sec = v["ms"] * 1e-3
            kernels[k] = {"ms_total": round(v["ms"], 4), "launches": v["launches"],
                          "avg_us": round(1e3 * v["ms"] / v["launches"], 3),
                          "alg_bytes_per_launch": round(alg / v["launches"]),
                          "alg_GBps": round(alg / sec / 1e9, 1),
                          "frac": round(alg / sec / 1e9 / HBM_PEAK_GBS, 4),
                          "compulsory_GBps": round(v["bytes"] / sec / 1e9, 1)}
        dom = max(kernels, key=lambda k: kernels[k]["ms_total"])
        d = kernels[dom]
        avg_s = d["avg_us"] * 1e-6
        achieved = d["alg_bytes_per_launch"] / avg_s / 1e9
        traffic = None
        tf = ROOT / "profiles" / f"traffic_{dom}_disc_source_{N}.json"
        if tf.exists():
            traffic = json.loads(tf.read_text()).get("bytes_per_launch")
        # the per-axis-pass model counts 8 B per cell for every axis transform
        # a launch performs; the spectral pass does its forward and inverse
        # transforms of a pencil in one HBM round trip, so its model bytes
        # exceed what it moves (compulsory bytes, and the PMC traffic)
        comp = kern[dom]["bytes"] / kern[dom]["launches"]
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic, "alg_bytes_per_launch": d["alg_bytes_per_launch"],
                    "model": "per-axis-pass (8 B per cell per axis transform)",
                    "compulsory_bytes_per_launch": round(comp),
                    "compulsory_frac": round(comp / avg_s / 1e9 / HBM_PEAK_GBS, 4),
                    "avg_launch_us": d["avg_us"]}
        log(f"[bench] dominant kernel {dom}: {achieved:.0f} GB/s algorithmic "
            f"({achieved / HBM_PEAK_GBS:.1%} of peak)")

    cpu = None
    if not args.no_cpu:
        cores = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        casts = ("single", "off") if args.cpu_fp64 else ("single",)
        log(f"[bench] CPU baseline: oracle {'/'.join(casts)}, 2 + {args.cpu_steps} steps per "
            f"regime, {cores} threads")
        cb = cpu_baseline(N, args.cpu_steps, cores, casts, n_src_on, Nt)
        f32 = cb["fp32"]
        cpu = {"value": round(f32["value"], 3), "unit": "Mcell-steps/s", "cores": cores,
               "kind": "port", "cpu_model": cpu_model(),
               "sample": f"{N}^3 workload, CPU restatement of the reference algorithm (NumPy/"
                         f"scipy.fft, single precision): median step of 2 warm-up + "
                         f"{args.cpu_steps} steps with the source on "
                         f"({f32['s_per_step_on']:.2f} s) and off ({f32['s_per_step_off']:.2f} s), "
                         f"weighted {n_src_on}/{Nt - n_src_on} as in the {Nt}-step run"}
        if "fp64" in cb:
            cpu["fp64_value"] = round(cb["fp64"]["value"], 3)

    acc = None
    if not args.no_accuracy:
        log("[bench] accuracy: 64^3 x 300 steps of the same recipe vs the FP64/FP32 oracle")
        acc = accuracy_check()

    c1 = None
    if not args.no_circle_2d:
        log("[bench] 2-D 128^2 lossless workload in full, GPU and CPU (double precision)")
        c1 = circle_2d_leg()
    c3 = None
    if args.bowl > 0:
        log(f"[bench] nonlinear bowl workload {args.bowl}^3 x {args.bowl_steps} steps")
        c3 = nonlinear_bowl_leg(args.bowl, 2000, args.bowl_steps)
        log(f"[bench] nonlinear bowl workload: {c3['value']} Mcell-steps/s")

    axs = None
    if args.axisym[0] > 0:
        log(f"[bench] kspaceFirstOrderAS {args.axisym[0]} x {args.axisym[1]} x "
            f"{args.axisym_steps} steps, GPU and CPU (single precision)")
        axs = axisym_leg(args.axisym[0], args.axisym[1], args.axisym_steps)
        log(f"[bench] kspaceFirstOrderAS: {axs['gpu_Mcell_steps_s']} Mcell-steps/s")

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
        "circle_2d_128": c1,
        "nonlinear_bowl_384": c3,
        "axisymmetric": axs,
        "device_gb": round(dev_gb, 3),
        "kernels": kernels,
        "batched": batched,
        "served": served,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
