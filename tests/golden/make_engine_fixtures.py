"""Writes the golden fixtures of the engine workloads at their real sizes,
from the oracle in single and double precision:

* tr_<seed>.npz: one photoacoustic problem of the batched time-reversal
  workload (workloads.batch_tr_problem: random cylinders or tubes,
  homogeneous c0 in [1450, 1600] m/s, a sensor on the face x = PML + 1):
  the forward run's p at the sensor and the time-reversal reconstruction's
  p_final, the reconstruction fed with that run's own forward data.
  Seeds 34 (2-D 256^2, 1207 steps) and 29 (3-D 64^3).
* served_<seed>_<steps>.npz: one request of the served workload
  (workloads.served_problem: heterogeneous c and rho, power-law absorption,
  B/A = 6, 1 MPa tone burst, full-grid p_max and p_rms) for the given number
  of steps; seed 34 is a 2-D 256^2 problem.

Every array is stored subsampled (the stride is in the file) with the
largest magnitude of the whole array, and a CRC of the inputs.

Run from the repository root:
    python tests/golden/make_engine_fixtures.py tr 34
    python tests/golden/make_engine_fixtures.py tr 29
    python tests/golden/make_engine_fixtures.py served 34 500
"""

import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path[:0] = [str(ROOT), str(ROOT / "k-wave-ii_amd"), str(HERE)]

from make_longrun import input_crc  # noqa: E402

TS_STRIDE = 4   # sensor points and time steps of the forward p series
GRID_STRIDE = 2  # cells along each axis of p_final, p_max and p_rms


def oracle_grid(kg):
    from oracle import kwave_oracle as ko
    args = []
    for n, d in zip(kg.shape, (kg.dx, kg.dy, kg.dz)):
        args += [n, d]
    g = ko.KGrid(*args)
    g.set_time(kg.Nt, kg.dt)
    return g


def sub_grid(v, shape):
    v = np.asarray(v, dtype=np.float64)
    if v.ndim == 1 and len(shape) > 1:
        v = v.reshape(shape, order="F")
    return v[tuple(slice(None, None, GRID_STRIDE) for _ in shape)]


def tr_problem(seed):
    from kwhip import workloads
    return workloads.batch_tr_problem(seed)


def run_tr(seed):
    from oracle import kwave_oracle as ko
    kg, med, src, sens, opts = tr_problem(seed)
    arrays = {}
    for cast, tag in (("single", "fp32"), ("off", "fp64")):
        t0 = time.time()
        g = oracle_grid(kg)
        kw = dict(data_cast=cast, pml_size=opts["PMLSize"], pml_alpha=opts["PMLAlpha"],
                  smooth=opts["Smooth"])
        fwd = np.asarray(ko.kspace_first_order(g, med, src, sens, **kw), dtype=np.float64)
        tr = ko.kspace_first_order(g, med, {}, {"mask": sens["mask"],
                                                "time_reversal_boundary_data": fwd}, **kw)
        print(f"tr seed {seed} {tag}: {time.time() - t0:.0f} s", flush=True)
        arrays[f"{tag}_fwd"] = fwd[::TS_STRIDE, ::TS_STRIDE].astype(np.float32)
        arrays[f"{tag}_fwd_absmax"] = np.float64(np.max(np.abs(fwd)))
        pf = np.asarray(tr, dtype=np.float64)
        arrays[f"{tag}_tr"] = sub_grid(pf, kg.shape).astype(np.float32)
        arrays[f"{tag}_tr_absmax"] = np.float64(np.max(np.abs(pf)))
    out = HERE / f"tr_{seed}.npz"
    np.savez_compressed(out, seed=seed, N=np.array(kg.shape), Nt=kg.Nt, crc=np.uint32(input_crc(med, src)),
                        ts_stride=TS_STRIDE, grid_stride=GRID_STRIDE, **arrays)
    print(out, flush=True)


def served(seed, steps):
    from kwhip import workloads
    return workloads.served_problem(seed, steps, kind="forward")


def run_served(seed, steps):
    from oracle import kwave_oracle as ko
    kg, med, src, sens, opts = served(seed, steps)
    arrays = {}
    for cast, tag in (("single", "fp32"), ("off", "fp64")):
        t0 = time.time()
        r = ko.kspace_first_order(oracle_grid(kg), med, src, sens, data_cast=cast,
                                  pml_size=opts["PMLSize"], pml_alpha=opts["PMLAlpha"],
                                  smooth=opts["Smooth"])
        print(f"served seed {seed} {tag}: {time.time() - t0:.0f} s", flush=True)
        for k in ("p_max", "p_rms"):
            v = np.asarray(r[k], dtype=np.float64)
            arrays[f"{tag}_{k}"] = sub_grid(v, kg.shape).astype(np.float32)
            arrays[f"{tag}_{k}_absmax"] = np.float64(np.max(np.abs(v)))
    out = HERE / f"served_{seed}_{steps}.npz"
    np.savez_compressed(out, seed=seed, N=np.array(kg.shape), Nt=kg.Nt,
                        crc=np.uint32(input_crc(med, src)), grid_stride=GRID_STRIDE, **arrays)
    print(out, flush=True)


if __name__ == "__main__":
    if sys.argv[1] == "tr":
        run_tr(int(sys.argv[2]))
    else:
        run_served(int(sys.argv[2]), int(sys.argv[3]))
