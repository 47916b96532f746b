"""The engine on the batched time-reversal and served workloads at their
real sizes, against oracle fixtures (tests/golden/make_engine_fixtures.py)
and against the same problems run alone.

* Batched photoacoustic time reversal: a 2-D 256^2 problem (1207 steps)
  and a 3-D 64^3 problem submitted together; each reconstruction is
  submitted when its forward run returns, with that run's sensor data.  The
  forward p at the sensor and the reconstruction's p_final: relative L-inf
  <= 1e-5 against the oracle in single precision, <= 1e-4 against double
  precision (each normalised by the largest magnitude of the whole array),
  and <= 1e-6 against the same chain run alone through
  kspaceFirstOrder2D/3D.
* Served request: a 2-D 256^2 request of the served workload (heterogeneous,
  absorbing, B/A = 6, 1 MPa source, 500 steps) submitted while a 3-D 96^3
  request of 2000 steps is on the device; p_max and p_rms to the same
  bars.
"""

import sys
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))

TOL32, TOL64, TOL_ALONE = 1e-5, 1e-4, 1e-6


def _fixture(name):
    path = GOLDEN / name
    if not path.exists():
        pytest.fail(f"missing fixture {path}: run tests/golden/make_engine_fixtures.py")
    return np.load(path)


def _err(got, ref, absmax):
    return float(np.max(np.abs(np.asarray(got, np.float64) - np.asarray(ref, np.float64))) / absmax)


def _solo(kg, med, src, sens, opts):
    import kwhip
    fn = kwhip.kspaceFirstOrder2D if kg.dim == 2 else kwhip.kspaceFirstOrder3D
    return fn(kg, med, src, sens, **opts)


def test_batched_time_reversal_real_sizes(hip_lib):
    import kwhip
    import make_engine_fixtures as mf
    from make_longrun import input_crc

    seeds = (34, 29)
    probs = {s: mf.tr_problem(s) for s in seeds}
    fx = {s: _fixture(f"tr_{s}.npz") for s in seeds}
    for s in seeds:
        assert int(fx[s]["crc"]) == input_crc(probs[s][1], probs[s][2])
        print(f"[engine-tr] seed {s}: grid {probs[s][0].shape}, {probs[s][0].Nt} steps", flush=True)
    fwd, rec = {}, {}
    with kwhip.Engine(steps_per_round=32) as eng:
        tf = {s: eng.submit(kg, med, src, sens, kg.dim, **opts)
              for s, (kg, med, src, sens, opts) in probs.items()}
        tt = {}
        for s in seeds:
            kg, med, src, sens, opts = probs[s]
            fwd[s] = np.asarray(eng.wait(tf[s]), np.float64)
            tt[s] = eng.submit(kg, med, {}, {"mask": sens["mask"],
                                             "time_reversal_boundary_data": fwd[s]}, kg.dim, **opts)
        for s in seeds:
            rec[s] = np.asarray(eng.wait(tt[s]), np.float64)
    for s in seeds:
        kg, med, src, sens, opts = probs[s]
        f, ts, gs = fx[s], int(fx[s]["ts_stride"]), int(fx[s]["grid_stride"])
        sub = tuple(slice(None, None, gs) for _ in kg.shape)
        for tag, tol in (("fp32", TOL32), ("fp64", TOL64)):
            e_f = _err(fwd[s][::ts, ::ts], f[f"{tag}_fwd"], float(f[f"{tag}_fwd_absmax"]))
            e_t = _err(rec[s][sub], f[f"{tag}_tr"], float(f[f"{tag}_tr_absmax"]))
            print(f"[engine-tr] seed {s} vs {tag} oracle: forward {e_f:.2e}, reconstruction {e_t:.2e}",
                  flush=True)
            assert e_f <= tol and e_t <= tol, (s, tag, e_f, e_t)
        alone_f = np.asarray(_solo(kg, med, src, sens, opts), np.float64)
        alone_t = np.asarray(_solo(kg, med, {}, {"mask": sens["mask"],
                                                 "time_reversal_boundary_data": alone_f}, opts),
                             np.float64)
        e_f = _err(fwd[s], alone_f, np.max(np.abs(alone_f)))
        e_t = _err(rec[s], alone_t, np.max(np.abs(alone_t)))
        print(f"[engine-tr] seed {s} vs alone: forward {e_f:.2e}, reconstruction {e_t:.2e}", flush=True)
        assert e_f <= TOL_ALONE and e_t <= TOL_ALONE, (s, e_f, e_t)


def test_served_request_beside_a_running_one(hip_lib):
    import time

    import kwhip
    import make_engine_fixtures as mf
    from kwhip import workloads
    from make_longrun import input_crc

    f = _fixture("served_34_500.npz")
    kg, med, src, sens, opts = mf.served(34, 500)
    assert int(f["crc"]) == input_crc(med, src)
    other = workloads.served_problem(53, 2000, kind="forward")
    print(f"[engine-served] 256^2 x 500 beside {other[0].shape} x {other[0].Nt}", flush=True)
    with kwhip.Engine(steps_per_round=32) as eng:
        t_other = eng.submit(other[0], other[1], other[2], other[3], other[0].dim, **other[4])
        # submit once the other request is stepping on the device
        t_end = time.time() + 60
        while eng.poll(t_other)["steps_done"] < 2 and eng.poll(t_other)["status"] != "done":
            assert time.time() < t_end, "the other request never started stepping"
            time.sleep(0.002)
        t = eng.submit(kg, med, src, sens, kg.dim, **opts)
        info_other = eng.poll(t_other)
        got = eng.wait(t)
        eng.wait(t_other)
    print(f"[engine-served] the other request was {info_other['status']} at submission "
          f"({info_other['steps_done']} of {info_other['steps_total']} steps)", flush=True)
    gs = int(f["grid_stride"])
    sub = tuple(slice(None, None, gs) for _ in kg.shape)
    alone = _solo(kg, med, src, sens, opts)
    for k in ("p_max", "p_rms"):
        g = np.asarray(got[k], np.float64).reshape(kg.shape, order="F")
        for tag, tol in (("fp32", TOL32), ("fp64", TOL64)):
            e = _err(g[sub], f[f"{tag}_{k}"], float(f[f"{tag}_{k}_absmax"]))
            print(f"[engine-served] {k} vs {tag} oracle: {e:.2e}", flush=True)
            assert e <= tol, (k, tag, e)
        a = np.asarray(alone[k], np.float64).reshape(kg.shape, order="F")
        e = _err(g, a, np.max(np.abs(a)))
        print(f"[engine-served] {k} vs alone: {e:.2e}", flush=True)
        assert e <= TOL_ALONE, (k, e)
