"""GPU parity at the benchmark sizes.

* The 256^3 benchmark workload (heterogeneous medium, power-law absorption,
  additive disc source, full-grid p_max/p_rms) for its first steps, against
  the oracle run on the same inputs (gates: 1e-5 vs single precision, 1e-4
  vs double precision).  The source signal starts inside the tone burst so
  that the source acts from the first step.
* A size-independent property at 256^3 that holds for any number of steps:
  with the PML off the grid is periodic and, for a homogeneous lossless
  medium, k-space stepping is exact, so step n holds P0(k) cos(c|k|(n-1)dt)
  (checked in FP64 with numpy's FFT; FP32 rounding gate 1e-4).
* The 2D 128^2 initial-value workload (p0 disc, 50 Cartesian points,
  nearest-point sensor) for all of its steps, against the oracle.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-4


def test_fullsize_workload_first_steps(hip_lib):
    import kwhip
    from kwhip import workloads
    from oracle import kwave_oracle as ko

    Nt = 3
    kg, med, src, sens, opts = workloads.disc_source_3d(256, Nt=Nt)
    src["p"] = src["p"][:, 250:]
    sens = {"mask": sens["mask"], "record": ["p_max", "p_rms", "p_final"]}
    print("[fullsize] 256^3 workload, 3 steps", flush=True)
    got = kwhip.kspaceFirstOrder3D(kg, med, src, sens, **opts)
    kgo = ko.KGrid(256, kg.dx, 256, kg.dy, 256, kg.dz)
    kgo.set_time(Nt, kg.dt)
    for cast, tol in (("single", TOL32), ("off", TOL64)):
        ref = ko.kspace_first_order(kgo, med, src, sens, data_cast=cast, pml_size=opts["PMLSize"],
                                    pml_alpha=opts["PMLAlpha"])
        for k in ("p_max", "p_rms", "p_final"):
            err = ko.rel_linf(np.asarray(got[k], np.float64), np.asarray(ref[k], np.float64))
            print(f"[fullsize] {cast} {k} {err:.2e}", flush=True)
            assert err <= tol, (cast, k, err)


def test_fullsize_periodic_exact(hip_lib):
    import kwhip

    N, dx, c, Nt = 256, 1e-4, 1500.0, 40
    dt = 0.3 * dx / c
    rng = np.random.default_rng(21)
    p0 = rng.standard_normal((N, N, N))
    kg = kwhip.kWaveGrid(N, dx, N, dx, N, dx)
    kg.setTime(Nt, dt)
    print("[fullsize] 256^3 periodic exact solution", flush=True)
    got = kwhip.kspaceFirstOrder3D(kg, {"sound_speed": c, "density": 1000.0}, {"p0": p0},
                                   {"mask": np.ones((N, N, N), dtype=bool), "record": ["p_final"]},
                                   PMLSize=2, PMLAlpha=0.0, Smooth=False)["p_final"]
    f = np.fft.fftfreq(N, dx) * 2 * math.pi
    k = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2)
    want = np.real(np.fft.ifftn(np.fft.fftn(p0) * np.cos(c * k * (Nt - 1) * dt)))
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    print(f"[fullsize] periodic rel L-inf {err:.2e}", flush=True)
    assert err <= TOL64


def test_fullsize_2d_initial_value_workload(hip_lib):
    import kwhip
    from kwhip import workloads
    from oracle import kwave_oracle as ko

    kg, med, src, sens, opts = workloads.circle_2d(128)
    print(f"[fullsize] 128^2 initial value, {kg.Nt} steps", flush=True)
    got = kwhip.kspaceFirstOrder2D(kg, med, src, sens, **opts)
    kgo = ko.KGrid(128, kg.dx, 128, kg.dy)
    kgo.set_time(kg.Nt, kg.dt)
    for cast, tol in (("single", TOL32), ("off", TOL64)):
        ref = ko.kspace_first_order(kgo, med, src, sens, data_cast=cast, pml_size=opts["PMLSize"],
                                    pml_alpha=opts["PMLAlpha"], cart_interp="nearest")
        err = ko.rel_linf(np.asarray(got, np.float64), np.asarray(ref, np.float64))
        print(f"[fullsize] 2D {cast} {err:.2e}", flush=True)
        assert err <= tol, (cast, err)
