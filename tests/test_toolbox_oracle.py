"""Pins of the new-toolbox oracle (oracle/toolbox_oracle.py) before it
checks the kwhip.toolbox adapter:

* the first run of Nt steps equals the legacy oracle run for Nt + 1 steps
  with the PML outside the grid, no smoothing, the same c_ref
  (TestAcousticSolverGrid.m:74-113 compares the two solvers the same way),
  in 1-D, 2-D and 3-D, lossless and absorbing, in double precision;
* restarting with a new time step (a smaller, then a larger one) equals one
  run at the first step in a homogeneous medium, where the k-space scheme is
  exact (TestAcousticSolver.m:455-616); dropping the split correction term
  or mis-signing it fails this;
* an initial particle velocity gives d'Alembert's solution
  (TestAcousticSolver.m:769-806);
* the recorded sensor data are the fields at the mask, times are k dt
  (TestAcousticSolver.m:1039-1082).
"""

import math

import numpy as np
import pytest

from oracle import kwave_oracle as ko
from oracle import toolbox_oracle as to


def _gauss(shape, dx, width=3.0):
    grids = np.meshgrid(*[(np.arange(n) - n // 2) * dx for n in shape], indexing="ij")
    r2 = sum(g ** 2 for g in grids)
    return np.exp(-r2 / (width * dx) ** 2)


def _legacy_first_run(shape, dx, pad, c, rho, p0, Nt, dt, alpha=None, y=None, c_ref=None):
    Np = [n + 2 * pad for n in shape]
    args = []
    for n in Np:
        args += [n, dx]
    kg = ko.KGrid(*args)
    kg.set_time(Nt + 1, dt)
    pad_e = lambda a: to.pad_edges(a, [pad] * len(shape), len(shape))  # noqa: E731
    med = {"sound_speed": c if np.ndim(c) == 0 else pad_e(c),
           "density": rho if np.ndim(rho) == 0 else pad_e(rho)}
    if c_ref is not None:
        med["sound_speed_ref"] = c_ref
    if alpha is not None:
        med["alpha_coeff"] = alpha
        med["alpha_power"] = y
    sensor = {"mask": np.ones(Np, dtype=bool), "record": ["p_final"]}
    out = ko.kspace_first_order(kg, med, {"p0": pad_e(p0)}, sensor, data_cast="off", smooth=False,
                                pml_size=pad, pml_inside=True)
    pf = np.asarray(out["p_final"]).reshape(Np, order="F")
    return pf[tuple(slice(pad, pad + n) for n in shape)]


@pytest.mark.parametrize("shape,absorbing", [((40,), False), ((24, 20), False), ((24, 20), True),
                                             ((12, 10, 14), True)])
def test_first_run_equals_legacy(shape, absorbing):
    rng = np.random.default_rng(3)
    dx, pad, Nt = 1e-4, 6, 30
    c = 1500 + 60 * rng.random(shape)
    rho = 1000 + 80 * rng.random(shape)
    p0 = _gauss(shape, dx)
    dt = 0.3 * dx / c.max()
    c_ref = float(np.max(to.pad_edges(c, [pad] * len(shape), len(shape))))
    # y = 1.5 is exact in single precision (the reference keeps absorptionPower single)
    alpha, y = (0.6, 1.5) if absorbing else (None, None)
    o = to.ToolboxOracle(shape, dx, pad, c, rho, p0, absorption_coeff=alpha, absorption_power=y,
                         absorption_type="on" if absorbing else "off")
    o.run(Nt, dt)
    ref = _legacy_first_run(shape, dx, pad, c, rho, p0, Nt, dt, alpha, y, c_ref)
    err = np.max(np.abs(o.pressure - ref)) / np.max(np.abs(ref))
    assert err < 1e-9, err


@pytest.mark.parametrize("dim", [1, 2])
def test_restart_with_new_step_equals_one_run(dim):
    n, dx, pad, c0 = 64, 4e-3, 10, 1500.0
    x = (np.arange(n) - n // 2) * dx
    prof = np.exp(-(x - 2e-3) ** 2 / (5 * dx) ** 2)
    shape = (n,) if dim == 1 else (n, 8)
    p0 = prof if dim == 1 else np.repeat(prof[:, None], 8, 1)
    padding = pad if dim == 1 else [pad, 0]
    dt = 0.3 * dx / c0
    Nt = 48

    def solver():
        return to.ToolboxOracle(shape, dx, padding, c0, 1000.0, p0)

    ref = solver()
    ref.run(Nt, dt)
    a = solver()
    a.run(2 * Nt // 3, dt)
    a.run(2 * Nt // 3, dt / 2)
    b = solver()
    b.run(2 * Nt // 3, dt)
    b.run(Nt // 6, 2 * dt)
    # p and the densities live at whole steps, so the three runs meet at the
    # same time; the staggered velocity is half the last step earlier, which
    # differs between them (the reference's test compares it with an
    # absolute 1e-6, far above its 1e-7 scale)
    for s in (a, b):
        for got, want in ((s.pressure, ref.pressure), (s.density_split, ref.density_split)):
            scale = np.max(np.abs(want))
            assert np.max(np.abs(got - want)) <= 1e-6 * scale
    # the split correction matters: without it the restart is off
    c = solver()
    c.run(2 * Nt // 3, dt)
    c._kappa_orig = c._kappa
    c._kappa = lambda dtv: (c._kappa_orig(dtv)[0], None)
    c.run(2 * Nt // 3, dt / 2)
    assert np.max(np.abs(c.pressure - ref.pressure)) > 1e-5 * np.max(np.abs(ref.pressure))


def test_initial_velocity_dalembert():
    n, dx, pad, c0, rho0 = 256, 4e-3, 20, 1500.0, 1000.0
    x = (np.arange(1, n + 1) - math.ceil((n + 1) / 2)) * dx
    g = lambda s: np.exp(-(s ** 2) / (5 * dx) ** 2)  # noqa: E731
    v0 = g(x) / (c0 * rho0)
    dt = 0.25 * dx / c0
    Nt = 150
    o = to.ToolboxOracle((n,), dx, pad, c0, rho0, g(x), u0=[v0])
    o.run(Nt, dt)
    T = Nt * dt
    v = lambda s: g(s) / (c0 * rho0)  # noqa: E731
    exact = (g(x - c0 * T) + c0 * rho0 * v(x - c0 * T) + g(x + c0 * T) -
             c0 * rho0 * v(x + c0 * T)) / 2
    assert np.max(np.abs(o.pressure - exact)) < 1e-6


def test_sensor_records_fields():
    shape, dx = (16, 14), 4e-3
    p0 = _gauss(shape, dx, 5.0)
    mask = np.ones(shape, dtype=bool)
    o = to.ToolboxOracle(shape, dx, 0, 1500.0, 1000.0, p0, mask=mask, velocity_sensor="on",
                         density_sensor="on", data_type="single")
    o.run(0, 1e-6)
    p, u, r, t = o.sensor_data()
    assert p.shape == (mask.sum(), 1) and u.shape == (mask.sum(), 2, 1) and t.shape == (1,)
    assert np.array_equal(p[:, 0], np.reshape(o.pressure, -1, order="F"))
    assert np.array_equal(u[:, :, 0], np.reshape(o.velocity, (-1, 2), order="F"))
    assert np.allclose(r[:, 0], np.reshape(o.density_split, (-1, 2), order="F").sum(1), atol=0)
    # run(Nt=0) leaves timeStepsTaken at 0 (run.m:97), so the next run starts
    # over: the initial condition is added again and recorded again, and the
    # sensor arrays grow past their preallocated width
    o.run(5, 1e-6)
    p, u, r, t = o.sensor_data()
    assert p.shape[1] == 7
    assert np.allclose(t, np.r_[0.0, np.arange(6) * np.float32(1e-6)], rtol=1e-6)
    assert np.allclose(p[:, 1], 2 * p[:, 0], rtol=1e-6)
    assert np.array_equal(p[:, -1], np.reshape(o.pressure, -1, order="F"))
