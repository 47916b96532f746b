"""Cases of the new-toolbox adapter (kwhip.toolbox) and their runs through
the adapter and through the toolbox oracle, shared by the emulator and GPU
tests."""

import numpy as np

from oracle import toolbox_oracle as to


def gauss(shape, dx, width=3.0, centre=None):
    centre = centre or [0.0] * len(shape)
    grids = np.meshgrid(*[(np.arange(n) - n // 2) * dx - c for n, c in zip(shape, centre)],
                        indexing="ij")
    return np.exp(-sum(g ** 2 for g in grids) / (width * dx) ** 2)


def case(name, shape, pad=5, seed=1):
    """A case: grid, medium, sources, sensor settings, absorption type and the
    runs [(Nt, dt), ...]."""
    rng = np.random.default_rng(seed)
    dim = len(shape)
    dx = 1e-4
    c = 1500 + 60 * rng.random(shape)
    rho = 1000 + 80 * rng.random(shape)
    dt = 0.3 * dx / c.max()
    cs = {"shape": shape, "dx": dx, "pad": pad, "c": c, "rho": rho, "p0": gauss(shape, dx),
          "u0": None, "alpha": None, "y": None, "absorption": "off",
          "mask": rng.random(shape) < 0.3, "velocity": "on", "density": "on", "ts": 1, "tss": 1,
          "runs": [(12, dt)]}
    if name == "homogeneous":
        cs["c"], cs["rho"] = 1500.0, 1000.0
    elif name == "restart_smaller":
        cs["runs"] = [(10, dt), (8, dt / 2)]
    elif name == "restart_larger":
        cs["runs"] = [(10, dt), (4, 1.5 * dt), (3, 1.5 * dt)]
    elif name == "absorbing":
        cs["alpha"], cs["y"], cs["absorption"] = 0.8 + 0.4 * rng.random(shape), 1.5, "on"
        cs["runs"] = [(10, dt), (5, dt / 2)]
    elif name == "no_dispersion":
        cs["alpha"], cs["y"], cs["absorption"] = 0.8, 1.3, "noDispersion"
    elif name == "no_absorption":
        cs["alpha"], cs["y"], cs["absorption"] = 0.8, 1.3, "noAbsorption"
    elif name == "initial_velocity":
        cs["u0"] = [gauss(shape, dx, 4.0, [dx] * dim) / (1500 * 1000) * (a + 1)
                    for a in range(dim)]
        cs["runs"] = [(10, dt), (5, dt)]
    elif name == "ongrid_every_3rd":
        cs["velocity"], cs["ts"], cs["density"] = "ongrid", 3, "off"
        cs["runs"] = [(10, dt), (7, dt)]
    elif name == "run_zero_first":
        cs["runs"] = [(0, dt), (6, dt), (4, dt / 2)]
    else:
        raise KeyError(name)
    return cs


NAMES = ["homogeneous", "restart_smaller", "restart_larger", "absorbing", "no_dispersion",
         "no_absorption", "initial_velocity", "ongrid_every_3rd", "run_zero_first"]


def run_adapter(cs):
    from kwhip import toolbox as tb
    kg = tb.Grid(cs["shape"], cs["dx"], cs["pad"])
    m = tb.AcousticMedium(kg)
    m.soundSpeed = cs["c"]
    m.density = cs["rho"]
    if cs["alpha"] is not None:
        m.absorptionCoeff = cs["alpha"]
        m.absorptionPower = cs["y"]
    s = tb.AcousticSource(kg)
    s.initialPressure = cs["p0"]
    if cs["u0"] is not None:
        s.initialVelocity = np.stack(cs["u0"], -1)
    se = tb.AcousticSensor(kg)
    se.mask = cs["mask"]
    se.velocitySensor = cs["velocity"]
    se.densitySensor = cs["density"]
    se.timeSteps = cs["ts"]
    se.timeStepSpacing = cs["tss"]
    sol = tb.AcousticSolver(kg, m, s, se)
    sol.absorptionType = cs["absorption"]
    for Nt, dt in cs["runs"]:
        sol.run(Nt=Nt, dt=dt)
    out = {"pressure": sol.pressure, "velocity": sol.velocity, "densitySplit": sol.densitySplit,
           "sensor.pressure": se.pressure, "sensor.velocity": se.velocity,
           "sensor.density": se.density, "sensor.times": se.times,
           "timeArray": sol.timeArray}
    sol.close()
    return out


def run_oracle(cs, data_type="single"):
    o = to.ToolboxOracle(cs["shape"], cs["dx"], cs["pad"], cs["c"], cs["rho"], cs["p0"],
                         u0=cs["u0"], absorption_coeff=cs["alpha"], absorption_power=cs["y"],
                         absorption_type=cs["absorption"], data_type=data_type,
                         mask=cs["mask"], velocity_sensor=cs["velocity"],
                         density_sensor=cs["density"], time_steps=cs["ts"],
                         time_step_spacing=cs["tss"])
    for Nt, dt in cs["runs"]:
        o.run(Nt, dt)
    P, U, R, T = o.sensor_data()
    return {"pressure": o.pressure, "velocity": o.velocity, "densitySplit": o.density_split,
            "sensor.pressure": P, "sensor.velocity": U,
            "sensor.density": R if cs["density"] == "on" else None, "sensor.times": T,
            "timeArray": np.asarray(o.time_array, dtype=np.float32)}


def compare(got, ref):
    """Relative L-inf per output (shapes must match); returns {name: err}."""
    errs = {}
    for k, r in ref.items():
        g = got[k]
        if r is None:
            assert g is None, k
            continue
        g, r = np.asarray(g, dtype=np.float64), np.asarray(r, dtype=np.float64)
        assert g.shape == r.shape, (k, g.shape, r.shape)
        scale = np.max(np.abs(r))
        errs[k] = float(np.max(np.abs(g - r)) / scale) if scale > 0 else float(np.max(np.abs(g)))
    return errs
