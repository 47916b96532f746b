"""The new toolbox's acoustic solver (kwave.toolbox: Grid, AcousticMedium,
AcousticSource, AcousticSensor, Settings, AcousticSolver) on the kwhip
kernels.

The reference's class solver (+kwave/+toolbox/@AcousticSolver,
@TimeDomainSolver, @FourierCollocation, @SplitFieldPML) steps the same
first-order k-space equations as kspaceFirstOrder1D/2D/3D on a grid padded by
the PML (Grid.gridPadding): a first run of Nt steps is the legacy loop run
for Nt + 1 steps with the PML outside the grid and no smoothing
(TestAcousticSolverGrid.m:74-113), and further runs continue from the same
state, with a new time step each time if asked (executeTimeStep.m:45-55,
143-176).  Here a run drives one device-resident simulation (kw_sim_* in
include/kwhip.h): the grid inputs are padded as the reference pads them
(edge values repeated, Grid/assignWithGridPadding.m), the first run creates
the simulation and steps Nt + 1 times, later runs extend it
(kw_sim_extend), arm the transition step when the time step changes
(kw_sim_change_dt), and an initial particle velocity is the first step
(kw_sim_init_velocity, then the transition from a zero step).  Sensors
record on the device every step; the records the reference keeps (every
sensor.timeSteps-th step) and its time points are picked on the host.

Differences from the reference, all deliberate:
* arithmetic is single precision whatever Settings.simulationDataType says
  (the reference's default is 'single'; 'double' runs here agree with it to
  single-precision accuracy);
* plotting options are accepted and ignored;
* the 'ongrid' velocity sensor unstaggers each component along its own
  axis in every dimension (the reference's 2-D and 3-D branch assigns a
  vector field to one component, @AcousticSensor/AcousticSensor.m:163-168,
  and stops with a size error; 1-D is the same);
* medium.absorptionPower = 1 with dispersion is rejected (the reference
  evaluates tan(pi/2)).
Output arrays are NumPy, MATLAB's column-major order kept: solver.pressure
is the grid (Nx,), (Nx, Ny) or (Nx, Ny, Nz); solver.velocity and
solver.densitySplit add a last axis of dim components; sensor.pressure and
sensor.density are [points x T], sensor.velocity [points x dim x T],
sensor.times [T], points in column-major mask order.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _abi
from .kgrid import kWaveGrid
from .solver import _parse_options, prepare

__all__ = ["Grid", "AcousticMedium", "AcousticSource", "AcousticSensor", "Settings",
           "AcousticSolver"]


def _error(ident, msg):
    raise ValueError(f"{ident}: {msg}")


class Grid:
    """Grid(gridSize, gridSpacing, gridPadding) (@Grid/Grid.m:101-121)."""

    def __init__(self, gridSize, gridSpacing, gridPadding=None):
        size = [int(v) for v in np.atleast_1d(gridSize)]
        if not 1 <= len(size) <= 3 or any(v < 1 for v in size):
            _error("Grid:incorrectInputSize", "gridSize must have 1 to 3 positive entries.")
        self.dimensions = len(size)
        sp = np.atleast_1d(np.asarray(gridSpacing, dtype=np.float64))
        if sp.size not in (1, self.dimensions):
            _error("Grid:incorrectInputSize",
                   "gridSpacing must be a scalar or the same length as gridSize.")
        pad = [0] * self.dimensions
        if gridPadding is not None:
            pd = np.atleast_1d(gridPadding)
            if pd.size not in (1, self.dimensions):
                _error("Grid:incorrectInputSize",
                       "gridSpacing must be a scalar or the same length as gridSize.")
            pad = [int(v) for v in np.broadcast_to(pd, (self.dimensions,))]
        self.gridSize = tuple(size + [1] * (3 - self.dimensions))
        self.gridSpacing = tuple(list(np.broadcast_to(sp, (self.dimensions,))) +
                                 [0.0] * (3 - self.dimensions))
        self.gridPadding = tuple(pad + [0] * (3 - self.dimensions))

    Nx = property(lambda s: s.gridSize[0])
    Ny = property(lambda s: s.gridSize[1])
    Nz = property(lambda s: s.gridSize[2])
    dx = property(lambda s: s.gridSpacing[0])
    dy = property(lambda s: s.gridSpacing[1])
    dz = property(lambda s: s.gridSpacing[2])
    totalGridPoints = property(lambda s: int(np.prod(s.gridSize)))

    def _vec(self, n, d):
        return (np.arange(1, n + 1) - math.ceil((n + 1) / 2)) * d

    xVec = property(lambda s: s._vec(s.Nx, s.dx))
    yVec = property(lambda s: s._vec(s.Ny, s.dy))
    zVec = property(lambda s: s._vec(s.Nz, s.dz))

    @property
    def shape(self):
        """The grid shape of a scalar field here: (Nx,), (Nx, Ny) or (Nx, Ny, Nz)."""
        return tuple(self.gridSize[:self.dimensions])

    def _check_field(self, value, name, vector=False):
        a = np.asarray(value)
        if a.size == 1:
            return a.astype(np.float64).reshape(())
        want = self.shape + ((self.dimensions,) if vector else ())
        if a.shape != want:
            # MATLAB sizes: a 1-D field may come as a row or column
            if self.dimensions == 1 and a.size == self.Nx * (self.dimensions if vector else 1):
                return a.reshape(want, order="F")
            _error("Grid:incorrectSize", f"{name} must be a scalar or of size {list(want)}.")
        return a


class _GridInput:
    """Holder of grid fields (@GridInput): values are kept unpadded and
    checked against the grid when set."""

    _fields: dict = {}
    _required: tuple = ()

    def __init__(self, kgrid):
        if not isinstance(kgrid, Grid):
            raise TypeError("kgrid must be a kwhip.toolbox.Grid")
        object.__setattr__(self, "kgrid", kgrid)
        object.__setattr__(self, "_values", {})

    def __setattr__(self, name, value):
        if name in self._fields:
            if value is None:
                self._values.pop(name, None)
                return
            vector = self._fields[name] == "vector"
            if self._fields[name] == "binary":
                a = np.asarray(value)
                if a.shape != self.kgrid.shape:
                    _error("Grid:incorrectSize", f"{name} must be of size {list(self.kgrid.shape)}.")
                if not np.all((a == 0) | (a == 1)):
                    _error("GridInput:invalidValue", f"{name} must be binary.")
                self._values[name] = a.astype(bool)
                return
            a = self.kgrid._check_field(value, name, vector)
            if not np.all(np.isfinite(a)) and name in ("initialPressure", "initialVelocity"):
                _error("GridInput:invalidValue", f"{name} must be finite.")
            self._values[name] = np.asarray(a, dtype=np.float64)
            return
        object.__setattr__(self, name, value)

    def __getattr__(self, name):
        if name in type(self)._fields:
            return self._values.get(name)
        raise AttributeError(name)

    def checkRequiredProperties(self):
        for name in self._required:
            if self._values.get(name) is None:
                _error("GridInput:missingInput", f"The property {name} must be defined.")


class AcousticMedium(_GridInput):
    """soundSpeed, density, absorptionCoeff, BonA (grid fields),
    soundSpeedReference, absorptionPower (@AcousticMedium)."""

    _fields = {"soundSpeed": "scalar", "density": "scalar", "absorptionCoeff": "scalar",
               "BonA": "scalar"}
    _required = ("soundSpeed", "density")

    def __init__(self, kgrid):
        super().__init__(kgrid)
        self.soundSpeedReference = None
        self.absorptionPower = None


class AcousticSource(_GridInput):
    """initialPressure (required) and initialVelocity (a vector field of the
    grid shape plus a last axis of dim components) (@AcousticSource)."""

    _fields = {"initialPressure": "scalar", "initialVelocity": "vector"}
    _required = ("initialPressure",)


class AcousticSensor(_GridInput):
    """mask (binary), pressureSensor 'on'/'off', velocitySensor
    'on'/'off'/'ongrid', densitySensor 'on'/'off', timeSteps (record every
    timeSteps-th step), timeStepSpacing (preallocated columns per step);
    outputs pressure, velocity, density, times (@AcousticSensor, @Sensor)."""

    _fields = {"mask": "binary"}
    _required = ("mask",)

    def __init__(self, kgrid):
        super().__init__(kgrid)
        self.pressureSensor = "on"
        self.velocitySensor = "off"
        self.densitySensor = "off"
        self.timeSteps = 1
        self.timeStepSpacing = 1
        self.pressure = None
        self.velocity = None
        self.density = None
        self.times = None

    @property
    def totalSensorPoints(self):
        return int(np.count_nonzero(self.mask)) if self.mask is not None else 0


class Settings:
    """simulationDataType ('single' or 'double'), plotSimulation,
    plotFrequency, plotScale (@Settings; plotting is not done here)."""

    def __init__(self):
        self.simulationDataType = "single"
        self.plotSimulation = False
        self.plotFrequency = 10
        self.plotScale = (-1.0, 1.0)


class _PML:
    """solver.pml: the PML absorption per axis (@SplitFieldPML pmlAlpha)."""

    def __init__(self):
        self.pmlAlpha = [2.0, 2.0, 2.0]


class AcousticSolver:
    """AcousticSolver(kgrid, medium, source, sensor=None, settings=None);
    run(Nt=, dt=) or run(CFL=, EndTime=) (@AcousticSolver/AcousticSolver.m:
    159-184, @TimeDomainSolver/run.m:58-101)."""

    def __init__(self, kgrid, medium, source, sensor=None, settings=None):
        if not isinstance(kgrid, Grid):
            raise TypeError("kgrid must be a kwhip.toolbox.Grid")
        if not isinstance(medium, AcousticMedium):
            _error("AcousticSolver:InvalidMediumType",
                   "medium must be an object of Medium or AcousticMedium class.")
        if sensor is not None and not isinstance(sensor, AcousticSensor):
            _error("AcousticSolver:InvalidSensorType",
                   "sensor must be an AcousticSensor object or be left empty [].")
        if not isinstance(source, AcousticSource):
            raise TypeError("source must be a kwhip.toolbox.AcousticSource")
        # Solver.m:87-122
        if medium.kgrid is not kgrid:
            _error("Solver:gridMismatch",
                   "The medium input references a different Grid object to the kgrid input.")
        if source.kgrid is not kgrid:
            _error("Solver:gridMismatch",
                   "The source input references a different Grid object to the kgrid input.")
        if sensor is not None and sensor.kgrid is not kgrid:
            _error("Solver:gridMismatch",
                   "The source input references a different Grid object to the kgrid input.")
        medium.checkRequiredProperties()
        source.checkRequiredProperties()
        self.kgrid, self.medium, self.source, self.sensor = kgrid, medium, source, sensor
        self.settings = settings if settings is not None else Settings()
        self.dimensions = kgrid.dimensions
        self.pml = _PML()
        self.absorptionType = "off"
        self.prevTimeStep = None
        self.timeArray = np.zeros(0, dtype=np.float32)
        self.timeStepsTaken = 0
        self._sim = None
        self._lib = None
        self._steps = 0         # device steps done
        self._ic_count = 0      # initial conditions applied (run(Nt=0) repeats them)
        self._records = 0       # sensor records kept (the reference's sensorIndex)
        self._width = 0

    # -- helpers ------------------------------------------------------------
    def _padded(self, a):
        g = self.kgrid
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 0:
            return float(a)
        pad = g.gridPadding[:g.dimensions]
        if not any(pad):
            return a
        return np.pad(a, [(p, p) for p in pad], mode="edge")

    def _padded_shape(self):
        g = self.kgrid
        return tuple(n + 2 * p for n, p in zip(g.shape, g.gridPadding))

    def _interior(self, a):
        g = self.kgrid
        return a[tuple(slice(p, p + n) for p, n in zip(g.gridPadding, g.shape))]

    def _absorption(self):
        """The operators setAbsorptionCoefficients.m builds, as medium inputs of
        the kernels (the same conversion, τ and η), or None when they are all
        zero (executeTimeStep.m:67-92)."""
        if self.absorptionType == "off":
            return None
        if self.absorptionType not in ("on", "noAbsorption", "noDispersion"):
            _error("AcousticSolver:invalidAbsorptionType",
                   "absorptionType must be 'off', 'on', 'noAbsorption' or 'noDispersion'.")
        m = self.medium
        if m.absorptionCoeff is None or m.absorptionPower is None:
            return None
        y = float(np.float32(m.absorptionPower))
        alpha = np.asarray(self._padded(m.absorptionCoeff), dtype=np.float64)
        alpha = np.where(np.isfinite(alpha), alpha, 0.0)
        no_disp = self.absorptionType == "noDispersion" or y in (0.0, 2.0)
        no_abs = self.absorptionType == "noAbsorption"
        if (no_abs and no_disp) or not np.any(alpha != 0):
            return None
        mode = "no_absorption" if no_abs else ("no_dispersion" if no_disp else None)
        if no_abs and no_disp:
            return None
        return alpha, y, mode

    def _auto_time_step(self, CFL, EndTime):
        """autoComputeTimeStep.m:40-69."""
        g, m = self.kgrid, self.medium
        CFL = 0.3 if CFL is None else float(CFL)
        c = np.asarray(m.soundSpeed, dtype=np.float64)
        cmin, cmax = float(np.min(c)), float(np.max(c))
        if EndTime is None:
            EndTime = float(np.linalg.norm(np.asarray(g.gridSize) * np.asarray(g.gridSpacing))) / cmin
        dx = min(g.gridSpacing[:g.dimensions])
        dt = CFL * dx / cmax
        Nt = int(math.ceil(EndTime / dt))
        dt = float(np.float32(EndTime / np.float32(Nt)))
        return Nt, dt

    def _create(self, Nt_total, dt, scale):
        g, m, src, sens = self.kgrid, self.medium, self.source, self.sensor
        dim = g.dimensions
        shp = self._padded_shape()
        args = []
        for n, d in zip(shp, g.gridSpacing):
            args += [n, d]
        kg = kWaveGrid(*args)
        kg.setTime(int(Nt_total), float(dt))
        c = self._padded(m.soundSpeed)
        rho = self._padded(m.density)
        c_ref = m.soundSpeedReference
        if c_ref is None:
            c_ref = float(np.max(c))  # setInitialConditions.m:36-38
            m.soundSpeedReference = c_ref
        med = {"sound_speed": c, "density": rho, "sound_speed_ref": float(c_ref)}
        ab = self._absorption()
        if ab is not None:
            alpha, y, mode = ab
            med["alpha_coeff"] = alpha
            med["alpha_power"] = y
            if mode:
                med["alpha_mode"] = mode
        source = {"p0": scale * np.asarray(self._padded(src.initialPressure)) *
                  np.ones(shp)}
        record = ["p_final"]
        sensor = None
        if sens is not None and sens.mask is not None and sens.totalSensorPoints > 0:
            mask = np.zeros(shp, dtype=bool)
            mask[tuple(slice(p, p + n) for p, n in zip(g.gridPadding, g.shape))] = sens.mask
            if sens.pressureSensor == "on":
                record.append("p")
            if sens.velocitySensor == "on":
                record.append("u")
            elif sens.velocitySensor == "ongrid":
                record.append("u_non_staggered")
            sensor = {"mask": mask, "record": record}
        opts = _parse_options((), {"PMLInside": True, "PMLSize": list(g.gridPadding[:dim]),
                                   "PMLAlpha": [float(a) for a in self.pml.pmlAlpha[:dim]],
                                   "Smooth": False, "DataCast": "single"})
        pp = prepare(kg, med, source, sensor if sensor is not None else {"record": ["p_final"]},
                     dim, opts)
        pp.prob.record_flags |= _abi.REC["p_final"]
        if sensor is not None and sens.densitySensor == "on":
            pp.prob.record_flags |= _abi.KW_REC_RHO
        self._pp = pp
        self._lib = _abi.load(pp.info.backend)
        h = C.c_void_p()
        _abi.call(self._lib.kw_sim_create, C.byref(pp.prob), C.byref(h))
        self._sim = h
        self._nsens = int(pp.prob.sensor_num_points)

    def _destroy(self):
        if self._sim is not None and self._lib is not None:
            self._lib.kw_sim_destroy(self._sim)
        self._sim = None

    def close(self):
        self._destroy()

    def __del__(self):
        try:
            self._destroy()
        except Exception:
            pass

    def _staggered_velocity(self, scale):
        """stagger.m 'fourier' 'forward' of each initial velocity component
        along its own axis, on the padded grid (executeTimeStep.m:134-139)."""
        g = self.kgrid
        u0 = np.asarray(self.source.initialVelocity, dtype=np.float64)
        out = []
        for a in range(g.dimensions):
            comp = u0[..., a] if u0.ndim > 0 else u0
            f = np.asarray(self._padded(comp)) * np.ones(self._padded_shape())
            n = f.shape[a]
            d = g.gridSpacing[a]
            k = np.fft.ifftshift(_wavenumbers(n, d))
            shp = [1] * f.ndim
            shp[a] = -1
            sh = np.exp(1j * k * d / 2).reshape(shp)
            out.append(scale * np.real(np.fft.ifftn(sh * np.fft.fftn(f))))
        return out

    # -- run -----------------------------------------------------------------
    def run(self, Nt=None, dt=None, CFL=None, EndTime=None):
        if (Nt is None) != (dt is None):
            _error("mustBeAllOrNoneEmpty", "Nt and dt must both be given or both be empty.")
        if Nt is None:
            Nt, dt = self._auto_time_step(CFL, EndTime)
        Nt = int(Nt)
        if Nt < 0 or not (float(dt) > 0) or not math.isfinite(float(dt)):
            _error("run:invalidInput", "Nt must be a non-negative integer and dt positive.")
        dt = float(dt)
        sens = self.sensor
        first = self.timeStepsTaken == 0
        ts = 1
        if sens is not None:
            ts = int(sens.timeSteps)
            if ts < 1:
                _error("Sensor:invalidTimeSteps", "timeSteps must be a positive integer.")
        if first:
            # run.m:81-86 (a run after run(Nt=0) starts over: the initial
            # condition is added to the state it left, which is the initial
            # condition itself, so it is that of the scaled inputs)
            self._ic_count += 1
            self._destroy()
            self._create(Nt + 1, dt, float(self._ic_count))
            if self.source.initialVelocity is not None:
                u = self._staggered_velocity(float(self._ic_count))
                arrs = [np.ascontiguousarray(np.reshape(v, -1, order="F")) for v in u]
                ptrs = (_abi.c_double_p * 3)(*([a.ctypes.data_as(_abi.c_double_p) for a in arrs] +
                                               [None] * (3 - len(arrs))))
                _abi.call(self._lib.kw_sim_init_velocity, self._sim, ptrs)
                if Nt > 0:
                    _abi.call(self._lib.kw_sim_change_dt, self._sim, C.c_double(0.0),
                              C.c_double(dt))
                    _abi.call(self._lib.kw_sim_step, self._sim, C.c_int64(Nt))
            else:
                _abi.call(self._lib.kw_sim_step, self._sim, C.c_int64(Nt + 1))
            start = 0  # a new device simulation
            self._steps = Nt + 1
            if sens is not None:
                # initialiseSensorData (AcousticSensor.m:66-124): Nt/timeStepSpacing
                # + 1 columns, or Nt/timeStepSpacing more after a run(Nt=0)
                tss = int(sens.timeStepSpacing)
                self._width = (self._width + Nt // tss) if self._records else Nt // tss + 1
            picks = [(j, j) for j in range(Nt + 1) if j % ts == 0]
            base = float(self.timeArray[-1]) if self.timeArray.size else None
        else:
            total = self._steps + Nt
            _abi.call(self._lib.kw_sim_extend, self._sim, C.c_int64(total))
            if np.float32(dt) != np.float32(self.prevTimeStep):
                _abi.call(self._lib.kw_sim_change_dt, self._sim, C.c_double(self.prevTimeStep),
                          C.c_double(dt))
            _abi.call(self._lib.kw_sim_step, self._sim, C.c_int64(Nt))
            start = self._steps
            self._steps = total
            # later runs preallocate nothing: the arrays grow as records come
            picks = [(j, j + 1) for j in range(Nt) if (j + 1) % ts == 0]
            base = float(self.timeArray[-1])
        if sens is not None and self._nsens > 0:
            self._collect(start, picks, base, np.float32(dt))
        # run.m:91-100
        self.prevTimeStep = float(np.float32(dt))
        self.timeStepsTaken += Nt
        step = np.float32(dt)
        new = np.arange(0 if not self.timeArray.size else 1, Nt + 1, dtype=np.float32) * step
        if self.timeArray.size:
            new = self.timeArray[-1] + new
        self.timeArray = np.concatenate([self.timeArray, new.astype(np.float32)])

    def _collect(self, start, picks, base, dt32):
        """Copy this run's records out of the device time series and keep the
        reference's: column `start + j` for each (j, k) of picks, at time
        base + dt k (AcousticSensor.m:66-173, executeTimeStep.m:244-252)."""
        sens, lib = self.sensor, self._lib
        ns, n_rec = self._nsens, self._steps
        out = _abi.Outputs()
        bufs = {}

        def want(name):
            a = np.zeros(ns * n_rec, dtype=np.float32)
            bufs[name] = a
            return a.ctypes.data_as(_abi.c_float_p)

        if sens.pressureSensor == "on":
            out.p = want("p")
        if sens.densitySensor == "on":
            out.rho = want("rho")
        dim = self.dimensions
        if sens.velocitySensor in ("on", "ongrid"):
            for a in range(dim):
                ptr = want(f"u{a}")
                if sens.velocitySensor == "on":
                    out.u[a] = ptr
                else:
                    out.u_non_staggered[a] = ptr
        _abi.call(lib.kw_sim_read, self._sim, C.byref(out))
        cols = [start + j for j, _ in picks]
        # timePoint in single: timeArray(end) + dt * k, or dt * k on the first run
        b32 = np.float32(0.0 if base is None else base)
        times = [float(b32 + dt32 * np.float32(k)) if base is not None else
                 float(dt32 * np.float32(k)) for _, k in picks]

        def grow(old, new_cols, tail):
            w = max(self._width, self._records + new_cols.shape[-1])
            arr = np.zeros(tail + (w,), dtype=np.float32)
            if old is not None:
                k = min(old.shape[-1], w)
                arr[..., :k] = old[..., :k]
            arr[..., self._records:self._records + new_cols.shape[-1]] = new_cols
            return arr

        if "p" in bufs:
            v = bufs["p"].reshape(ns, n_rec, order="F")[:, cols]
            sens.pressure = grow(sens.pressure, v, (ns,))
        if "rho" in bufs:
            v = bufs["rho"].reshape(ns, n_rec, order="F")[:, cols]
            sens.density = grow(sens.density, v, (ns,))
        if "u0" in bufs:
            v = np.stack([bufs[f"u{a}"].reshape(ns, n_rec, order="F")[:, cols]
                          for a in range(dim)], 1)
            sens.velocity = grow(sens.velocity, v, (ns, dim))
        t = np.asarray(times, dtype=np.float32)[None, :]
        sens.times = grow(None if sens.times is None else sens.times[None, :], t, (1,))[0]
        self._records += len(cols)

    # -- fields (AcousticSolver.m:188-205) -------------------------------------
    def _fields(self):
        if self._sim is None:
            return None
        shp = self._padded_shape()
        n = int(np.prod(shp))
        dim = self.dimensions
        p = np.zeros(n, dtype=np.float32)
        u = [np.zeros(n, dtype=np.float32) for _ in range(dim)]
        r = [np.zeros(n, dtype=np.float32) for _ in range(dim)]
        up = (_abi.c_float_p * 3)(*([a.ctypes.data_as(_abi.c_float_p) for a in u] +
                                    [None] * (3 - dim)))
        rp = (_abi.c_float_p * 3)(*([a.ctypes.data_as(_abi.c_float_p) for a in r] +
                                    [None] * (3 - dim)))
        _abi.call(self._lib.kw_sim_get_fields, self._sim, p.ctypes.data_as(_abi.c_float_p), up,
                  rp)
        f = lambda a: self._interior(a.reshape(shp, order="F"))  # noqa: E731
        return f(p), np.stack([f(a) for a in u], -1), np.stack([f(a) for a in r], -1)

    @property
    def pressure(self):
        f = self._fields()
        return None if f is None else f[0]

    @property
    def velocity(self):
        f = self._fields()
        return None if f is None else f[1]

    @property
    def densitySplit(self):
        f = self._fields()
        return None if f is None else f[2]


def _wavenumbers(n, d):
    """Grid/getWavenumbers.m."""
    if n == 1:
        return np.zeros(1)
    if n % 2 == 0:
        nx = np.arange(-n // 2, n // 2) / n
    else:
        nx = np.arange(-(n - 1) // 2, (n - 1) // 2 + 1) / n
    nx[n // 2] = 0.0
    return (2 * math.pi / d) * nx
