"""CPU restatement of the new toolbox's acoustic solver (kwave.toolbox.
AcousticSolver) for tests.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the
kwhip.toolbox adapter, never by the product path.

Restates, in NumPy with whole-grid FFTs as the reference does:

* AcousticSolver/executeTimeStep.m:35-260 (the time loop: initial
  conditions, the transition step after a change of time step, the usual
  step, the absorption terms, sensor recording);
* AcousticSolver/setkSpaceCorrection.m:36-57 (κ, and the two-step κ and
  kappaSplit when the step changes), kappaSplitCorrection.m:48-62;
* AcousticSolver/setAbsorptionCoefficients.m:38-70 (τ, η with the dB
  conversion), executeTimeStep.m:67-92 (NaN-safe operators);
* FourierCollocation.m:117-160 (shift-derivative operators on the padded
  grid), gradient.m:60-108, divergenceSplit.m:64-113, fracLaplacian.m:63-90,
  stagger.m:60-130 ('fourier' and 'linInterpolate'), sinc.m;
* SplitFieldPML/getQuarticPMLProfile.m:60-92, setupQuarticPML.m, applyPML.m;
* Grid/getWavenumbers.m, Grid/assignWithGridPadding.m with expandMatrix.m
  (grid inputs padded by repeating their edge values);
* TimeDomainSolver/run.m:58-101 (repeated runs, prevTimeStep in single
  precision, timeArray) and AcousticSensor.m:66-173 / Sensor.m:73-99
  (recorded pressure, summed density, staggered or 'ongrid' velocity).

Parity pins (tests/test_toolbox_oracle.py): the first run equals the legacy
oracle run for Nt + 1 steps with the PML outside the grid
(TestAcousticSolverGrid.m:74-113 makes the same comparison); restarts with a
new time step equal one run at the first step in a homogeneous medium
(TestAcousticSolver.m:455-616, the k-space scheme being exact there); the
initial-velocity path matches d'Alembert's solution
(TestAcousticSolver.m:769-806); the recorded sensor data equal the fields at
the mask (TestAcousticSolver.m:1039-1082).
"""

import math

import numpy as np


def wavenumbers(n, d):
    """Grid/getWavenumbers.m: centred wavenumbers of an axis (the Nyquist
    term first for even n; the centre entry 0)."""
    if n == 1:
        return np.zeros(1)
    if n % 2 == 0:
        nx = np.arange(-n // 2, n // 2) / n
    else:
        nx = np.arange(-(n - 1) // 2, (n - 1) // 2 + 1) / n
    nx[n // 2] = 0.0
    return (2 * math.pi / d) * nx


def sinc(x):
    """FourierCollocation/sinc.m."""
    z = x == 0
    return np.sin(x + math.pi * z) / (x + z) + z


def quartic_pml(n, d, dt, c, size, alpha, staggered):
    """SplitFieldPML/getQuarticPMLProfile.m."""
    prof = np.ones(n)
    if size == 0:
        return prof
    x = np.arange(1, size + 1, dtype=np.float64)
    if staggered:
        left = alpha * (c / d) * (((x + 0.5) - size - 1) / (0 - size)) ** 4
        right = alpha * (c / d) * ((x + 0.5) / size) ** 4
    else:
        left = alpha * (c / d) * ((x - size - 1) / (0 - size)) ** 4
        right = alpha * (c / d) * (x / size) ** 4
    prof[:size] = np.exp(-left * dt / 2)
    prof[n - size:] = np.exp(-right * dt / 2)
    return prof


def pad_edges(a, pad, dim):
    """Grid/assignWithGridPadding.m with expandMatrix.m (edge values repeated);
    scalars stay scalars."""
    a = np.asarray(a, dtype=np.float64)
    if a.size == 1 or not any(pad[:dim]):
        return a
    return np.pad(a, [(p, p) for p in pad[:dim]], mode="edge")


class ToolboxOracle:
    """One AcousticSolver: grid_size / spacing / padding per dimension,
    medium values (scalars or arrays of grid_size), initial pressure and
    optional initial velocity (a list of dim arrays), an optional binary
    sensor mask with the sensors on, 'single' or 'double' arithmetic."""

    def __init__(self, grid_size, spacing, padding, sound_speed, density, p0, u0=None,
                 absorption_coeff=None, absorption_power=None, absorption_type="off",
                 sound_speed_ref=None, pml_alpha=2.0, data_type="double", mask=None,
                 pressure_sensor="on", velocity_sensor="off", density_sensor="off",
                 time_steps=1, time_step_spacing=1):
        self.dim = len(grid_size)
        self.N = list(grid_size)
        self.pad = list(padding) if np.ndim(padding) else [int(padding)] * self.dim
        self.d = list(spacing) if np.ndim(spacing) else [float(spacing)] * self.dim
        self.Np = [n + 2 * p for n, p in zip(self.N, self.pad)]
        self.dt_f = np.float32 if data_type == "single" else np.float64
        self.shape3 = tuple(self.Np + [1] * (3 - self.dim))
        r = lambda a: self._f(np.reshape(pad_edges(a, self.pad, self.dim), -1, order="F")  # noqa: E731
                              .reshape(self.shape3, order="F")) if np.size(a) > 1 else self._f(a)
        self.c = r(sound_speed)
        self.rho = r(density)
        self.p0 = r(p0)
        self.u0 = None if u0 is None else [r(u) for u in u0]
        self.c_ref = float(sound_speed_ref) if sound_speed_ref is not None else float(np.max(self.c))
        self.alpha = None if absorption_coeff is None else r(absorption_coeff)
        self.y = absorption_power
        self.absorption_type = absorption_type
        self.pml_alpha = [pml_alpha] * 3 if np.ndim(pml_alpha) == 0 else list(pml_alpha)
        # padded wavenumbers (FourierCollocation.m:117-160)
        self.kvec = [wavenumbers(n, d) for n, d in zip(self.Np, self.d)]
        kk = np.zeros(self.shape3)
        for a in range(self.dim):
            kk = kk + self._along(self.kvec[a] ** 2, a)
        self.k = np.sqrt(kk)
        self.dpos, self.dneg, self.spos, self.sneg = [], [], [], []
        for a in range(self.dim):
            kv, d = self.kvec[a], self.d[a]
            self.dpos.append(self._along(np.fft.ifftshift(1j * kv * np.exp(1j * kv * d / 2)), a))
            self.dneg.append(self._along(np.fft.ifftshift(1j * kv * np.exp(-1j * kv * d / 2)), a))
            self.spos.append(self._along(np.fft.ifftshift(np.exp(1j * kv * d / 2)), a))
            self.sneg.append(self._along(np.fft.ifftshift(np.exp(-1j * kv * d / 2)), a))
        self.rho_sg = self._stagger_lin(self.rho) if np.size(self.rho) > 1 else \
            [self.rho] * self.dim
        # state (executeTimeStep; setInitialConditions.m)
        self.p = self._f(np.zeros(self.shape3))
        self.rho_split = [self._f(np.zeros(self.shape3)) for _ in range(self.dim)]
        self.u = [self._f(np.zeros(self.shape3)) for _ in range(self.dim)]
        self.prev_dt = None
        self.steps_taken = 0
        self.time_array = []
        # sensor
        self.mask = None if mask is None else np.asarray(mask, dtype=bool)
        self.sensors = (pressure_sensor, velocity_sensor, density_sensor)
        self.time_steps, self.tss = time_steps, time_step_spacing
        self.rec_p, self.rec_u, self.rec_rho, self.times = [], [], [], []
        self.width = 0
        self.time_point = 0.0

    # -- helpers ------------------------------------------------------------
    def _f(self, a):
        return np.asarray(a, dtype=self.dt_f)

    def _along(self, v, a):
        shp = [1, 1, 1]
        shp[a] = -1
        return np.reshape(v, shp)

    def _ifft(self, x):
        return self._f(np.real(np.fft.ifftn(x, axes=(0, 1, 2))))

    def _fft(self, x):
        return np.fft.fftn(x, axes=(0, 1, 2))

    def _stagger_lin(self, f):
        """stagger.m 'linInterpolate' 'forward': midpoints along each axis, the
        last point (outside the grid) keeping its own value."""
        out = []
        for a in range(self.dim):
            g = np.array(f, dtype=np.float64)
            s = [slice(None)] * 3
            s2 = [slice(None)] * 3
            s[a] = slice(0, -1)
            s2[a] = slice(1, None)
            mid = (g[tuple(s)] + g[tuple(s2)]) / 2
            g[tuple(s)] = mid
            out.append(self._f(g))
        return out

    def _kappa(self, dt):
        """setkSpaceCorrection.m:36-57: (κ, kappaSplit)."""
        ck = self.c_ref * self.k
        # dt (double) against the single prevTimeStep: compared in single
        if self.prev_dt is not None and np.float32(dt) != np.float32(self.prev_dt):
            d2, d1 = dt, float(self.prev_dt)
            with np.errstate(divide="ignore", invalid="ignore"):
                ks = 2 * (np.cos(ck * d2 / 2) / np.cos(ck * d1 / 2) - 1) / (d1 + d2)
                kap = 2 * (np.sin(ck * d2 / 2) + np.sin(ck * d1 / 2) *
                           (np.cos(ck * d2 / 2) / np.cos(ck * d1 / 2))) / (ck * (d1 + d2))
            kap[np.isnan(kap)] = 1
            ks[np.isnan(ks)] = 1
            return self._f(np.fft.ifftshift(kap)), self._f(np.fft.ifftshift(ks))
        return self._f(np.fft.ifftshift(sinc(ck * dt / 2))), None

    def _pml(self, dt, staggered):
        return [self._f(self._along(quartic_pml(self.Np[a], self.d[a], dt, self.c_ref,
                                                self.pad[a], self.pml_alpha[a], staggered), a))
                for a in range(self.dim)]

    def _grad(self, f):
        fk = self._fft(f)
        return [self._ifft(self.dpos[a] * self.kappa * fk) for a in range(self.dim)]

    def _div_split(self, u):
        return [self._ifft(self.dneg[a] * self.kappa * self._fft(u[a])) for a in range(self.dim)]

    def _frac_lap(self, f, s):
        with np.errstate(divide="ignore"):
            kp = np.fft.ifftshift(self.k ** (2 * s))
        kp[np.isinf(kp)] = 0
        return self._ifft(kp * self._fft(f))

    def _absorption(self):
        """setAbsorptionCoefficients.m with the NaN-safe operators of
        executeTimeStep.m:67-92: (has, tau, eta, y)."""
        if self.absorption_type == "off" or self.alpha is None:
            return False, None, None, None
        y = float(np.float32(self.y))
        a = 5 * self.alpha * ((1e-6) / (2 * math.pi)) ** y / math.log10(math.e)
        if self.absorption_type == "noDispersion" or y == 0 or y == 2:
            eta = np.zeros(self.shape3)
        else:
            eta = -2 * a * self.c ** y * math.tan(math.pi * y / 2)
        tau = np.zeros(self.shape3) if self.absorption_type == "noAbsorption" else \
            -2 * a * self.c ** (y - 1)
        tau = np.where(np.isfinite(tau), tau, 0) * np.ones(self.shape3)
        eta = np.where(np.isfinite(eta), eta, 0) * np.ones(self.shape3)
        has = bool(np.any((tau != 0) | (eta != 0)))
        return has, self._f(tau), self._f(eta), y

    def _eos(self, dt):
        self.p = self._f(self.c ** 2 * sum(self.rho_split))
        has, tau, eta, y = self._absorption()
        if has:
            div = sum(self._div_split(self.u))
            self.p = self._f(self.p + self.c ** 2 * (
                tau * self._frac_lap(self.rho * div, y / 2 - 1) +
                eta * self._frac_lap(sum(self.rho_split), y / 2 - 0.5)))

    def _mass(self, dt):
        div = self._div_split(self.u)
        self.rho_split = [self._f(self.pml[a] * (self.pml[a] * self.rho_split[a] -
                                                 dt * self.rho * div[a]))
                          for a in range(self.dim)]

    def _momentum(self, dt, correction=None):
        g = self._grad(self.p)
        new = []
        for a in range(self.dim):
            inner = self.pml_sg[a] * self.u[a] - (dt / self.rho_sg[a]) * g[a]
            if correction is not None:
                inner = inner + dt * correction[a]
            new.append(self._f(self.pml_sg[a] * inner))
        self.u = new

    def _interior(self, f):
        s = tuple(slice(p, p + n) for p, n in zip(self.pad, self.N))
        return f[s + (0,) * (3 - self.dim)] if self.dim < 3 else f[s]

    def _record(self):
        if self.mask is None:
            return
        m = self.mask.reshape(-1, order="F")
        pick = lambda f: np.reshape(self._interior(f), -1, order="F")[m]  # noqa: E731
        ps, vs, ds = self.sensors
        if ps == "on":
            self.rec_p.append(pick(self.p).astype(np.float32))
        if ds == "on":
            self.rec_rho.append(sum(pick(r) for r in self.rho_split).astype(np.float32))
        if vs == "on":
            self.rec_u.append(np.stack([pick(u) for u in self.u], 1).astype(np.float32))
        elif vs == "ongrid":
            self.rec_u.append(np.stack([pick(self._ifft(self.sneg[a] * self._fft(self.u[a])))
                                        for a in range(self.dim)], 1).astype(np.float32))
        self.times.append(self.time_point)

    # -- run (TimeDomainSolver/run.m, AcousticSolver/executeTimeStep.m) -------
    def run(self, Nt, dt):
        Nt = int(Nt)
        first = self.steps_taken == 0
        # initialiseSensorData runs only while timeStepsTaken is 0 (run.m:81-86):
        # Nt/timeStepSpacing + 1 columns, or that many more minus one after a
        # run(Nt=0); later runs let the arrays grow as records come
        if first:
            self.width = (self.width + Nt // self.tss) if self.rec_p or self.rec_u or \
                self.rec_rho or self.times else Nt // self.tss + 1
        cur = float(dt)
        # prevTimeStep is single: mixed arithmetic is single
        if self.prev_dt is not None:
            dtv = float((np.float32(cur) + np.float32(self.prev_dt)) / np.float32(2))
        else:
            dtv = cur
        self.kappa, self.ksplit = self._kappa(cur)
        self.pml = self._pml(dtv, False)
        self.pml_sg = self._pml(dtv, True)
        adj = 0
        n_iter = Nt
        if first:
            n_iter += 1
            adj = 1
        for t in range(1, n_iter + 1):
            if t == 1 and first:
                self.p = self._f(self.p + self.p0)
                self.rho_split = [self._f(r + self.p0 / (self.dim * self.c ** 2))
                                  for r in self.rho_split]
                if self.u0 is None:
                    g = self._grad(self.p)
                    self.u = [self._f((dtv / self.rho_sg[a]) * g[a] / 2) for a in range(self.dim)]
                else:
                    for a in range(self.dim):
                        st = self._ifft(self.spos[a] * self._fft(self.u0[a]))
                        self.u[a] = self._f(self.u[a] + st)
                    self.prev_dt = 0.0
                    dtv = cur / 2
                    self.pml = self._pml(dtv, False)
                    self.pml_sg = self._pml(dtv, True)
                    self.kappa, self.ksplit = self._kappa(cur)
            elif (t == 1 and not first) or (t == 2 and first and self.u0 is not None):
                corr = None
                if self.ksplit is not None:
                    pu = [self.pml_sg[a] * self.u[a] for a in range(self.dim)]
                    corr = [self._ifft(self.ksplit * self._fft(pu[a])) for a in range(self.dim)]
                self._momentum(dtv, corr)
                dtv = cur
                self.prev_dt = float(np.float32(cur))
                self.kappa, self.ksplit = self._kappa(cur)
                self.pml = self._pml(cur, False)
                self.pml_sg = self._pml(cur, True)
                self._mass(dtv)
                self._eos(dtv)
            else:
                self._momentum(dtv)
                self._mass(dtv)
                self._eos(dtv)
            if self.mask is not None and (t - adj) % self.time_steps == 0:
                k = np.float32(dtv) * np.float32(t - adj)
                self.time_point = float(np.float32(self.time_array[-1]) + k) if self.time_array \
                    else float(k)
                self._record()
        self.prev_dt = float(np.float32(cur))
        self.steps_taken += Nt
        step = np.float32(cur)
        if not self.time_array:
            self.time_array = [float(np.float32(i) * step) for i in range(Nt + 1)]
        else:
            last = self.time_array[-1]
            self.time_array += [last + float(np.float32(i) * step) for i in range(1, Nt + 1)]

    # -- outputs ------------------------------------------------------------
    @property
    def pressure(self):
        return self._interior(self.p)

    @property
    def velocity(self):
        return np.stack([self._interior(u) for u in self.u], -1)

    @property
    def density_split(self):
        return np.stack([self._interior(r) for r in self.rho_split], -1)

    def sensor_data(self):
        """(pressure [points x T], velocity [points x dim x T], density
        [points x T], times [T]) with T the reference's array width (columns
        past the records stay zero)."""
        def cols(lst, shape_tail):
            w = max(self.width, len(lst))
            if not lst:
                return None
            out = np.zeros(lst[0].shape + (w,), dtype=np.float32)
            for i, v in enumerate(lst):
                out[..., i] = v
            return out
        times = np.zeros(max(self.width, len(self.times)), dtype=np.float32)
        times[:len(self.times)] = self.times
        return cols(self.rec_p, ()), cols(self.rec_u, ()), cols(self.rec_rho, ()), times
