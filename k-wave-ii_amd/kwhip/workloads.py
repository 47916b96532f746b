"""Synthetic workloads of the benchmark configurations, as reference-style
(kgrid, medium, source, sensor, options) inputs.

* ``circle_2d``: 2D 128x128, dx = dy = 0.1 mm, homogeneous lossless medium, PML 20
  (alpha 2), smoothed disc p0 of radius 5 at the centre, CFL 0.3,
  t_end = 1.5 x diagonal / c0, 50 Cartesian sensor points on a circle of
  radius 40 grid points recording p.
* ``disc_source_3d``: 3D N^3 (N = 256 in the benchmark), dx = 0.1 mm, sound speed
  1500 + 100 g clipped to [1400, 1600] m/s and density 1000 + 100 g clipped to
  [900, 1100] kg/m^3 with g Gaussian-smoothed (sigma 4 points) uniform noise
  (seeds 0 and 1), power-law absorption 0.75 dB/(MHz^1.5 cm), PML 10, a 1 MHz
  10-cycle Hann-windowed 1 Pa tone burst on a disc of radius 30 points in the
  plane z = PML + 2, CFL 0.3, 1000 steps, full-grid p_max and p_rms.
* ``nonlinear_bowl_3d``: 3D N^3 (N = 384 in the precision study), the disc_source_3d medium recipe
  with B/A = 6, a 0.5 MHz 5-cycle Hann-windowed 1 MPa tone burst on a
  spherical-cap bowl (radius of curvature 150 points, aperture 120 points
  at N = 384, scaled with N) whose rear sits on the axis at z = PML + 2,
  facing +z, CFL 0.3, 2000 steps; the sensor is a 64 x 64 plane through the
  bowl's geometric focus, recording p (and p_max, p_rms).
* ``batch_tr_problem``: one problem of the batched photoacoustic workload:
  a 2D grid from 256^2 to 1024^2 or a 3D grid from 64^3 to 192^3 (sizes
  2^a 3^b 5^c), homogeneous c0 ~ U[1450, 1600] m/s, p0 = 5-20 random discs
  (2D) or tubes (3D) of radius 1-4 points and amplitude U[0.5, 1], a line
  (2D) or plane (3D) sensor on one face just inside the PML, time from
  makeTime (CFL 0.3); the forward run records p, the time-reversal run
  re-injects it.
* ``served_problem``: one request of the served workload: a 2D grid from
  256^2 to 1024^2 or a 3D grid from 96^3 to 256^3 with the heterogeneous
  medium recipe (smoothed-noise c and rho, power-law absorption y = 1.5,
  B/A = 6) and a 1 MHz tone-burst disc source, a given number of steps,
  p_max and p_rms on the whole grid.
"""

from __future__ import annotations

import math

import numpy as np

from .kgrid import kWaveGrid


def smoothed_noise(shape, seed, sigma):
    """U(-1, 1) noise filtered with a periodic Gaussian of standard deviation
    `sigma` grid points, normalised to max |g| = 1 (float64)."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(-1.0, 1.0, size=shape)
    G = np.fft.rfftn(g)
    ks = [np.fft.fftfreq(n) for n in shape[:-1]] + [np.fft.rfftfreq(shape[-1])]
    filt = np.ones(G.shape)
    for ax, k in enumerate(ks):
        sh = [1] * len(shape)
        sh[ax] = -1
        filt = filt * np.exp(-2.0 * (math.pi ** 2) * (sigma ** 2) * k.reshape(sh) ** 2)
    g = np.fft.irfftn(G * filt, s=shape, axes=tuple(range(len(shape))))
    return g / np.max(np.abs(g))


def disc_mask(Nx, Ny, cx, cy, radius):
    """makeDisc-style binary disc (1-based centre, distance <= radius)."""
    x = np.arange(1, Nx + 1)[:, None]
    y = np.arange(1, Ny + 1)[None, :]
    return ((x - cx) ** 2 + (y - cy) ** 2 <= radius ** 2).astype(np.float64)


def cart_circle(radius, num_points):
    """makeCartCircle (makeCartCircle.m:81-99): angles (0:N-1) 2 pi/N + pi/2,
    points [r cos a; -r sin a]."""
    a = np.arange(num_points) * 2 * math.pi / num_points + math.pi / 2
    return np.stack([radius * np.cos(a), -radius * np.sin(a)])


def circle_2d(N=128):
    dx = 1e-4
    c0, rho0 = 1500.0, 1000.0
    kg = kWaveGrid(N, dx, N, dx)
    t_end = 1.5 * math.sqrt(2.0) * N * dx / c0
    kg.makeTime(c0, 0.3, t_end)
    medium = {"sound_speed": c0, "density": rho0}
    source = {"p0": disc_mask(N, N, N // 2 + 1, N // 2 + 1, 5)}
    sensor = {"mask": cart_circle(40 * dx, 50)}
    opts = {"PMLSize": 20, "PMLAlpha": 2.0, "CartInterp": "nearest"}
    return kg, medium, source, sensor, opts


def disc_source_3d(N=256, Nt=1000, pml=10, src_radius=30, seeds=(0, 1)):
    dx = 1e-4
    shape = (N, N, N)
    c = np.clip(1500.0 + 100.0 * smoothed_noise(shape, seeds[0], 4.0), 1400.0, 1600.0)
    rho = np.clip(1000.0 + 100.0 * smoothed_noise(shape, seeds[1], 4.0), 900.0, 1100.0)
    kg = kWaveGrid(N, dx, N, dx, N, dx)
    dt = 0.3 * dx / float(c.max())
    kg.setTime(Nt, dt)
    medium = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5}
    f0, cycles = 1e6, 10
    n_sig = int(math.floor(cycles / f0 / dt)) + 1
    t = np.arange(n_sig) * dt
    sig = np.sin(2 * math.pi * f0 * t) * np.hanning(n_sig)
    mask = np.zeros(shape)
    mask[:, :, pml + 1] = disc_mask(N, N, N // 2 + 1, N // 2 + 1, src_radius)
    source = {"p": sig[None, :], "p_mask": mask, "p_mode": "additive"}
    sensor = {"mask": np.ones(shape, dtype=bool), "record": ["p_max", "p_rms"]}
    opts = {"PMLSize": pml, "PMLAlpha": 2.0}
    return kg, medium, source, sensor, opts


def bowl_mask(shape, apex, radius, diameter, focus):
    """Binary spherical cap ("bowl") on a grid: the cells within half a cell
    of the sphere of the given radius of curvature whose centre lies at
    `radius` from `apex` towards `focus`, limited to the cap of aperture
    `diameter` around the apex (0-based cell coordinates, lengths in
    cells)."""
    apex = np.asarray(apex, dtype=np.float64)
    axis = np.asarray(focus, dtype=np.float64) - apex
    axis = axis / np.linalg.norm(axis)
    centre = apex + radius * axis
    half = diameter / 2.0
    if half > radius:
        raise ValueError("the aperture must not exceed twice the radius of curvature")
    # the cap: points of the sphere at most `depth` from the apex along the axis
    depth = radius - math.sqrt(radius * radius - half * half)
    lo = [max(0, int(math.floor(min(apex[a], centre[a]) - radius - 1))) for a in range(3)]
    hi = [min(shape[a], int(math.ceil(max(apex[a], centre[a]) + radius + 2))) for a in range(3)]
    g = np.meshgrid(*[np.arange(lo[a], hi[a], dtype=np.float64) for a in range(3)], indexing="ij")
    d = [g[a] - centre[a] for a in range(3)]
    r = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
    along = sum((g[a] - apex[a]) * axis[a] for a in range(3))
    sel = (np.abs(r - radius) <= 0.5) & (along <= depth + 0.5)
    mask = np.zeros(shape)
    mask[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = sel
    return mask


def nonlinear_bowl_3d(N=384, Nt=2000, pml=10, seeds=(0, 1), plane=64):
    """Precision-study workload (nonlinear absorbing, bowl source); returns
    (kgrid, medium, source, sensor, opts) and the 0-based z index of the
    sensor plane."""
    dx = 1e-4
    shape = (N, N, N)
    c = np.clip(1500.0 + 100.0 * smoothed_noise(shape, seeds[0], 4.0), 1400.0, 1600.0)
    rho = np.clip(1000.0 + 100.0 * smoothed_noise(shape, seeds[1], 4.0), 900.0, 1100.0)
    kg = kWaveGrid(N, dx, N, dx, N, dx)
    dt = 0.3 * dx / float(c.max())
    kg.setTime(Nt, dt)
    medium = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5,
              "BonA": 6.0}
    f0, cycles = 0.5e6, 5
    n_sig = int(math.floor(cycles / f0 / dt)) + 1
    t = np.arange(n_sig) * dt
    sig = 1e6 * np.sin(2 * math.pi * f0 * t) * np.hanning(n_sig)
    # radius of curvature and aperture (150 and 120 points at N = 384),
    # rounded half away from zero
    radius = math.floor(150.0 * N / 384 + 0.5)
    diameter = math.floor(120.0 * N / 384 + 0.5)
    cx = N // 2
    apex = (cx, cx, pml + 1)
    mask = bowl_mask(shape, apex, radius, diameter, (cx, cx, N - 1))
    source = {"p": sig[None, :], "p_mask": mask, "p_mode": "additive"}
    zf = pml + 1 + radius
    if zf >= N - pml:
        raise ValueError("grid too small for the bowl")
    m = np.zeros(shape, dtype=bool)
    h = plane // 2
    m[cx - h:cx - h + plane, cx - h:cx - h + plane, zf] = True
    sensor = {"mask": m, "record": ["p", "p_max", "p_rms"]}
    opts = {"PMLSize": pml, "PMLAlpha": 2.0}
    return kg, medium, source, sensor, opts, zf


def fft_friendly(lo, hi):
    """Sizes 2^a 3^b 5^c in [lo, hi], sorted."""
    out = []
    for n in range(lo, hi + 1):
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return out


SIZES_2D = fft_friendly(256, 1024)
SIZES_3D = fft_friendly(64, 192)
SIZES_3D_SERVED = fft_friendly(96, 256)


def _random_p0(rng, shape):
    """5-20 discs (2D) or tubes along a random axis (3D), radius 1-4 points,
    amplitude U[0.5, 1]."""
    dim = len(shape)
    p0 = np.zeros(shape)
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    for _ in range(int(rng.integers(5, 21))):
        r = rng.uniform(1.0, 4.0)
        amp = rng.uniform(0.5, 1.0)
        c = [rng.uniform(0.25 * n, 0.75 * n) for n in shape]
        if dim == 2:
            d2 = (grids[0] - c[0]) ** 2 + (grids[1] - c[1]) ** 2
        else:
            ax = int(rng.integers(0, 3))
            others = [a for a in range(3) if a != ax]
            d2 = sum((grids[a] - c[a]) ** 2 for a in others)
            lo, hi = sorted(rng.uniform(0.2 * shape[ax], 0.8 * shape[ax], size=2))
            d2 = np.where((grids[ax] >= lo) & (grids[ax] <= hi), d2, np.inf)
        p0 = np.where(d2 <= r * r, np.maximum(p0, amp), p0)
    return p0


def batch_tr_problem(seed, dim=None):
    """(kgrid, medium, source, sensor, opts) of one batched forward problem;
    the time-reversal run uses the same grid and medium with
    sensor.time_reversal_boundary_data set to the forward run's p."""
    rng = np.random.default_rng(seed)
    if dim is None:
        dim = int(rng.integers(2, 4))
    N = int(rng.choice(SIZES_2D if dim == 2 else SIZES_3D))
    shape = (N,) * dim
    dx = 1e-4
    c0 = float(rng.uniform(1450.0, 1600.0))
    pml = 20 if dim == 2 else 10
    kg = kWaveGrid(*sum(([n, dx] for n in shape), []))
    kg.makeTime(c0, 0.3)
    medium = {"sound_speed": c0, "density": 1000.0}
    source = {"p0": _random_p0(rng, shape)}
    mask = np.zeros(shape, dtype=bool)
    mask[pml] = True  # the face x = PML + 1 (1-based)
    sensor = {"mask": mask}
    opts = {"PMLSize": pml, "PMLAlpha": 2.0, "Smooth": False}
    return kg, medium, source, sensor, opts


def served_problem(seed, steps, dim=None, kind=None):
    """(kgrid, medium, source, sensor, opts) of one served request: a forward
    simulation of a tone burst from a disc (line) source recording the
    whole-grid p_max and p_rms, or (three requests in ten) a time-reversal
    reconstruction re-emitting bursts from that disc; kind = "forward" or
    "reconstruction" fixes which."""
    rng = np.random.default_rng(seed)
    if dim is None:
        dim = int(rng.integers(2, 4))
    N = int(rng.choice(SIZES_2D if dim == 2 else SIZES_3D_SERVED))
    shape = (N,) * dim
    dx = 1e-4
    c = np.clip(1500.0 + 100.0 * smoothed_noise(shape, int(rng.integers(1 << 30)), 4.0),
                1400.0, 1600.0)
    rho = np.clip(1000.0 + 100.0 * smoothed_noise(shape, int(rng.integers(1 << 30)), 4.0),
                  900.0, 1100.0)
    kg = kWaveGrid(*sum(([n, dx] for n in shape), []))
    dt = 0.3 * dx / float(c.max())
    kg.setTime(int(steps), dt)
    medium = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5,
              "BonA": 6.0}
    f0 = 1e6
    n_sig = min(int(steps), int(math.floor(10 / f0 / dt)) + 1)
    t = np.arange(n_sig) * dt
    sig = 1e6 * np.sin(2 * math.pi * f0 * t) * np.hanning(n_sig)
    pml = 20 if dim == 2 else 10
    mask = np.zeros(shape)
    r = max(4, N // 8)
    if dim == 2:
        mask[pml + 1, N // 2 - r:N // 2 + r + 1] = 1
    else:
        mask[:, :, pml + 1] = disc_mask(N, N, N // 2 + 1, N // 2 + 1, r)
    opts = {"PMLSize": pml, "PMLAlpha": 2.0, "Smooth": False}
    tr = rng.random() < 0.3
    if kind is not None:
        tr = kind == "reconstruction"
    if tr:
        # a time-reversal reconstruction: the disc (line) re-emits recorded
        # bursts, one per point with its own delay
        n_pts = int(np.count_nonzero(mask))
        delays = rng.integers(0, max(1, int(steps) // 3), size=n_pts)
        data = np.zeros((n_pts, int(steps)))
        for i, dl in enumerate(delays):
            m = min(n_sig, int(steps) - int(dl))
            data[i, dl:dl + m] = 1e5 * sig[:m] / 1e6
        sensor = {"mask": mask.astype(bool), "time_reversal_boundary_data": data}
        return kg, medium, {}, sensor, opts
    source = {"p": sig[None, :], "p_mask": mask, "p_mode": "additive"}
    sensor = {"mask": np.ones(shape, dtype=bool), "record": ["p_max", "p_rms"]}
    return kg, medium, source, sensor, opts
