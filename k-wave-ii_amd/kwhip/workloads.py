"""Synthetic workloads of the benchmark configurations, as reference-style
(kgrid, medium, source, sensor, options) inputs.

* ``cfg1``: 2D 128x128, dx = dy = 0.1 mm, homogeneous lossless medium, PML 20
  (alpha 2), smoothed disc p0 of radius 5 at the centre, CFL 0.3,
  t_end = 1.5 x diagonal / c0, 50 Cartesian sensor points on a circle of
  radius 40 grid points recording p.
* ``cfg2``: 3D N^3 (N = 256 in the benchmark), dx = 0.1 mm, sound speed
  1500 + 100 g clipped to [1400, 1600] m/s and density 1000 + 100 g clipped to
  [900, 1100] kg/m^3 with g Gaussian-smoothed (sigma 4 points) uniform noise
  (seeds 0 and 1), power-law absorption 0.75 dB/(MHz^1.5 cm), PML 10, a 1 MHz
  10-cycle Hann-windowed 1 Pa tone burst on a disc of radius 30 points in the
  plane z = PML + 2, CFL 0.3, 1000 steps, full-grid p_max and p_rms.
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
    g = np.fft.irfftn(G * filt, s=shape)
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


def cfg1(N=128):
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


def cfg2(N=256, Nt=1000, pml=10, src_radius=30, seeds=(0, 1)):
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
