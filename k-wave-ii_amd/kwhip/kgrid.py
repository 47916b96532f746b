"""kWaveGrid: the grid object of the reference, for uniform 1D/2D/3D grids.

Mirrors +kwave/+legacy/kWaveGrid.m: constructor (:342-436), wavenumber
vectors (makeDim, :779-797), position vectors (:494-506), t_array
(:524-530), makeTime (:612-673) and setTime.
"""

from __future__ import annotations

import math

import numpy as np


def _make_dim(N: int, d: float) -> np.ndarray:
    # kWaveGrid.m:779-797: nx = (-N/2:N/2-1)/N (even) or (-(N-1)/2:(N-1)/2)/N,
    # middle value forced to zero, k = (2 pi/dx) nx
    if N % 2 == 0:
        n = np.arange(-(N // 2), N // 2, dtype=np.float64)
    else:
        n = np.arange(-((N - 1) // 2), (N - 1) // 2 + 1, dtype=np.float64)
    nx = n / N
    nx[N // 2] = 0.0
    return (2.0 * math.pi / d) * nx


class kWaveGrid:  # noqa: N801 - reference class name
    """kWaveGrid(Nx, dx), kWaveGrid(Nx, dx, Ny, dy) or
    kWaveGrid(Nx, dx, Ny, dy, Nz, dz).  1-D grid fields are arrays of shape
    (Nx,)."""

    def __init__(self, Nx, dx, Ny=None, dy=None, Nz=None, dz=None):
        self.Nx, self.dx = int(Nx), float(dx)
        if Ny is None or dy is None:
            self.Ny, self.dy = 1, 0.0
            self.dim = 1
        else:
            self.Ny, self.dy = int(Ny), float(dy)
            self.dim = 2
        if Nz is not None and self.dim == 2:
            self.Nz, self.dz = int(Nz), float(dz)
            self.dim = 3
        else:
            self.Nz, self.dz = 1, 0.0
        self.kx_vec = _make_dim(self.Nx, self.dx)
        self.ky_vec = _make_dim(self.Ny, self.dy) if self.dim >= 2 else np.zeros(1)
        self.kz_vec = _make_dim(self.Nz, self.dz) if self.dim == 3 else np.zeros(1)
        self.Nt = "auto"
        self.dt = "auto"

    # -- sizes --------------------------------------------------------------
    @property
    def shape(self):
        return ((self.Nx,), (self.Nx, self.Ny), (self.Nx, self.Ny, self.Nz))[self.dim - 1]

    @property
    def total_grid_points(self):
        return int(np.prod(self.shape))

    @property
    def spacings(self):
        return (self.dx, self.dy, self.dz)[:self.dim]

    @property
    def x_size(self):
        return self.Nx * self.dx

    @property
    def y_size(self):
        return self.Ny * self.dy

    @property
    def z_size(self):
        return self.Nz * self.dz

    @property
    def x_vec(self):
        return self.x_size * self.kx_vec * self.dx / (2 * math.pi)

    @property
    def y_vec(self):
        return self.y_size * self.ky_vec * self.dy / (2 * math.pi)

    @property
    def z_vec(self):
        return self.z_size * self.kz_vec * self.dz / (2 * math.pi)

    @property
    def k_max(self):
        vecs = [self.kx_vec, self.ky_vec, self.kz_vec][:self.dim]
        return min(float(np.max(np.abs(v))) for v in vecs)

    # -- time ---------------------------------------------------------------
    @property
    def t_array(self):
        if isinstance(self.Nt, str) or isinstance(self.dt, str):
            return "auto"
        return np.arange(self.Nt) * self.dt

    @t_array.setter
    def t_array(self, t):
        if isinstance(t, str) and t == "auto":
            self.Nt, self.dt = "auto", "auto"
            return
        t = np.asarray(t, dtype=np.float64).ravel()
        self.setTime(len(t), float(t[1] - t[0]) if len(t) > 1 else 0.0)

    def setTime(self, Nt, dt):  # noqa: N802 - reference method name
        self.Nt, self.dt = int(Nt), float(dt)

    def makeTime(self, c, cfl=0.3, t_end=None):  # noqa: N802 - reference method name
        """kWaveGrid.m:612-673: dt from the CFL number and the largest sound
        speed, Nt from t_end (default: grid diagonal at the smallest speed)."""
        c = np.asarray(c, dtype=np.float64)
        c_max, c_min = float(c.max()), float(c.min())
        if t_end is None:
            sizes = [n * d for n, d in zip(self.shape, self.spacings)]
            t_end = math.sqrt(sum(s * s for s in sizes)) / c_min
        dt = cfl * min(self.spacings) / c_max
        Nt = math.floor(t_end / dt) + 1
        ratio = t_end / dt
        if math.floor(ratio) != math.ceil(ratio) and (t_end - math.trunc(ratio) * dt) == 0:
            Nt += 1
        self.Nt, self.dt = int(Nt), float(dt)
        return self.t_array, self.dt
