"""Pins of the oracle's axisymmetric solver (kspace_first_order_as):

* a pressure pulse exp(-R^2 / w^2) centred on the axis, with zero particle
  velocity, is a spherical wave whose pressure is [(R - ct) g(R - ct) +
  (R + ct) g(R + ct)] / (2R); the solver follows it inside the grid away
  from the PML to the scheme's discretisation error, which falls as the
  pulse is better resolved (w = 4, 6, 8 points);
* the same run with the sign of the u_r / r term flipped (a plausible
  mistake, planted in the oracle here) is off by order one.
"""

import numpy as np

import as_cases as ac
from oracle import kwave_oracle as ko


def _run(Nx, Ny, dx, c0, Nt, dt, w):
    p0, exact, inside = ac.spherical_wave(Nx, Ny, dx, c0, Nt, dt, w)
    kg = ko.KGrid(Nx, dx, Ny, dx)
    kg.set_time(Nt, dt)
    pf = ko.kspace_first_order_as(kg, {"sound_speed": c0, "density": 1000.0}, {"p0": p0},
                                  {"mask": np.ones((Nx, Ny), bool), "record": ["p_final"]},
                                  pml_size=16, smooth=False)["p_final"]
    return float(np.max(np.abs(pf - exact)[inside]) / np.max(np.abs(exact)))


def test_spherical_wave_converges():
    Nx, Ny, dx, c0 = 128, 64, 1e-4, 1500.0
    dt = 0.2 * dx / c0
    errs = [_run(Nx, Ny, dx, c0, 60, dt, w * dx) for w in (4, 6, 8)]
    assert errs[0] < 1.5e-3 and errs[1] < 6e-4 and errs[2] < 3e-4, errs
    assert errs[0] > errs[1] > errs[2], errs


def test_planted_sign_error_is_caught(monkeypatch):
    orig = ko.as_mirror_u

    def flipped(f, signs=(1, 1, -1, -1)):
        if signs == (1, -1, -1, 1):
            signs = (-1, 1, 1, -1)
        return orig(f, signs)

    monkeypatch.setattr(ko, "as_mirror_u", flipped)
    Nx, Ny, dx, c0 = 128, 64, 1e-4, 1500.0
    assert _run(Nx, Ny, dx, c0, 60, 0.2 * dx / c0, 6 * dx) > 0.1
