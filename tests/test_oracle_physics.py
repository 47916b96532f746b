"""Pins the oracle's time loop against physics it must reproduce (CPU only).

Every check compares with something the oracle does not compute the same
way: a closed-form solution, a textbook law, or a second route through the
solver whose equality follows from the reference's definitions.  Each is
chosen so that a plausible slip fails it (a wrong sign of tau or eta, a
dropped nonlinear term, a factor of 2 in a source scaling, a swapped axis
spacing, an off-by-one in the time index of a source or of the time-reversal
data).

Mirrors the analytic tests of +kwave/+tests/+unit/TestAcousticSolver.m
(plane waves equal across dimensions :49-317, axis permutation :455-614,
d'Alembert initial value problem :769-806).
"""

import math

import numpy as np
import pytest

from oracle import kwave_oracle as ko

W = 2  # FFT workers per oracle run


def run(kg, medium, source, sensor, **kw):
    kw.setdefault("smooth", (False,))
    return ko.kspace_first_order(kg, medium, source, sensor, fft_workers=W, **kw)


def kmag(shape, spacings):
    """|k| in FFT order, built from numpy's fftfreq."""
    acc = 0.0
    for ax, (n, d) in enumerate(zip(shape, spacings)):
        sh = [1] * len(shape)
        sh[ax] = n
        acc = acc + (2 * math.pi * np.fft.fftfreq(n, d)).reshape(sh) ** 2
    return np.sqrt(acc)


# ---------------------------------------------------------------------------
# homogeneous lossless: exact solutions
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("shape,sp", [((48,), (1e-4,)), ((45,), (1.1e-4,)),
                                      ((24, 20), (1e-4, 1.2e-4)),
                                      ((15, 16), (1e-4, 1e-4)),
                                      ((12, 10, 9), (1e-4, 0.9e-4, 1.1e-4))])
def test_periodic_initial_value_is_exact(shape, sp):
    """With the PML switched off the grid is periodic, and k-space time
    stepping with kappa = sinc(c k dt/2) is exact for c = c_ref: step n
    holds P0(k) cos(c |k| (n-1) dt), for every mode including Nyquist."""
    c, Nt, dt = 1500.0, 30, 2.5e-8
    kg = ko.KGrid(*[v for pair in zip(shape, sp) for v in pair])
    kg.set_time(Nt, dt)
    rng = np.random.default_rng(len(shape))
    p0 = rng.standard_normal(shape)
    got = run(kg, {"sound_speed": c, "density": 1000.0}, {"p0": p0},
              {"mask": np.ones(shape)}, pml_size=2, pml_alpha=0.0)
    P0, K = np.fft.fftn(p0), kmag(shape, sp)
    for n in (0, 1, 2, 7, Nt - 1):
        want = np.real(np.fft.ifftn(P0 * np.cos(c * K * n * dt))).ravel(order="F")
        assert np.max(np.abs(got[:, n] - want)) < 1e-10 * np.max(np.abs(p0))


@pytest.mark.parametrize("dim,axis", [(2, 0), (2, 1), (3, 2)])
def test_dalembert_plane_wave(dim, axis):
    """A Gaussian p0 that varies along one axis splits into two half-amplitude
    pulses travelling at c (TestAcousticSolver.m:769-806 without the initial
    velocity, which the legacy solver does not take)."""
    N, dx, c, cfl, Nt, w = 256, 4e-3, 1500.0, 0.25, 150, 5.0
    dt = cfl * dx / c
    shape = [6] * dim
    shape[axis] = N
    args = []
    for n in shape:
        args += [n, dx]
    kg = ko.KGrid(*args)
    kg.set_time(Nt, dt)
    x = (np.arange(N) - N // 2) * dx
    prof = lambda s: np.exp(-(s / (w * dx)) ** 2)   # noqa: E731
    sh = [1] * dim
    sh[axis] = N
    p0 = np.broadcast_to(prof(x).reshape(sh), shape).copy()
    pml_size = [0] * dim
    pml_size[axis] = 20
    pml_alpha = [0.0] * dim
    pml_alpha[axis] = 2.0
    sd = run(kg, {"sound_speed": c, "density": 1000.0}, {"p0": p0},
             {"mask": np.ones(shape), "record": ["p_final"]},
             pml_size=pml_size, pml_alpha=pml_alpha)
    T = (Nt - 1) * dt
    want = 0.5 * (prof(x - c * T) + prof(x + c * T))
    got = np.moveaxis(sd["p_final"], axis, 0).reshape(N, -1)
    assert np.max(np.abs(got - want[:, None])) < 1e-9


def test_dirichlet_pressure_source_sets_pressure_each_step():
    """A dirichlet p source holds p at its points to the signal value of the
    same step (signal column t at step t), for a linear lossless medium."""
    shape = (20, 18)
    kg = ko.KGrid(20, 1e-4, 18, 1e-4)
    kg.set_time(25, 2e-8)
    rng = np.random.default_rng(3)
    pm = np.zeros(shape)
    pm[5, 7] = pm[11, 3] = pm[14, 12] = 1
    sig = rng.standard_normal((3, 25))
    c = 1500 + 50 * rng.random(shape)
    sd = run(kg, {"sound_speed": c, "density": 1000.0},
             {"p_mask": pm, "p": sig, "p_mode": "dirichlet"}, {"mask": pm}, pml_size=3)
    assert np.allclose(sd, sig, rtol=1e-12, atol=1e-12)


def test_dirichlet_velocity_source_sets_velocity_each_step():
    shape = (16, 12, 10)
    kg = ko.KGrid(16, 1e-4, 12, 1e-4, 10, 1e-4)
    kg.set_time(12, 2e-8)
    um = np.zeros(shape)
    um[4, 5, 6] = um[9, 2, 3] = 1
    rng = np.random.default_rng(4)
    ux, uz = rng.standard_normal((2, 12)), rng.standard_normal((2, 12))
    sd = run(kg, {"sound_speed": 1500.0, "density": 1000.0},
             {"u_mask": um, "ux": ux, "uz": uz, "u_mode": "dirichlet"},
             {"mask": um, "record": ["u"]})
    assert np.allclose(sd["ux"], ux, rtol=1e-12, atol=1e-14)
    assert np.allclose(sd["uz"], uz, rtol=1e-12, atol=1e-14)
    assert np.max(np.abs(sd["uy"])) > 0       # uy is driven by the pressure field


# ---------------------------------------------------------------------------
# invariances of the full heterogeneous, absorbing, nonlinear loop
# ---------------------------------------------------------------------------

def _medium(shape, rng, absorbing=True, nonlinear=True):
    c = 1500 + 100 * ko.smooth_mat(rng.standard_normal(shape), True)
    rho = 1000 + 80 * ko.smooth_mat(rng.standard_normal(shape), True)
    med = {"sound_speed": c, "density": rho}
    if absorbing:
        med.update(alpha_coeff=0.75 + 0.2 * rng.random(shape), alpha_power=1.5)
    if nonlinear:
        med["BonA"] = 6 + rng.random(shape)
    return med


def _compact_p0(shape, rng, margin):
    p0 = np.zeros(shape)
    sl = tuple(slice(margin, n - margin) for n in shape)
    p0[sl] = rng.standard_normal(p0[sl].shape)
    return 1e5 * p0


def test_axis_permutation_3d():
    """Relabelling the axes (sizes, spacings, PML settings, medium and
    source together) permutes the result; a swapped spacing, PML axis or
    staggered-density axis breaks it."""
    shape, sp = (12, 10, 14), (1e-4, 1.3e-4, 0.8e-4)
    rng = np.random.default_rng(5)
    med = _medium(shape, rng)
    p0 = _compact_p0(shape, rng, 3)
    pml_size, pml_alpha = (3, 2, 4), (2.0, 1.5, 2.5)
    kg = ko.KGrid(shape[0], sp[0], shape[1], sp[1], shape[2], sp[2])
    kg.set_time(20, 1.5e-8)
    sd = run(kg, med, {"p0": p0}, {"mask": np.ones(shape), "record": ["p_final", "u_final"]},
             pml_size=pml_size, pml_alpha=pml_alpha)
    perm = (2, 0, 1)
    kp = ko.KGrid(*[v for a in perm for v in (shape[a], sp[a])])
    kp.set_time(20, 1.5e-8)
    medp = {k: np.transpose(v, perm) for k, v in med.items() if k != "alpha_power"}
    medp["alpha_power"] = 1.5
    sdp = run(kp, medp, {"p0": np.transpose(p0, perm)},
              {"mask": np.ones(kp.shape), "record": ["p_final", "u_final"]},
              pml_size=[pml_size[a] for a in perm], pml_alpha=[pml_alpha[a] for a in perm])
    ref = np.max(np.abs(sd["p_final"]))
    assert np.max(np.abs(sdp["p_final"] - np.transpose(sd["p_final"], perm))) < 1e-10 * ref
    names = ["ux_final", "uy_final", "uz_final"]
    for new_ax, old_ax in enumerate(perm):
        a, b = sdp[names[new_ax]], np.transpose(sd[names[old_ax]], perm)
        assert np.max(np.abs(a - b)) < 1e-10 * np.max(np.abs(b))


def test_2d_equals_3d_extruded():
    """A 3D problem that does not vary along z (periodic in z, no z PML)
    equals the 2D problem in every z plane, with heterogeneity, power-law
    absorption and nonlinearity.  Catches any dimension-specific term."""
    rng = np.random.default_rng(6)
    s2 = (20, 18)
    med2 = _medium(s2, rng)
    p0 = _compact_p0(s2, rng, 4)
    sens = np.zeros(s2)
    sens[2:18:3, 1:17:5] = 1
    kg2 = ko.KGrid(20, 1e-4, 18, 1.2e-4)
    kg2.set_time(30, 1.5e-8)
    rec = ["p", "p_max", "p_rms", "p_final"]
    a = run(kg2, med2, {"p0": p0}, {"mask": sens, "record": rec}, pml_size=3, pml_alpha=2.0)
    Nz = 5
    ext = lambda v: np.repeat(np.asarray(v)[:, :, None], Nz, axis=2)   # noqa: E731
    med3 = {k: (ext(v) if np.ndim(v) == 2 else v) for k, v in med2.items()}
    sens3 = np.zeros(s2 + (Nz,))
    sens3[:, :, 2] = sens
    kg3 = ko.KGrid(20, 1e-4, 18, 1.2e-4, Nz, 0.7e-4)
    kg3.set_time(30, 1.5e-8)
    b = run(kg3, med3, {"p0": ext(p0)}, {"mask": sens3, "record": rec},
            pml_size=(3, 3, 1), pml_alpha=(2.0, 2.0, 0.0))
    for key in ("p", "p_max", "p_rms"):
        assert np.max(np.abs(a[key] - b[key])) < 1e-10 * np.max(np.abs(a[key]))
    for z in range(Nz):
        assert np.max(np.abs(a["p_final"] - b["p_final"][:, :, z])) < 1e-10 * np.max(np.abs(a["p_final"]))


# ---------------------------------------------------------------------------
# power-law absorption and dispersion (Treeby & Cox 2010)
# ---------------------------------------------------------------------------

def _mode_history(alpha0, y, m, mode=None, N=64, Nt=160):
    """p(t) of the single periodic mode cos(2 pi m x / L) (amplitude at x=0)."""
    dx, c = 1e-4, 1500.0
    dt = 0.3 * dx / c
    kg = ko.KGrid(N, dx, 4, dx)
    kg.set_time(Nt, dt)
    p0 = np.cos(2 * math.pi * m * np.arange(N) / N)[:, None] * np.ones((1, 4))
    med = {"sound_speed": c, "density": 1000.0, "alpha_coeff": alpha0, "alpha_power": y}
    if mode:
        med["alpha_mode"] = mode
    mask = np.zeros((N, 4))
    mask[0, 0] = 1
    a = run(kg, med, {"p0": p0}, {"mask": mask}, pml_size=2, pml_alpha=0.0)[0]
    # one mode of a linear two-field scheme: a[n+1] = c1 a[n] + c2 a[n-1]
    # exactly (after the forced first step); c2 = -r^2, c1 = 2 r cos(theta)
    h = a[1:]
    A = np.stack([h[1:-1], h[:-2]], 1)
    (c1, c2), *_ = np.linalg.lstsq(A, h[2:], rcond=None)
    r = math.sqrt(-c2)
    theta = math.acos(max(-1.0, min(1.0, c1 / (2 * r))))
    k = 2 * math.pi * m / (N * dx)
    return -math.log(r) / dt, theta / dt, c, k


def _theory(alpha0, y, c, k, dispersion=True, absorption=True):
    """Roots of s^2 + 2 a c^(y+1) k^y s + c^2 k^2 (1 - eta k^(y-1)) = 0 with
    eta = 2 a c^y tan(pi y / 2), a in Np/((rad/s)^y m)."""
    a = ko.db2neper(alpha0, y)
    g = a * c ** (y + 1) * k ** y if absorption else 0.0
    eta = 2 * a * c ** y * math.tan(math.pi * y / 2) if dispersion else 0.0
    return g, math.sqrt(c * c * k * k * (1 - eta * k ** (y - 1)) - g * g)


@pytest.mark.parametrize("y", [1.5, 1.3])
def test_absorption_decay_rate_and_dispersion(y):
    for m in (2, 4):
        g, w, c, k = _mode_history(1.0, y, m)
        g_th, w_th = _theory(1.0, y, c, k)
        assert g == pytest.approx(g_th, rel=2e-3)
        # phase-speed change from dispersion, relative to the lossless c k
        assert (w - c * k) == pytest.approx(w_th - c * k, rel=0.08)
    g2, *_ = _mode_history(1.0, y, 2)
    g4, *_ = _mode_history(1.0, y, 4)
    assert g4 / g2 == pytest.approx(2 ** y, rel=2e-3)


def test_absorption_modes():
    g, w, c, k = _mode_history(1.0, 1.5, 4, "no_absorption")
    _, w_th = _theory(1.0, 1.5, c, k, absorption=False)
    assert abs(g) < 1e-6
    assert (w - c * k) == pytest.approx(w_th - c * k, rel=5e-3)
    g, w, c, k = _mode_history(1.0, 1.5, 4, "no_dispersion")
    g_th, _ = _theory(1.0, 1.5, c, k, dispersion=False)
    assert g == pytest.approx(g_th, rel=2e-3)
    _, w_disp = _theory(1.0, 1.5, c, k)
    assert abs(w - c * k) < 0.1 * abs(w_disp - c * k)
    # Stokes: alpha_power is taken as 2 whatever is given
    g, w, c, k = _mode_history(1.0, 1.1, 4, "stokes")
    g_th, _ = _theory(1.0, 2.0, c, k, dispersion=False)
    assert g == pytest.approx(g_th, rel=2e-3)
    # no absorption at all: the exact lossless phase c k dt per step
    g, w, c, k = _mode_history(0.0, 1.5, 4)
    assert abs(g) < 1e-6 and w == pytest.approx(c * k, rel=1e-12)


def test_alpha_sign_flips_absorption():
    g, _, c, k = _mode_history(1.0, 1.5, 4)
    dx, N = 1e-4, 64
    kg = ko.KGrid(N, dx, 4, dx)
    kg.set_time(160, 0.3 * dx / 1500.0)
    p0 = np.cos(2 * math.pi * 4 * np.arange(N) / N)[:, None] * np.ones((1, 4))
    mask = np.zeros((N, 4))
    mask[0, 0] = 1
    med = {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 1.0, "alpha_power": 1.5,
           "alpha_sign": [-1, 1]}
    a = run(kg, med, {"p0": p0}, {"mask": mask}, pml_size=2, pml_alpha=0.0)[0]
    # with tau negated the mode grows at the same rate
    env = np.max(np.abs(a[-60:])) / np.max(np.abs(a[1:61]))
    dt = 0.3 * dx / 1500.0
    assert env == pytest.approx(math.exp(g * 99 * dt), rel=0.05)


# ---------------------------------------------------------------------------
# heterogeneous medium and nonlinearity
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("c2,r2", [(1875.0, 1600.0), (1500.0, 3000.0), (1200.0, 1000.0)])
def test_reflection_and_transmission_coefficients(c2, r2):
    """A pulse at normal incidence on a planar interface: reflected and
    transmitted pressure amplitudes (Z2-Z1)/(Z2+Z1) and 2 Z2/(Z1+Z2)."""
    N, Ny, dx, c1, r1, w = 512, 2, 1e-4, 1500.0, 1000.0, 6.0
    dt = 0.3 * dx / max(c1, c2)
    i = np.arange(N)
    xi, src = N // 2 + 40, N // 2
    c = np.where(i < xi, c1, c2)[:, None] * np.ones((1, Ny))
    rho = np.where(i < xi, r1, r2)[:, None] * np.ones((1, Ny))
    p0 = np.exp(-((i - src) / w) ** 2)[:, None] * np.ones((1, Ny))
    mask = np.zeros((N, Ny))
    mask[src - 60, 0] = mask[xi + 60, 0] = 1
    Nt = int((200 * dx / min(c1, c2)) / dt)
    kg = ko.KGrid(N, dx, Ny, dx)
    kg.set_time(Nt, dt)
    left, right = run(kg, {"sound_speed": c, "density": rho}, {"p0": p0}, {"mask": mask},
                      pml_size=(20, 0), pml_alpha=(2.0, 0.0))
    t = np.arange(Nt) * dt
    split = t < 100 * dx / c1
    direct, refl = left[split], left[~split]
    pk = lambda v: v[np.argmax(np.abs(v))]   # noqa: E731
    Z1, Z2 = r1 * c1, r2 * c2
    assert pk(direct) == pytest.approx(0.5, rel=1e-4)
    assert t[np.argmax(direct)] == pytest.approx(60 * dx / c1, abs=dt)
    assert pk(refl) / pk(direct) == pytest.approx((Z2 - Z1) / (Z2 + Z1), rel=0.025)
    assert pk(right) / pk(direct) == pytest.approx(2 * Z2 / (Z1 + Z2), rel=0.005)


def _second_harmonic(BonA, dist=100):
    N, Ny, dx, c, rho, P, ppw = 384, 2, 1e-4, 1500.0, 1000.0, 1e6, 24
    dt = 0.3 * dx / c
    f = c / (ppw * dx)
    spp = 1 / (f * dt)
    Nt = int((dist + 5) * dx / c / dt + 10 * spp)
    kg = ko.KGrid(N, dx, Ny, dx)
    kg.set_time(Nt, dt)
    t = np.arange(Nt) * dt
    ramp = 0.5 - 0.5 * np.cos(math.pi * np.clip(t * f / 2, 0, 1))
    pm = np.zeros((N, Ny))
    pm[40, :] = 1
    mask = np.zeros((N, Ny))
    mask[40 + dist, 0] = 1
    med = {"sound_speed": c, "density": rho}
    if BonA is not None:
        med["BonA"] = BonA
    d = run(kg, med, {"p_mask": pm, "p": (P * ramp * np.sin(2 * math.pi * f * t))[None, :],
                      "p_mode": "dirichlet"}, {"mask": mask},
            pml_size=(20, 0), pml_alpha=(2.0, 0.0))[0]
    n = int(round(6 * spp))
    ts, seg = t[-n:], d[-n:]
    w1 = 2 * math.pi * f
    A = np.stack([fn(h * w1 * ts) for h in (1, 2, 3) for fn in (np.cos, np.sin)]
                 + [np.ones_like(ts)], 1)
    co = np.linalg.lstsq(A, seg, rcond=None)[0]
    p1, p2 = math.hypot(co[0], co[1]), math.hypot(co[2], co[3])
    # Fubini, small distance: p2 = beta w p1^2 x / (2 rho c^3), beta = 1 + B/2A
    beta = 1 + (BonA or 0.0) / 2
    return p1, p2, beta * w1 * p1 ** 2 * (dist * dx) / (2 * rho * c ** 3)


def test_second_harmonic_growth():
    p1, p2, th = _second_harmonic(6.0)
    assert p2 == pytest.approx(th, rel=0.02)
    q1, q2, qth = _second_harmonic(0.0)      # convective nonlinearity alone (beta = 1)
    assert q2 == pytest.approx(qth, rel=0.02)
    l1, l2, _ = _second_harmonic(None)       # linear: no harmonic
    assert l2 < 1e-6 * l1


# ---------------------------------------------------------------------------
# sources, time reversal and sensor bookkeeping
# ---------------------------------------------------------------------------

def test_time_reversal_equals_shifted_dirichlet_source():
    """Time reversal enforces the flipped boundary data on p (and on the split
    densities) at the start of each step and once more after the last one;
    for a linear lossless medium that is a dirichlet p source one step earlier
    with the data column order reversed and the first enforced column zero."""
    shape = (24, 20)
    kg = ko.KGrid(24, 1e-4, 20, 1e-4)
    Nt = 30
    kg.set_time(Nt, 2e-8)
    rng = np.random.default_rng(8)
    sens = np.zeros(shape)
    sens[3, 2:18] = 1
    sens[4:20, 17] = 1
    ns = int(sens.sum())
    data = rng.standard_normal((ns, Nt))
    data[:, -1] = 0      # first column used after the flip
    med = {"sound_speed": 1500 + 60 * rng.random(shape), "density": 1000.0}
    tr = run(kg, med, {}, {"mask": sens, "time_reversal_boundary_data": data}, pml_size=3)
    kd = ko.KGrid(24, 1e-4, 20, 1e-4)
    kd.set_time(Nt - 1, 2e-8)
    sd = run(kd, med, {"p_mask": sens, "p": data[:, ::-1][:, 1:], "p_mode": "dirichlet"},
             {"mask": sens, "record": ["p_final"]}, pml_size=3)
    assert np.max(np.abs(tr - sd["p_final"])) < 1e-12 * np.max(np.abs(tr))


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_additive_source_is_filtered_no_correction_source(dim):
    """An 'additive' p source equals an 'additive-no-correction' source on
    every cell whose signal is the source mask filtered by
    cos(c_ref k dt / 2) in k-space times the time signal; both are scaled by
    2 dt / (dim c dx)."""
    shape = {1: (40,), 2: (16, 14), 3: (10, 12, 9)}[dim]
    sp = (1e-4, 1.1e-4, 0.9e-4)[:dim]
    args = [v for pair in zip(shape, sp) for v in pair]
    rng = np.random.default_rng(9)
    c = 1500.0
    Nt, dt = 20, 2e-8
    mask = np.zeros(shape)
    mask[tuple(rng.integers(2, n - 2, 5) for n in shape)] = 1
    s = rng.standard_normal(Nt)
    med = {"sound_speed": c, "density": 1000.0, "alpha_coeff": 0.5, "alpha_power": 1.5}
    kg = ko.KGrid(*args)
    kg.set_time(Nt, dt)
    a = run(kg, med, {"p_mask": mask, "p": s[None, :], "p_mode": "additive"},
            {"mask": np.ones(shape), "record": ["p_final", "p_max"]}, pml_size=2)
    K = np.real(np.fft.ifftn(np.cos(c * kmag(shape, sp) * dt / 2) * np.fft.fftn(mask)))
    sig = K.ravel(order="F")[:, None] * s[None, :]
    kg.set_time(Nt, dt)
    b = run(kg, med, {"p_mask": np.ones(shape), "p": sig, "p_mode": "additive-no-correction"},
            {"mask": np.ones(shape), "record": ["p_final", "p_max"]}, pml_size=2)
    ref = np.max(np.abs(a["p_final"]))
    assert ref > 0
    assert np.max(np.abs(a["p_final"] - b["p_final"])) < 1e-11 * ref
    assert np.max(np.abs(a["p_max"] - b["p_max"])) < 1e-11 * np.max(np.abs(a["p_max"]))


def test_source_linearity_and_scaling():
    """Linear medium: doubling the signal doubles the field; the additive
    scaling 2 dt/(dim c dx) makes a no-correction point source independent
    of how the same total signal is split over two coincident series."""
    shape = (18, 16)
    kg = ko.KGrid(18, 1e-4, 16, 1e-4)
    kg.set_time(15, 2e-8)
    rng = np.random.default_rng(10)
    m = np.zeros(shape)
    m[5, 6] = m[12, 9] = 1
    s = rng.standard_normal((2, 15))
    med = {"sound_speed": 1500.0, "density": 1000.0}
    sens = {"mask": np.ones(shape), "record": ["p_final"]}
    one = run(kg, med, {"p_mask": m, "p": s, "p_mode": "additive-no-correction"}, sens, pml_size=3)
    two = run(kg, med, {"p_mask": m, "p": 2 * s, "p_mode": "additive-no-correction"}, sens, pml_size=3)
    assert np.allclose(two["p_final"], 2 * one["p_final"], rtol=1e-12, atol=1e-20)
    # labelled mask: both points read series 1 and 2 by label
    lab = m.copy()
    lab[5, 6], lab[12, 9] = 2, 1
    sw = run(kg, med, {"p_mask": lab, "p": s[::-1], "p_mode": "additive-no-correction"}, sens, pml_size=3)
    assert np.allclose(sw["p_final"], one["p_final"], rtol=1e-12, atol=1e-20)


def test_sensor_statistics_match_time_series():
    rng = np.random.default_rng(11)
    shape = (20, 16, 12)
    med = _medium(shape, rng, absorbing=True, nonlinear=False)
    p0 = _compact_p0(shape, rng, 3)
    sens = np.zeros(shape)
    sens[::3, ::4, ::5] = 1
    kg = ko.KGrid(20, 1e-4, 16, 1e-4, 12, 1e-4)
    kg.set_time(25, 1.5e-8)
    rec = ["p", "p_max", "p_min", "p_rms", "p_max_all", "p_min_all", "p_final",
           "u", "u_max", "u_min", "u_rms", "u_max_all", "u_min_all", "u_final"]
    sd, st = run(kg, med, {"p0": p0}, {"mask": sens, "record": rec, "record_start_index": 6},
                 pml_size=3, return_state=True)
    p = sd["p"]
    assert p.shape == (int(sens.sum()), 20)
    assert np.array_equal(sd["p_max"], p.max(1)) and np.array_equal(sd["p_min"], p.min(1))
    assert np.allclose(sd["p_rms"], np.sqrt(np.mean(p ** 2, 1)), rtol=1e-12)
    assert np.array_equal(sd["p_final"], st.p)
    assert np.array_equal(sd["p_final"].ravel(order="F")[ko.find_F(sens)], p[:, -1])
    for u, v in zip(("ux", "uy", "uz"), st.u):
        ts = sd[u]
        assert np.array_equal(sd[u + "_max"], ts.max(1)) and np.array_equal(sd[u + "_min"], ts.min(1))
        assert np.allclose(sd[u + "_rms"], np.sqrt(np.mean(ts ** 2, 1)), rtol=1e-12)
        assert np.array_equal(sd[u + "_final"], v)
        assert np.array_equal(sd[u + "_max_all"].ravel(order="F")[ko.find_F(sens)], ts.max(1))
    assert np.array_equal(sd["p_max_all"].ravel(order="F")[ko.find_F(sens)], p.max(1))
    assert np.array_equal(sd["p_min_all"].ravel(order="F")[ko.find_F(sens)], p.min(1))


def test_record_start_index_drops_leading_steps():
    shape = (16, 16)
    kg = ko.KGrid(16, 1e-4, 16, 1e-4)
    kg.set_time(20, 2e-8)
    rng = np.random.default_rng(12)
    p0 = _compact_p0(shape, rng, 3)
    sens = np.zeros(shape)
    sens[4, :] = 1
    full = run(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0}, {"mask": sens}, pml_size=3)
    late = run(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0},
               {"mask": sens, "record_start_index": 8}, pml_size=3)
    assert np.array_equal(late, full[:, 7:])


def test_single_precision_tracks_double():
    """'DataCast','single' stays within FP32 rounding growth of the FP64 run."""
    rng = np.random.default_rng(13)
    shape = (32, 28)
    med = _medium(shape, rng)
    p0 = _compact_p0(shape, rng, 5)
    sens = np.zeros(shape)
    sens[6, 4:24] = 1
    kg = ko.KGrid(32, 1e-4, 28, 1e-4)
    kg.set_time(60, 1.5e-8)
    d = run(kg, med, {"p0": p0}, {"mask": sens}, pml_size=4)
    s = run(kg, med, {"p0": p0}, {"mask": sens}, pml_size=4, data_cast="single")
    assert s.dtype == np.float32
    assert 1e-9 < ko.rel_linf(s, d) < 1e-5


def test_pml_absorbs_outgoing_energy():
    """Waves leaving the domain are absorbed: after the pulse has crossed the
    PML the remaining field is a small fraction of the initial one."""
    shape = (64, 64)
    kg = ko.KGrid(64, 1e-4, 64, 1e-4)
    kg.set_time(400, 2e-8)
    x = np.arange(64) - 32
    p0 = np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / 9.0)
    sd = run(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0},
             {"mask": np.ones(shape), "record": ["p_final"]}, pml_size=12, pml_alpha=2.0)
    assert np.max(np.abs(sd["p_final"])) < 0.01 * np.max(p0)
    sd0 = run(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": p0},
              {"mask": np.ones(shape), "record": ["p_final"]}, pml_size=12, pml_alpha=0.0)
    assert np.max(np.abs(sd0["p_final"])) > 0.03 * np.max(p0)


@pytest.mark.parametrize("dim,axis", [(2, 0), (2, 1), (3, 2)])
def test_standing_wave_velocity_and_intensity(dim, axis):
    """Periodic lossless standing wave p0 = cos(k x_a): the k-space scheme
    is exact, so p = cos(k x) cos(w t_n) with t_n = (n-1) dt, the velocity
    on the non-staggered grid is sin(k x) sin(w (t_n - dt/2)) / (rho c)
    (at the grid points, half a step behind p), and after the half-step
    Fourier shift in time of kspaceFirstOrder_saveIntensity.m the intensity
    is I = p u = cos(k x) sin(k x) cos(w t) sin(w t) / (rho c); two whole
    periods are recorded, so I_avg = 0.  The other components vanish."""
    N, dx, c, rho, m = 32, 1e-4, 1500.0, 1000.0, 3
    k = 2 * math.pi * m / (N * dx)
    w = c * k
    Nt = 40
    dt = 2 * math.pi / w / 20
    shape = [4] * dim
    shape[axis] = N
    args = []
    for n in shape:
        args += [n, dx]
    kg = ko.KGrid(*args)
    kg.set_time(Nt, dt)
    x = np.arange(N) * dx
    sh = [1] * dim
    sh[axis] = N
    p0 = np.broadcast_to(np.cos(k * x).reshape(sh), shape).copy()
    sd = run(kg, {"sound_speed": c, "density": rho}, {"p0": p0},
             {"mask": np.ones(shape), "record": ["p", "u_non_staggered", "I", "I_avg"]},
             pml_size=2, pml_alpha=0.0)
    xa = np.broadcast_to(x.reshape(sh), shape).ravel(order="F")[:, None]
    t = (np.arange(Nt) * dt)[None, :]
    comp = "xyz"[axis]
    assert np.max(np.abs(sd["p"] - np.cos(k * xa) * np.cos(w * t))) < 1e-12
    u = np.sin(k * xa) * np.sin(w * (t - dt / 2)) / (rho * c)
    assert np.max(np.abs(sd[f"u{comp}_non_staggered"] - u)) < 1e-12 * np.max(np.abs(u))
    intensity = np.cos(k * xa) * np.sin(k * xa) * np.cos(w * t) * np.sin(w * t) / (rho * c)
    assert np.max(np.abs(sd[f"I{comp}"] - intensity)) < 1e-12 * np.max(np.abs(intensity))
    assert np.max(np.abs(sd[f"I{comp}_avg"])) < 1e-12 * np.max(np.abs(intensity))
    for other in "xyz"[:dim]:
        if other != comp:
            assert np.max(np.abs(sd[f"u{other}_non_staggered"])) < 1e-12 * np.max(np.abs(u))
            assert np.max(np.abs(sd[f"I{other}"])) < 1e-12 * np.max(np.abs(intensity))


@pytest.mark.parametrize("dim", [2, 3])
def test_pml_outside_equals_expanded_problem(dim):
    """'PMLInside', false is the simulation of the problem grown by the PML
    on every side (medium maps by their edge values, p0 smoothed on the
    user's grid and the masks by zeros), with the _final and _all grids cut
    back to the user's grid and the sensor series unchanged; the expanded
    problem is built here with numpy.pad."""
    shape = (24, 40) if dim == 2 else (12, 18, 10)
    pml = 4 if dim == 2 else 3
    dx, Nt, dt = 1e-4, 20, 1.5e-8
    rng = np.random.default_rng(dim)
    c = 1500 + 100 * rng.uniform(size=shape)
    rho = 1000 + 100 * rng.uniform(size=shape)
    p0 = np.zeros(shape)
    p0[tuple(slice(n // 2 - 2, n // 2 + 2) for n in shape)] = 1.0
    mask = rng.uniform(size=shape) < 0.2
    rec = ["p", "p_final", "p_max_all", "u_final"]
    args = []
    for n in shape:
        args += [n, dx]
    kg = ko.KGrid(*args)
    kg.set_time(Nt, dt)
    med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.5, "alpha_power": 1.5}
    got = run(kg, med, {"p0": p0}, {"mask": mask, "record": rec}, pml_size=pml,
              pml_inside=False, smooth=(True, False, False))
    pad = [(pml, pml)] * dim
    args = []
    for n in shape:
        args += [n + 2 * pml, dx]
    kge = ko.KGrid(*args)
    kge.set_time(Nt, dt)
    mede = dict(med, sound_speed=np.pad(c, pad, mode="edge"), density=np.pad(rho, pad, mode="edge"))
    p0e = np.pad(ko.smooth_mat(p0, True), pad)
    want = run(kge, mede, {"p0": p0e}, {"mask": np.pad(mask, pad), "record": rec},
               pml_size=pml, smooth=(False, False, False))
    inside = tuple(slice(pml, pml + n) for n in shape)
    assert np.array_equal(got["p"], want["p"])
    for k in ("p_final", "p_max_all", "ux_final"):
        assert got[k].shape == shape
        assert np.array_equal(got[k], want[k][inside])


@pytest.mark.parametrize("dim", [2, 3])
def test_split_field_of_irrotational_velocity(dim):
    """In a homogeneous periodic medium the velocity is the gradient of a
    scalar at every step (u_k = -dt/rho0 i k kappa exp(i k d/2) P summed
    over steps), so its shear part vanishes and its compressional part is
    the velocity on the non-staggered grid."""
    shape = (16, 12) if dim == 2 else (12, 10, 8)
    args = []
    for n in shape:
        args += [n, 1e-4]
    kg = ko.KGrid(*args)
    kg.set_time(12, 2e-8)
    rng = np.random.default_rng(5)
    sd = run(kg, {"sound_speed": 1500.0, "density": 1000.0}, {"p0": rng.standard_normal(shape)},
             {"mask": np.ones(shape), "record": ["u_non_staggered", "u_split_field"]},
             pml_size=2, pml_alpha=0.0)
    for u in ["ux", "uy", "uz"][:dim]:
        scale = np.max(np.abs(sd[u + "_non_staggered"]))
        assert np.max(np.abs(sd[u + "_split_s"])) < 1e-12 * scale
        assert np.max(np.abs(sd[u + "_split_p"] - sd[u + "_non_staggered"])) < 1e-12 * scale


@pytest.mark.parametrize("dim", [2, 3])
def test_cartesian_linear_at_grid_nodes_and_midpoints(dim):
    """'CartInterp' 'linear' on points that are grid nodes returns what
    'nearest' returns for them (in the points' order), and at the midpoint
    of a grid edge the mean of the two nodes' series: both hold for every
    triangulation of the grid."""
    shape = (14, 12) if dim == 2 else (10, 12, 8)
    args = []
    for n in shape:
        args += [n, 1e-4]
    kg = ko.KGrid(*args)
    kg.set_time(10, 2e-8)
    rng = np.random.default_rng(11)
    p0 = np.zeros(shape)
    p0[tuple(slice(n // 2 - 2, n // 2 + 2) for n in shape)] = rng.uniform(0.5, 1.5, (4,) * dim)
    nodes = np.array([[2, -3, 0, 1][:4], [-1, 2, 0, -3][:4], [1, 0, -2, 2][:4]][:dim]) * 1e-4
    rec = ["p", "p_max", "u", "u_non_staggered"]
    med = {"sound_speed": 1500.0, "density": 1000.0}
    lin = run(kg, med, {"p0": p0}, {"mask": nodes, "record": rec}, pml_size=2, cart_interp="linear")
    near = run(kg, med, {"p0": p0}, {"mask": nodes, "record": rec}, pml_size=2, cart_interp="nearest")
    for k in near:
        assert np.allclose(lin[k], near[k], rtol=0, atol=1e-12 * np.max(np.abs(near[k]))), k
    # midpoints of the edges between node 0 and its neighbour along each axis
    mids = []
    for a in range(dim):
        m = nodes[:, 0].copy()
        m[a] += 0.5e-4
        mids.append(m)
    ends = []
    for a in range(dim):
        e = nodes[:, 0].copy()
        e[a] += 1e-4
        ends.append(e)
    pts = np.stack(mids + ends + [nodes[:, 0]], axis=1)
    lin = run(kg, med, {"p0": p0}, {"mask": pts, "record": ["p"]}, pml_size=2, cart_interp="linear")
    p = lin["p"]
    for a in range(dim):
        want = 0.5 * (p[-1] + p[dim + a])
        assert np.allclose(p[a], want, rtol=0, atol=1e-12 * np.max(np.abs(p)))


@pytest.mark.parametrize("kind", ["p0", "p_no_correction", "p_dirichlet", "ux_additive",
                                  "time_reversal"])
def test_1d_equals_y_invariant_2d(kind):
    """kspaceFirstOrder1D against the 2-D loop (pinned above) on a problem
    that does not vary along y, with no PML along y: for linear problems,
    the 2-D densities add up to the 1-D one (the y part holds its initial
    share and changes only where a p source feeds it), the p sources scaled
    by 1/dim go into each component, and a sensor row of the 2-D run equals
    the 1-D sensors.  Heterogeneous c and rho, power-law absorption, PML
    along x.  (The k-space filtered 'additive' p source also reaches the
    PML cells, where the split PML damps only the x part of the 2-D density,
    so it is pinned on its own by the filtered-source test above.)"""
    Nx, Ny, dx, Nt, dt = 64, 6, 1e-4, 40, 2e-8
    rng = np.random.default_rng(7)
    c1 = 1500 + 60 * np.sin(np.arange(Nx) / 6.0)
    r1 = 1000 + 50 * np.cos(np.arange(Nx) / 4.0)
    med1 = {"sound_speed": c1, "density": r1, "alpha_coeff": 0.6, "alpha_power": 1.4}
    med2 = {"sound_speed": np.repeat(c1[:, None], Ny, 1), "density": np.repeat(r1[:, None], Ny, 1),
            "alpha_coeff": 0.6, "alpha_power": 1.4}
    sens_x = np.array([20, 31, 40, 47])
    m1 = np.zeros(Nx)
    m1[sens_x] = 1
    m2 = np.zeros((Nx, Ny))
    m2[sens_x, 0] = 1
    sig = np.sin(2 * np.pi * 2e6 * np.arange(Nt) * dt) * np.hanning(Nt)
    src1, src2 = {}, {}
    if kind == "p0":
        p0 = np.zeros(Nx)
        p0[28:36] = rng.uniform(0.5, 1.5, 8)
        src1["p0"], src2["p0"] = p0, np.repeat(p0[:, None], Ny, 1)
    elif kind.startswith("p_") or kind == "ux_additive":
        sm1 = np.zeros(Nx)
        sm1[30:33] = 1
        key = "p" if kind.startswith("p_") else "ux"
        mode = {"p_dirichlet": "dirichlet", "p_no_correction": "additive-no-correction"}.get(
            kind, "additive")
        pre = "p" if key == "p" else "u"
        src1 = {key: sig, pre + "_mask": sm1, pre + "_mode": mode}
        src2 = {key: sig, pre + "_mask": np.repeat(sm1[:, None], Ny, 1), pre + "_mode": mode}
    kw = dict(pml_alpha=2.0)
    if kind == "time_reversal":
        data = rng.standard_normal((4, Nt))
        kg1 = ko.KGrid(Nx, dx)
        kg1.set_time(Nt, dt)
        kg2 = ko.KGrid(Nx, dx, Ny, dx)
        kg2.set_time(Nt, dt)
        got1 = run(kg1, med1, {}, {"mask": m1, "time_reversal_boundary_data": data}, pml_size=8, **kw)
        # the boundary data on every row (column-major order: row by row)
        m2all = np.repeat(m1[:, None], Ny, 1)
        got2 = run(kg2, med2, {}, {"mask": m2all, "time_reversal_boundary_data": np.tile(data, (Ny, 1))},
                   pml_size=[8, 0], **kw)
        assert np.allclose(got1, got2[:, 0], rtol=0, atol=1e-10 * np.max(np.abs(got2)))
        return
    kg1 = ko.KGrid(Nx, dx)
    kg1.set_time(Nt, dt)
    kg2 = ko.KGrid(Nx, dx, Ny, dx)
    kg2.set_time(Nt, dt)
    rec = ["p", "u", "p_final"]
    got1 = run(kg1, med1, src1, {"mask": m1, "record": rec}, pml_size=8, **kw)
    got2 = run(kg2, med2, src2, {"mask": m2, "record": rec}, pml_size=[8, 0], **kw)
    for k1, k2 in (("p", "p"), ("ux", "ux"), ("p_final", "p_final")):
        a, b = np.asarray(got1[k1]), np.asarray(got2[k2])
        if k1 == "p_final":
            b = b[:, 0]
        assert np.allclose(a, b, rtol=0, atol=1e-10 * np.max(np.abs(b))), k1
    assert np.max(np.abs(got2["uy"])) < 1e-12 * np.max(np.abs(got2["ux"]))


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_cuboid_corners_equal_full_grid_record(dim):
    """A cuboid-corners sensor returns, per cuboid, the full-grid record at
    the cuboid's points reshaped to the cuboid (time last); overlapping
    cuboids both see their shared points; whole-grid outputs go to the
    first cuboid; invalid corners raise the reference's errors."""
    shape = {1: (40,), 2: (20, 18), 3: (12, 10, 9)}[dim]
    args = []
    for n in shape:
        args += [n, 1e-4]
    kg = ko.KGrid(*args)
    kg.set_time(12, 2e-8)
    rng = np.random.default_rng(13)
    p0 = np.zeros(shape)
    p0[tuple(slice(n // 2 - 2, n // 2 + 2) for n in shape)] = rng.uniform(0.5, 1.5, (4,) * dim)
    med = {"sound_speed": 1500.0, "density": 1000.0}
    lo1, hi1 = [3, 4, 2][:dim], [9, 8, 6][:dim]
    lo2, hi2 = [6, 2, 3][:dim], [12, 5, 7][:dim]
    corners = np.array([lo1 + hi1, lo2 + hi2], dtype=float).T
    rec = ["p", "p_max", "u", "p_max_all"]
    cub = run(kg, med, {"p0": p0}, {"mask": corners, "record": rec}, pml_size=2)
    full = run(kg, med, {"p0": p0}, {"mask": np.ones(shape), "record": rec}, pml_size=2)
    assert isinstance(cub, list) and len(cub) == 2
    for q, (lo, hi) in enumerate(((lo1, hi1), (lo2, hi2))):
        sl = tuple(slice(a - 1, b) for a, b in zip(lo, hi))
        for k in ("p", "ux", "p_max"):
            g = np.asarray(full[k]).reshape(shape + np.asarray(full[k]).shape[1:], order="F")
            assert np.array_equal(cub[q][k], g[sl]), (q, k)
    assert np.array_equal(cub[0]["p_max_all"], full["p_max_all"]) and "p_max_all" not in cub[1]
    with pytest.raises(ValueError, match="integers"):
        run(kg, med, {"p0": p0}, {"mask": corners + 0.5}, pml_size=2)
    with pytest.raises(ValueError, match="x2 => x1"):
        run(kg, med, {"p0": p0}, {"mask": corners[list(range(dim, 2 * dim)) + list(range(dim))]},
            pml_size=2)
    bad = corners.copy()
    bad[dim, 0] = shape[0] + 1
    with pytest.raises(ValueError, match="within the grid"):
        run(kg, med, {"p0": p0}, {"mask": bad}, pml_size=2)


@pytest.mark.parametrize("shape", [(24, 20), (25, 21)])
def test_directivity_on_a_plane_wave(shape):
    """2-D sensor directivity on a periodic standing wave cos(kx0 x): the
    field stays one mode (+-kx0, 0), so every recorded p is the raw p times
    the mean of the filter at the two bins.  Even sizes: 'gradient' gives
    |cos(theta)|, 'pressure' sinc(|sin(theta)| kx0 size/2), and size 0 the
    raw p.  Odd sizes: the reference moves the centred filter to FFT order
    with fftshift, which reads the filter one bin off, at kx = (m+1) dk and
    (1-m) dk and ky = dky; the factor is the mean of |kx|/|k| there."""
    Nx, Ny = shape
    dx = dy = 1e-4
    kg = ko.KGrid(Nx, dx, Ny, dy)
    kg.set_time(15, 2e-8)
    m = 3
    x = np.arange(Nx) * dx
    p0 = np.repeat(np.cos(2 * np.pi * m * x / (Nx * dx))[:, None], Ny, 1)
    mask = np.zeros(shape)
    pts = [(2, 3), (7, 5), (11, 9), (16, 14)]
    for i, j in pts:
        mask[i, j] = 1
    thetas = [0.0, np.pi / 6, np.pi / 2, 2.0]
    ang = np.zeros(shape)
    for (i, j), th in zip(pts, thetas):
        ang[i, j] = th
    med = {"sound_speed": 1500.0, "density": 1000.0}
    raw = run(kg, med, {"p0": p0}, {"mask": mask, "record": ["p"]}, pml_size=2, pml_alpha=0.0)["p"]
    order = np.argsort([i + Nx * j for i, j in pts])   # sensor rows in column-major order
    th_rows = np.array(thetas)[order]
    kx0 = 2 * np.pi * m / (Nx * dx)
    dkx, dky = 2 * np.pi / (Nx * dx), 2 * np.pi / (Ny * dy)

    def directional(pattern, size):
        sens = {"mask": mask, "record": ["p"], "directivity_angle": ang,
                "directivity_pattern": pattern, "directivity_size": size}
        return run(kg, med, {"p0": p0}, sens, pml_size=2, pml_alpha=0.0)["p"]

    scale = np.max(np.abs(raw))
    got = directional("gradient", 1e-3)
    if Nx % 2 == 0 and Ny % 2 == 0:
        want = np.abs(np.cos(th_rows))
    else:
        def g(kx, ky, th):
            return abs(np.sin(th) * ky + np.cos(th) * kx) / np.hypot(kx, ky)
        want = np.array([(g((m + 1) * dkx, dky, th) + g((1 - m) * dkx, dky, th)) / 2 for th in th_rows])
    assert np.allclose(got, want[:, None] * raw, rtol=0, atol=1e-9 * scale)
    if Nx % 2 == 0 and Ny % 2 == 0:
        size = 3 * dx
        got = directional("pressure", size)
        want = np.array([ko.sinc(abs(np.sin(th)) * kx0 * size / 2) for th in th_rows])
        assert np.allclose(got, want[:, None] * raw, rtol=0, atol=1e-9 * scale)
        assert np.allclose(directional("pressure", 0.0), raw, rtol=0, atol=1e-12 * scale)


def test_alpha_filter_selects_absorbed_modes():
    """medium.alpha_filter (centred k-space order) multiplies both absorption
    operators: on a periodic grid, a filter that is 1 at the centred bins of
    +-k1 and 0 elsewhere absorbs mode k1 as without a filter and leaves mode
    k2 lossless (by linearity the run is the sum of the two); a filter of
    ones changes nothing."""
    N, dx, Nt, dt = 48, 1e-4, 40, 2e-8
    kg = ko.KGrid(N, dx)
    kg.set_time(Nt, dt)
    x = np.arange(N) * dx
    m1, m2 = 5, 9
    f1 = np.cos(2 * np.pi * m1 * x / (N * dx))
    f2 = np.cos(2 * np.pi * m2 * x / (N * dx))
    absorbing = {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 40.0, "alpha_power": 1.5}
    lossless = {"sound_speed": 1500.0, "density": 1000.0}
    sens = {"mask": np.ones(N), "record": ["p"]}
    kw = dict(pml_size=2, pml_alpha=0.0)
    af = np.zeros(N)
    af[N // 2 + m1] = af[N // 2 - m1] = 1.0     # centred index of +-k1
    got = run(kg, dict(absorbing, alpha_filter=af), {"p0": f1 + f2}, sens, **kw)["p"]
    want = run(kg, absorbing, {"p0": f1}, sens, **kw)["p"] + run(kg, lossless, {"p0": f2}, sens, **kw)["p"]
    assert np.max(np.abs(got - want)) < 1e-11 * np.max(np.abs(want))
    plain = run(kg, absorbing, {"p0": f1 + f2}, sens, **kw)["p"]
    assert np.max(np.abs(got - plain)) > 1e-3 * np.max(np.abs(plain))   # mode k2 was absorbed there
    ones = run(kg, dict(absorbing, alpha_filter=np.ones(N)), {"p0": f1 + f2}, sens, **kw)["p"]
    assert np.max(np.abs(ones - plain)) < 1e-12 * np.max(np.abs(plain))


@pytest.mark.parametrize("kind", ["p", "ux"])
def test_frequency_ref_source_matches_kspace_corrected_tone(kind):
    """source.p_frequency_ref / u_frequency_ref replace the k-space source
    filter cos(c k dt/2) by its value cos(w dt/2) at one frequency: for a
    tone at that frequency the radiated wave matches the k-space corrected
    ('additive') source to a small fraction of the 5 % the correction
    itself makes at this coarse time step."""
    N, dx = 256, 1e-4
    c = 1500.0
    dt = 0.5 * dx / c
    Nt = 400
    kg = ko.KGrid(N, dx)
    kg.set_time(Nt, dt)
    f0 = 3e6
    t = np.arange(Nt) * dt
    ramp = np.clip(t / (8 / f0), 0, 1)
    sig = ramp * np.sin(2 * np.pi * f0 * t)
    if kind == "ux":
        sig = sig * 1e-6
    m = np.zeros(N)
    m[40] = 1
    sens = {"mask": (np.arange(N) == 200).astype(float), "record": ["p"]}
    med = {"sound_speed": c, "density": 1000.0}
    pre = "p" if kind == "p" else "u"

    def go(mode, fref=None):
        src = {kind: sig, pre + "_mask": m, pre + "_mode": mode}
        if fref is not None:
            src[pre + "_frequency_ref"] = fref
        return run(kg, med, src, sens, pml_size=20)["p"][0, -100:]

    ref = go("additive")
    amp = np.max(np.abs(ref))
    plain = go("additive-no-correction")
    wcorr = go("additive", f0)
    assert np.max(np.abs(plain - ref)) > 0.03 * amp
    assert np.max(np.abs(wcorr - ref)) < 0.1 * np.max(np.abs(plain - ref))
    with pytest.raises(ValueError, match="maximum frequency"):
        go("additive", 10 * f0)
