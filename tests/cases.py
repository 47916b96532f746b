"""Problem builders shared by the emulator and GPU parity tests.

Each builder returns (kgrid_args, Nt, dt, medium, source, sensor, options):
plain dicts with the reference's field names, so that the same inputs go to
the oracle (oracle.kwave_oracle.kspace_first_order) and to the product
(kwhip.kspaceFirstOrder2D/3D).
"""

from __future__ import annotations

import math
import zlib

import numpy as np


def disc(shape, centre, radius):
    """Binary disc/ball (makeDisc/makeBall convention: distance <= radius,
    1-based centre)."""
    grids = np.meshgrid(*[np.arange(1, n + 1) for n in shape], indexing="ij")
    r2 = sum((g - c) ** 2 for g, c in zip(grids, centre))
    return (r2 <= radius ** 2).astype(np.float64)


def smooth_noise(shape, seed, sigma):
    """Uniform noise smoothed with a periodic Gaussian (FFT), scaled to
    max |g| = 1."""
    rng = np.random.default_rng(seed)
    g = rng.uniform(-1, 1, size=shape)
    ks = np.meshgrid(*[np.fft.fftfreq(n) for n in shape], indexing="ij")
    k2 = sum(k * k for k in ks)
    g = np.real(np.fft.ifftn(np.fft.fftn(g) * np.exp(-2 * (math.pi ** 2) * (sigma ** 2) * k2)))
    return g / np.max(np.abs(g))


def hetero_medium(shape, seed=0, sigma=2.0):
    c = np.clip(1500 + 100 * smooth_noise(shape, seed, sigma), 1400, 1600)
    rho = np.clip(1000 + 100 * smooth_noise(shape, seed + 1, sigma), 900, 1100)
    return c, rho


def tone(Nt, dt, f0, cycles, amp=1.0):
    t = np.arange(Nt) * dt
    n = int(math.floor(cycles / f0 / dt)) + 1
    s = np.zeros(Nt)
    m = min(n, Nt)
    w = np.hanning(n)[:m]
    s[:m] = amp * np.sin(2 * math.pi * f0 * t[:m]) * w
    return s[:m]


def case(name, dim=2, N=None):
    """Named parity cases.  Returns a dict with keys grid, Nt, dt, medium,
    source, sensor, opts.

    A name ending in "_off" is the named source case with every signal cut
    to SIGNAL_OFF_LEN time points, fewer than Nt: the source is applied at
    steps 1..SIGNAL_OFF_LEN and the remaining steps run with it switched off
    (kspaceFirstOrder3D.m:806,930: flags.source_p >= t_index)."""
    if name.endswith("_off"):
        cs = case(name[:-4], dim, N)
        src = dict(cs["source"])
        for nm in ("p", "ux", "uy", "uz"):
            if nm in src:
                s = np.atleast_2d(np.asarray(src[nm]))
                assert s.shape[1] > SIGNAL_OFF_LEN
                src[nm] = s[:, :SIGNAL_OFF_LEN] if np.ndim(cs["source"][nm]) == 2 else s[0, :SIGNAL_OFF_LEN]
        cs["source"] = src
        return cs
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    if N is None:
        N = {1: (96,), 2: (32, 30), 3: (16, 18, 20)}[dim]
        if "pml_outside" in name:
            # the grid grows by the PML on both sides: FFT-friendly after it
            N = {1: (40,), 2: (24, 40), 3: (12, 18, 10)}[dim]
        if name == "pml_auto":
            # sizes with factors above 5: 'PMLSize' 'auto' picks PMLs that
            # make the grown grid FFT-friendly
            N = {1: (53,), 2: (37, 29), 3: (13, 11, 17)}[dim]
    shape = tuple(N)
    dx = 1e-4
    c0 = 1500.0
    dt = 0.3 * dx / 1600.0
    Nt = 24
    med = {"sound_speed": c0, "density": 1000.0}
    src = {}
    pml = 4 if dim <= 2 else 3
    opts = {"PMLSize": pml, "PMLAlpha": 2.0, "Smooth": False}
    centre = [n // 2 + 1 for n in shape]
    sens = {"mask": np.ones(shape), "record": ["p", "p_final", "p_max", "p_min", "p_rms"]}
    if name == "p0_lossless":
        src["p0"] = 2.0 * disc(shape, centre, 3)
        opts["Smooth"] = True
    elif name == "p0_hetero_absorbing":
        c, rho = hetero_medium(shape, 3)
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5}
        src["p0"] = disc(shape, centre, 3)
        opts["Smooth"] = True
    elif name == "p0_nonlinear_absorbing":
        c, rho = hetero_medium(shape, 5)
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.5, "alpha_power": 1.3,
               "BonA": 6.0}
        src["p0"] = 1e6 * disc(shape, centre, 3)
    elif name == "p0_nonlinear_lossless":
        c, rho = hetero_medium(shape, 7)
        med = {"sound_speed": c, "density": rho, "BonA": 5.0 + rng.uniform(0, 2, size=shape)}
        src["p0"] = 2e6 * disc(shape, centre, 3)
    elif name == "stokes":
        med = {"sound_speed": c0, "density": 1000.0, "alpha_coeff": 0.5, "alpha_mode": "stokes"}
        src["p0"] = disc(shape, centre, 2)
    elif name == "no_dispersion":
        c, rho = hetero_medium(shape, 9)
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 2.0, "alpha_power": 1.0,
               "alpha_mode": "no_dispersion"}
        src["p0"] = disc(shape, centre, 2)
    elif name == "no_absorption":
        med = {"sound_speed": c0, "density": 1000.0, "alpha_coeff": 2.0, "alpha_power": 1.7,
               "alpha_mode": "no_absorption"}
        src["p0"] = disc(shape, centre, 2)
    elif name.startswith("psrc_"):
        mode = {"psrc_add": "additive", "psrc_dir": "dirichlet", "psrc_nocorr": "additive-no-correction",
                "psrc_many": "additive", "psrc_label": "additive", "psrc_many_dir": "dirichlet"}[name]
        c, rho = hetero_medium(shape, 11)
        med = {"sound_speed": c, "density": rho}
        m = np.zeros(shape)
        # a plane of points (a few points in 1-D)
        sl = tuple([slice(pml + 2, pml + (3 if dim > 1 else 6))] +
                   [slice(pml + 2, n - pml - 2) for n in shape[1:]])
        m[sl] = 1
        npts = int(m.sum())
        sig = tone(Nt, dt, 2e6, 3)
        if name in ("psrc_many", "psrc_many_dir"):
            sigs = np.stack([sig * (1 + 0.1 * i) for i in range(npts)])
            src["p"] = sigs
        elif name == "psrc_label":
            lab = m.copy()
            flat = np.flatnonzero(lab.ravel(order="F"))
            labels = (np.arange(len(flat)) % 3) + 1
            lf = lab.ravel(order="F")
            lf[flat] = labels
            m = lf.reshape(shape, order="F")
            src["p"] = np.stack([sig, 0.5 * sig, -sig])
        else:
            src["p"] = sig
        src["p_mask"] = m
        src["p_mode"] = mode
        sens = {"mask": _sparse_mask(shape, rng), "record": ["p", "p_max", "p_min", "p_rms", "u",
                                                               "u_max", "u_min", "u_rms"]}
    elif name.startswith("usrc_"):
        mode = {"usrc_add": "additive", "usrc_dir": "dirichlet",
                "usrc_nocorr": "additive-no-correction", "usrc_many": "additive"}[name]
        m = np.zeros(shape)
        sl = tuple([slice(pml + 3, pml + (4 if dim > 1 else 7))] +
                   [slice(pml + 2, n - pml - 2) for n in shape[1:]])
        m[sl] = 1
        npts = int(m.sum())
        sig = tone(Nt, dt, 2e6, 2, amp=1e-3)
        if name == "usrc_many":
            src["ux"] = np.stack([sig * (1 + 0.05 * i) for i in range(npts)])
            if dim >= 2:
                src["uy"] = np.stack([0.3 * sig * (1 - 0.05 * i) for i in range(npts)])
        else:
            src["ux"] = sig
            if dim >= 2:
                src["uy"] = 0.5 * sig
        if dim == 3:
            src["uz"] = -0.25 * (src["ux"])
        src["u_mask"] = m
        src["u_mode"] = mode
        sens = {"mask": _sparse_mask(shape, rng),
                "record": ["p", "u", "u_final", "u_max_all", "u_min_all", "p_max_all", "p_min_all"],
                "record_start_index": 5}
    elif name == "time_reversal":
        m = np.zeros(shape)
        m[pml + 1, ...] = 1
        m[..., pml + 1] = 1
        npts = int(m.sum())
        data = np.stack([tone(Nt, dt, 1.5e6, 2, amp=1 + 0.2 * np.sin(i)) for i in range(npts)])
        if data.shape[1] < Nt:
            data = np.pad(data, ((0, 0), (0, Nt - data.shape[1])))
        sens = {"mask": m, "time_reversal_boundary_data": data + 0.01 * rng.standard_normal((npts, Nt))}
    elif name == "intensity":
        # the intensities alone: p and u_non_staggered are recorded for them
        # and dropped again (kspaceFirstOrder_saveIntensity.m)
        c, rho = hetero_medium(shape, 11)
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5}
        src["p0"] = disc(shape, centre, 3)
        sens = {"mask": _sparse_mask(shape, rng), "record": ["I", "I_avg", "p_max"]}
    elif name == "u_nonstaggered":
        c, rho = hetero_medium(shape, 13)
        med = {"sound_speed": c, "density": rho}
        m = np.zeros(shape)
        m[tuple(slice(n // 2 - 2, n // 2 + 2) for n in shape)] = 1
        src["p"] = tone(Nt, dt, 2e6, 2)
        src["p_mask"] = m
        sens = {"mask": _sparse_mask(shape, rng), "record": ["p", "u", "u_non_staggered", "I"],
                "record_start_index": 4}
    elif name == "u_split":
        # compressional and shear parts of the velocity: a velocity source
        # whose components vary along the source plane drives a rotational
        # field, and the density gradients add vorticity
        c, rho = hetero_medium(shape, 19)
        med = {"sound_speed": c, "density": rho}
        m = np.zeros(shape)
        sl = tuple([slice(pml + 3, pml + 5)] + [slice(pml + 2, n - pml - 2) for n in shape[1:]])
        m[sl] = 1
        npts = int(m.sum())
        sig = tone(Nt, dt, 2e6, 2, amp=1e-3)
        src["ux"] = np.stack([sig * np.cos(0.7 * i) for i in range(npts)])
        src["uy"] = np.stack([sig * np.sin(0.4 * i + 0.3) for i in range(npts)])
        if dim == 3:
            src["uz"] = np.stack([0.5 * sig * np.cos(0.3 * i) for i in range(npts)])
        src["u_mask"] = m
        src["u_mode"] = "additive"
        sens = {"mask": _sparse_mask(shape, rng), "record": ["p", "u_split_field", "u_non_staggered"],
                "record_start_index": 3}
    elif name == "pml_outside":
        # 'PMLInside', false: heterogeneous maps grow by edge replication,
        # p0 is smoothed before the expansion, _all/_final cut back
        c, rho = hetero_medium(shape, 17)
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5}
        src["p0"] = disc(shape, centre, 3)
        opts["PMLInside"] = False
        opts["Smooth"] = True
        sens = {"mask": _sparse_mask(shape, rng),
                "record": ["p", "p_final", "p_max_all", "u_final", "p_rms", "u_non_staggered"]}
    elif name == "pml_outside_source":
        m = np.zeros(shape)
        m[tuple(slice(1, 3) for _ in shape)] = 1   # at the user's grid edge
        src["p"] = tone(Nt, dt, 2e6, 2)
        src["p_mask"] = m
        opts["PMLInside"] = False
        # Cartesian points near the source, given out of grid order (1-based
        # grid index i sits at x = (i - floor(N/2) - 1) dx)
        idx = np.array([[6, 3, 8, 4, 5, 7, 3, 9], [3, 6, 4, 8, 5, 2, 7, 3], [4, 2, 5, 3, 6, 4, 2, 5]])
        pts = [(idx[a] - shape[a] // 2 - 1) * dx for a in range(dim)]
        sens = {"mask": np.stack(pts), "record": ["p", "u", "p_min_all"]}
        opts["CartInterp"] = "nearest"
    elif name in ("cartesian_linear", "cartesian_linear_pml_outside"):
        # points between grid nodes, interpolated on the simplices of the
        # declared triangulation; the second grows the grid by the PML
        c, rho = hetero_medium(shape, 23)
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5}
        src["p0"] = disc(shape, centre, 3)
        npt = 11
        ang = np.arange(npt) * 2 * math.pi / npt + 0.3
        r = (min(shape) // 2 - pml - 1.37) * dx
        pts = [r * np.cos(ang), -r * np.sin(ang)][:dim]
        if dim == 3:
            pts.append(0.43 * dx * np.sin(2 * ang))
        pts = np.stack(pts)
        pts[:, 0] = [(n - 1 - n // 2) * dx for n in shape]   # the last grid node
        rec = ["p", "p_max", "p_min", "p_rms", "u", "u_rms", "u_non_staggered", "I_avg"]
        if name == "cartesian_linear_pml_outside":
            opts["PMLInside"] = False
            pts[:, 1] = [(n - 1 - n // 2 + 0.4) * dx for n in shape]   # between the user's edge and the PML
            rec = ["p", "p_max", "u", "p_final"]
        sens = {"mask": pts, "record": rec}
        opts["CartInterp"] = "linear"
    elif name in ("directivity_pressure", "directivity_gradient"):
        # 2-D sensor directivity: a line of sensors facing the source at a
        # few angles, a full-grid sensor with one angle per column
        c, rho = hetero_medium(shape, 37)
        med = {"sound_speed": c, "density": rho}
        src["p0"] = disc(shape, centre, 3)
        pattern = name.split("_")[1]
        if pattern == "pressure":
            m = _sparse_mask(shape, rng, 0.2)
            ang = np.zeros(shape)
            ang[...] = (np.arange(shape[0]) % 4)[:, None] * (np.pi / 5)
            sens = {"mask": m, "record": ["p", "p_max", "I_avg"], "directivity_angle": ang,
                    "directivity_size": 4 * dx}
        else:
            ang = np.repeat((np.arange(shape[1]) % 3)[None, :] * 0.7, shape[0], 0)
            sens = {"mask": np.ones(shape), "record": ["p", "p_rms"], "directivity_angle": ang,
                    "directivity_pattern": "gradient"}
    elif name == "source_frequency_ref":
        # the k-space source correction at one frequency (a scaled
        # 'additive-no-correction' source), for p and u sources
        c, rho = hetero_medium(shape, 43)
        med = {"sound_speed": c, "density": rho}
        m = np.zeros(shape)
        m[tuple(slice(n // 2 - 1, n // 2 + 1) for n in shape)] = 1
        sig = tone(Nt, dt, 2e6, 2)
        src = {"p": sig, "p_mask": m, "p_frequency_ref": 2e6,
               "ux": 1e-3 * sig, "u_mask": m, "u_frequency_ref": 1.5e6}
        sens = {"mask": _sparse_mask(shape, rng, 0.3), "record": ["p", "u_max"]}
    elif name == "pml_auto":
        c, rho = hetero_medium(shape, 47)
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.5, "alpha_power": 1.5}
        src["p0"] = disc(shape, centre, 3)
        opts["PMLInside"] = False
        opts["PMLSize"] = "auto"
        opts["PMLRange"] = [4, 14]
        sens = {"mask": _sparse_mask(shape, rng, 0.3), "record": ["p", "p_final", "p_max_all"]}
    elif name == "alpha_filter":
        # absorption compensation: the power-law operators times a centred
        # k-space window (not symmetric for odd sizes)
        c, rho = hetero_medium(shape, 41)
        g = np.meshgrid(*[np.linspace(-1, 1, n) for n in shape], indexing="ij")
        af = np.clip(1.2 - sum(v * v for v in g), 0, 1) * (1 + 0.1 * g[0])
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 3.0, "alpha_power": 1.5,
               "alpha_filter": af}
        src["p0"] = disc(shape, centre, 3)
    elif name == "freq_response":
        # the recorded p filtered by a Gaussian around 2 MHz (intensity and
        # p_max are taken from the unfiltered p)
        c, rho = hetero_medium(shape, 31)
        med = {"sound_speed": c, "density": rho}
        src["p0"] = disc(shape, centre, 2)
        sens = {"mask": _sparse_mask(shape, rng), "record": ["p", "p_max", "I_avg"],
                "frequency_response": [2e6, 60]}
    elif name in ("cuboid_corners", "cuboid_corners_pml_outside"):
        # two overlapping cuboids of sensor points (1-based, inclusive)
        c, rho = hetero_medium(shape, 29)
        med = {"sound_speed": c, "density": rho}
        src["p0"] = disc(shape, centre, 3)
        lo1 = [n // 2 - 4 for n in shape]
        hi1 = [n // 2 + 1 for n in shape]
        lo2 = [n // 2 - 1 for n in shape]
        hi2 = [min(n, n // 2 + 5) for n in shape]
        corners = np.array([lo1 + hi1, lo2 + hi2], dtype=float).T
        rec = ["p", "p_max", "p_rms", "u", "u_non_staggered", "p_final", "u_max_all", "I_avg"]
        if name == "cuboid_corners_pml_outside":
            opts["PMLInside"] = False
            rec = ["p", "u_max", "p_max_all"]
        sens = {"mask": corners, "record": rec}
    elif name.startswith("cref_"):
        # medium.sound_speed_ref other than the default maximum
        # (kspaceFirstOrder_setSoundSpeedRef.m:30-50): kappa, the source
        # correction and the PML all follow c_ref
        c, rho = hetero_medium(shape, 53)
        ref = {"cref_min": "min", "cref_mean": "mean", "cref_max": "max", "cref_scalar": 1530.0}[name]
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5,
               "sound_speed_ref": ref}
        m = np.zeros(shape)
        m[tuple(slice(n // 2 - 1, n // 2 + 1) for n in shape)] = 1
        src = {"p": tone(Nt, dt, 2e6, 3), "p_mask": m}
        sens = {"mask": _sparse_mask(shape, rng, 0.3), "record": ["p", "p_max", "u_rms", "p_final"]}
    elif name.startswith("alpha_sign"):
        # medium.alpha_sign (kspaceFirstOrder_createAbsorptionVariables.m:
        # 78-81): sign(alpha_sign) multiplies tau and eta, sign(0) = 0
        c, rho = hetero_medium(shape, 59)
        sign = {"alpha_sign": [-1.0, 1.0], "alpha_sign_zero": [2.0, 0.0]}[name]
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 3.0, "alpha_power": 1.5,
               "alpha_sign": sign}
        src["p0"] = disc(shape, centre, 3)
    elif name in ("no_kspace", "no_sg"):
        # 'UsekSpace' false: kappa = 1 and no source correction filter;
        # 'UseSG' false: the gradients at the grid points and rho0 not
        # interpolated (kspaceFirstOrder3D.m:585-607,629-661)
        c, rho = hetero_medium(shape, 61)
        med = {"sound_speed": c, "density": rho, "alpha_coeff": 0.75, "alpha_power": 1.5}
        m = np.zeros(shape)
        m[tuple(slice(n // 2 - 1, n // 2 + 1) for n in shape)] = 1
        src = {"p": tone(Nt, dt, 2e6, 3), "p_mask": m,
               "ux": 1e-3 * tone(Nt, dt, 1.5e6, 2), "u_mask": m}
        opts["UsekSpace" if name == "no_kspace" else "UseSG"] = False
        sens = {"mask": _sparse_mask(shape, rng, 0.3),
                "record": ["p", "p_max", "u", "u_non_staggered", "p_final"]}
    elif name == "cartesian_nearest":
        src["p0"] = disc(shape, centre, 3)
        npt = 12
        ang = np.arange(npt) * 2 * math.pi / npt + math.pi / 2
        r = (min(shape) // 2 - pml - 1) * dx
        pts = [r * np.cos(ang), -r * np.sin(ang)][:dim]
        if dim == 3:
            pts.append(np.zeros(npt))
        sens = {"mask": np.stack(pts), "record": ["p", "p_max", "p_rms"]}
        opts["CartInterp"] = "nearest"
    else:
        raise KeyError(name)
    return {"shape": shape, "dx": dx, "Nt": Nt, "dt": dt, "medium": med, "source": src,
            "sensor": sens, "opts": opts, "dim": dim}


def _sparse_mask(shape, rng, frac=0.1):
    m = (rng.uniform(size=shape) < frac).astype(np.float64)
    m.ravel()[0] = 1
    if len(shape) == 1:
        # 1-D: also every other point near the sources, which sit a few
        # points inside the PML, so the sensors see more than rounding
        m[6:22:2] = 1
    return m


SIGNAL_OFF_LEN = 10

# source cases whose signals end before Nt
OFF_CASES = ["psrc_add_off", "psrc_dir_off", "psrc_nocorr_off", "psrc_many_off", "psrc_label_off",
             "psrc_many_dir_off", "usrc_add_off", "usrc_dir_off", "usrc_nocorr_off", "usrc_many_off"]
# options that change the operator tables
OPTION_CASES = ["cref_min", "cref_mean", "cref_scalar", "alpha_sign", "alpha_sign_zero",
                "no_kspace", "no_sg"]

CASES_2D = ["p0_lossless", "p0_hetero_absorbing", "p0_nonlinear_absorbing", "p0_nonlinear_lossless",
            "stokes", "no_dispersion", "no_absorption", "psrc_add", "psrc_dir", "psrc_nocorr",
            "psrc_many", "psrc_label", "psrc_many_dir", "usrc_add", "usrc_dir", "usrc_nocorr",
            "usrc_many", "time_reversal", "cartesian_nearest", "intensity", "u_nonstaggered", "u_split",
            "pml_outside", "pml_outside_source", "cartesian_linear", "cartesian_linear_pml_outside",
            "cuboid_corners", "cuboid_corners_pml_outside", "freq_response",
            "directivity_pressure", "directivity_gradient", "alpha_filter", "source_frequency_ref",
            "pml_auto"] + OFF_CASES + OPTION_CASES
# sensor directivity exists in 2-D only
CASES_3D = [c for c in CASES_2D if not c.startswith("directivity")]
# kspaceFirstOrder1D: every case whose inputs exist in 1-D
CASES_1D = ["p0_lossless", "p0_hetero_absorbing", "p0_nonlinear_absorbing", "p0_nonlinear_lossless",
            "stokes", "no_dispersion", "psrc_add", "psrc_dir", "psrc_nocorr", "psrc_many",
            "psrc_label", "usrc_add", "usrc_dir", "usrc_many", "time_reversal", "cartesian_nearest",
            "cartesian_linear", "intensity", "u_nonstaggered", "pml_outside", "cartesian_linear_pml_outside",
            "cuboid_corners", "cuboid_corners_pml_outside", "freq_response", "alpha_filter",
            "source_frequency_ref", "pml_auto"] + \
    [c for c in OFF_CASES if c[:-4] in ("psrc_add", "psrc_dir", "psrc_nocorr", "psrc_many", "psrc_label",
                                        "usrc_add", "usrc_dir", "usrc_many")] + OPTION_CASES


def oracle_opts(opts):
    """Map reference option names to the oracle's keyword names."""
    m = {"PMLSize": "pml_size", "PMLAlpha": "pml_alpha", "Smooth": "smooth", "PMLRange": "pml_range",
         "CartInterp": "cart_interp", "PMLInside": "pml_inside", "UsekSpace": "use_kspace",
         "UseSG": "use_sg"}
    return {m[k]: v for k, v in opts.items() if k in m}


def run_oracle(cs, data_cast="single"):
    from oracle import kwave_oracle as ko
    shape, dx = cs["shape"], cs["dx"]
    args = []
    for n in shape:
        args += [n, dx]
    kg = ko.KGrid(*args)
    kg.set_time(cs["Nt"], cs["dt"])
    return ko.kspace_first_order(kg, cs["medium"], cs["source"], cs["sensor"],
                                 data_cast=data_cast, **oracle_opts(cs["opts"]))


def run_kwhip(cs, backend):
    """The product on the inputs of a case; backend "emu" runs the same host
    code over the CPU lane emulator (test harness only)."""
    import kwhip
    from kwhip import _abi
    shape, dx = cs["shape"], cs["dx"]
    args = []
    for n in shape:
        args += [n, dx]
    kg = kwhip.kWaveGrid(*args)
    kg.setTime(cs["Nt"], cs["dt"])
    fn = {1: kwhip.kspaceFirstOrder1D, 2: kwhip.kspaceFirstOrder2D, 3: kwhip.kspaceFirstOrder3D}[cs["dim"]]
    with _abi._backend_override(backend):
        return fn(kg, cs["medium"], cs["source"], cs["sensor"], **cs["opts"])


def compare(got, ref):
    """Relative L-inf per output field; returns {field: err}."""
    from oracle import kwave_oracle as ko
    if isinstance(ref, list):
        # cuboid corners: one record per cuboid
        assert isinstance(got, list) and len(got) == len(ref)
        errs = {}
        for q, (g, r) in enumerate(zip(got, ref)):
            assert set(g) == set(r), (q, sorted(g), sorted(r))
            for k, e in compare(g, r).items():
                assert np.shape(g[k]) == np.shape(r[k]), (q, k)
                errs[f"{q}.{k}"] = e
        return errs
    if not isinstance(ref, dict):
        got, ref = {"p": got}, {"p": ref}
    errs = {}
    for k, r in ref.items():
        g = got[k]
        errs[k] = ko.rel_linf(np.asarray(g, dtype=np.float64), np.asarray(r, dtype=np.float64))
    return errs


def prepared(cs):
    """The kw_problem of a case (the host precompute of kspaceFirstOrder1D/
    2D/3D), for calls straight through the C ABI."""
    import kwhip
    from kwhip import solver
    shape, dx = cs["shape"], cs["dx"]
    args = []
    for n in shape:
        args += [n, dx]
    kg = kwhip.kWaveGrid(*args)
    kg.setTime(cs["Nt"], cs["dt"])
    opts = solver._parse_options((), dict(cs["opts"]))
    return solver.prepare(kg, cs["medium"], cs["source"], cs["sensor"], cs["dim"], opts)


def run_entry(lib, fn_name, pp):
    """Run a prepared problem through the C-ABI entry point `fn_name`
    (kw_run or kw_run_noalloc) of `lib`; outputs as a dict of fields."""
    import ctypes as C
    from kwhip import _abi, solver
    out, arrays = solver.alloc_outputs(pp)
    _abi.call(getattr(lib, fn_name), C.byref(pp.prob), C.byref(out))
    res = solver.shape_outputs(pp, arrays)
    return res if isinstance(res, dict) else {"p": res}
