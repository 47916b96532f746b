"""Cases of kspaceFirstOrderAS (the axisymmetric solver), shared by the
emulator and GPU tests: x axial, y radial from the axis; 'WSWA-FFT' unless
a case says otherwise."""

import numpy as np


def case(name, Nx=32, Ny=16, seed=2, Nt=24, sym="WSWA-FFT"):
    rng = np.random.default_rng(seed)
    dx = 1e-4
    x = (np.arange(Nx) - Nx // 2) * dx
    r = np.arange(Ny) * dx
    X, R = np.meshgrid(x, r, indexing="ij")
    c = 1500 + 80 * rng.random((Nx, Ny))
    rho = 1000 + 90 * rng.random((Nx, Ny))
    cs = {"Nx": Nx, "Ny": Ny, "dx": dx, "Nt": Nt,
          "medium": {"sound_speed": c, "density": rho},
          "p0": np.exp(-(X ** 2 + R ** 2) / (3 * dx) ** 2),
          "sensor": {"mask": rng.random((Nx, Ny)) < 0.3,
                     "record": ["p", "p_max", "p_rms", "p_final", "u", "u_final"]},
          "pml": [6, 4], "smooth": False, "sym": sym}
    cs["dt"] = 0.25 * dx / c.max()
    cs["source"] = {"p0": cs.pop("p0")}
    t = np.arange(cs["Nt"]) * cs["dt"]
    burst = np.sin(2 * np.pi * t / (8 * cs["dt"])) * np.exp(-((t - 8 * cs["dt"]) / (5 * cs["dt"])) ** 2)
    src_mask = rng.random((Nx, Ny)) < 0.06
    src_mask[Nx // 2, 0] = src_mask[Nx // 3, 0] = True      # points on the axis
    n_src = int(src_mask.sum())
    if name == "homogeneous":
        cs["medium"] = {"sound_speed": 1500.0, "density": 1000.0}
        cs["dt"] = 0.25 * dx / 1500.0
    elif name == "nonlinear":
        cs["medium"]["BonA"] = 6 + rng.random((Nx, Ny))
        cs["source"]["p0"] = cs["source"]["p0"] * 1e6
    elif name == "stokes":
        cs["medium"]["alpha_coeff"] = 0.5 + rng.random((Nx, Ny))
    elif name == "stokes_nonlinear":
        cs["medium"]["alpha_coeff"] = 0.75
        cs["medium"]["BonA"] = 5.0
        cs["source"]["p0"] = cs["source"]["p0"] * 1e6
    elif name == "smoothed":
        cs["smooth"] = True
        cs["source"]["p0"] = (np.sqrt(X ** 2 + R ** 2) < 4 * dx).astype(float)
    elif name == "source_additive":
        # one series per point
        cs["source"] = {"p_mask": src_mask, "p_mode": "additive",
                        "p": rng.standard_normal((n_src, 1)) * burst}
    elif name == "source_dirichlet":
        cs["source"] = {"p_mask": src_mask, "p_mode": "dirichlet", "p": burst[None, :12]}
    elif name == "source_labelled":
        lab = np.zeros((Nx, Ny))
        lab[src_mask] = 1 + (rng.random(n_src) < 0.5)
        cs["source"] = {"p_mask": lab, "p_mode": "additive-no-correction",
                        "p": np.stack([burst, -0.5 * burst])}
    elif name == "source_frequency_ref":
        cs["source"] = {"p_mask": src_mask, "p": burst[None], "p_frequency_ref": 1e6}
    elif name == "source_stokes_nonlinear":
        cs["medium"]["alpha_coeff"] = 0.5 + rng.random((Nx, Ny))
        cs["medium"]["BonA"] = 6.0
        cs["source"] = {"p_mask": src_mask, "p": 1e6 * burst[None]}
    elif name == "velocity_additive":
        # axial and radial velocity sources, one series per point
        cs["source"] = {"u_mask": src_mask, "u_mode": "additive",
                        "ux": rng.standard_normal((n_src, 1)) * burst,
                        "uy": rng.standard_normal((n_src, 1)) * burst}
    elif name == "velocity_dirichlet":
        cs["source"] = {"u_mask": src_mask, "u_mode": "dirichlet", "ux": 1e-4 * burst[None, :12],
                        "uy": -1e-4 * burst[None, :12]}
    elif name == "velocity_labelled":
        lab = np.zeros((Nx, Ny))
        lab[src_mask] = 1 + (rng.random(n_src) < 0.5)
        cs["source"] = {"u_mask": lab, "u_mode": "additive-no-correction",
                        "uy": np.stack([burst, -0.5 * burst])}
    elif name == "velocity_radial_additive":
        # the k-space filtered radial source, points up to the last radial
        # column (whose 'WSWS' value takes the filtered source too)
        m = src_mask.copy()
        m[Nx // 2, Ny - 2:] = True
        cs["source"] = {"u_mask": m, "u_mode": "additive",
                        "uy": rng.standard_normal((int(m.sum()), 1)) * burst}
    elif name in ("velocity_radial_dirichlet", "velocity_radial_nocorr"):
        # radial (and axial) velocity sources away from the last two radial
        # columns, which 'WSWS-FFT' takes in these modes
        m = src_mask.copy()
        m[:, Ny - 2:] = False
        n_m = int(m.sum())
        if name == "velocity_radial_dirichlet":
            cs["source"] = {"u_mask": m, "u_mode": "dirichlet", "uy": 1e-4 * burst[None, :12],
                            "ux": -5e-5 * burst[None, :12]}
        else:
            cs["source"] = {"u_mask": m, "u_mode": "additive-no-correction",
                            "uy": rng.standard_normal((n_m, 1)) * burst}
    elif name == "velocity_axial_only":
        cs["medium"]["alpha_coeff"] = 0.5 + rng.random((Nx, Ny))
        cs["source"] = {"u_mask": src_mask, "ux": burst[None], "u_frequency_ref": 1e6}
    elif name == "pml_outside":
        # the PML outside the grid, smoothed p0, Stokes absorption
        cs["pml_inside"] = False
        cs["smooth"] = True
        cs["medium"]["alpha_coeff"] = 0.5 + rng.random((Nx, Ny))
        cs["sensor"]["record"] = cs["sensor"]["record"] + ["p_max_all"]
    elif name == "pml_outside_source":
        cs["pml_inside"] = False
        cs["source"] = {"p_mask": src_mask, "p": burst[None]}
    elif name in ("cartesian_linear", "cartesian_nearest", "cartesian_tr"):
        # Cartesian points (x, r): distinct grid nodes plus offsets below half
        # a cell, so the nearest nodes do not repeat
        n_pts = 40
        cells = rng.choice((Nx - 12) * (Ny - 6), n_pts, replace=False)
        ix, iy = cells % (Nx - 12) + 6, cells // (Nx - 12)
        xs = (ix - Nx // 2 + rng.uniform(-0.45, 0.45, n_pts)) * dx
        rs = (iy + rng.uniform(0, 0.45, n_pts)) * dx
        cart = np.stack([xs, rs])
        cs["cart_interp"] = "nearest" if name == "cartesian_nearest" else "linear"
        if name == "cartesian_tr":
            cs["source"] = {}
            cs["sensor"] = {"mask": cart, "time_reversal_boundary_data":
                            rng.standard_normal((n_pts, 1)) * burst}
        else:
            # point records only: the final field is normalised by the largest
            # recorded p_max, which a few points need not see
            cs["sensor"] = {"mask": cart, "record": ["p", "p_max", "p_rms", "u"]}
    elif name == "time_reversal":
        sm = cs["sensor"]["mask"]
        cs["source"] = {}
        cs["sensor"] = {"mask": sm, "time_reversal_boundary_data":
                        rng.standard_normal((int(sm.sum()), 1)) * burst}
    elif name != "hetero":
        raise KeyError(name)
    return cs


NAMES = ["homogeneous", "hetero", "nonlinear", "stokes", "stokes_nonlinear", "smoothed",
         "source_additive", "source_dirichlet", "source_labelled", "source_frequency_ref",
         "source_stokes_nonlinear", "time_reversal", "velocity_additive", "velocity_dirichlet",
         "velocity_labelled", "velocity_axial_only", "pml_outside", "pml_outside_source",
         "cartesian_linear", "cartesian_nearest", "cartesian_tr", "velocity_radial_dirichlet",
         "velocity_radial_nocorr", "velocity_radial_additive"]
# the cases run with 'WSWS-FFT' as well (it takes 'dirichlet' and
# 'additive-no-correction' radial velocity sources away from the last two
# radial columns)
WS_NAMES = ["hetero", "stokes_nonlinear", "smoothed", "source_additive", "source_dirichlet",
            "time_reversal", "velocity_axial_only", "pml_outside", "cartesian_linear",
            "velocity_radial_dirichlet", "velocity_radial_nocorr", "velocity_radial_additive"]


def run_kwhip(cs):
    import kwhip
    kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.setTime(cs["Nt"], cs["dt"])
    out = kwhip.kspaceFirstOrderAS(kg, dict(cs["medium"]), dict(cs["source"]), dict(cs["sensor"]),
                                   PMLSize=cs["pml"], Smooth=cs["smooth"],
                                   RadialSymmetry=cs.get("sym", "WSWA-FFT"),
                                   PMLInside=cs.get("pml_inside", True),
                                   CartInterp=cs.get("cart_interp", "linear"))
    return {"p_final": out} if "time_reversal_boundary_data" in cs["sensor"] else out


def run_oracle(cs, data_cast):
    from oracle import kwave_oracle as ko
    kg = ko.KGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.set_time(cs["Nt"], cs["dt"])
    sym = "WSWS-FFT" if cs.get("sym", "WSWA-FFT").startswith("WSWS") else "WSWA-FFT"
    return ko.kspace_first_order_as(kg, cs["medium"], cs["source"], cs["sensor"],
                                    data_cast=data_cast, pml_size=cs["pml"], smooth=cs["smooth"],
                                    radial_symmetry=sym, pml_inside=cs.get("pml_inside", True),
                                    cart_interp=cs.get("cart_interp", "linear"))


def compare(got, ref):
    """Relative L-inf per output.  The final fields are normalised by the
    run's peak (p_final by the largest recorded p_max, u_final by the
    largest recorded velocity): a spreading wave leaves them far below the
    peak, where single-precision implementations differ by more than 1e-5
    of their own maximum (the oracle's single- and double-precision runs
    of the 128 x 64 cases differ by 1.5e-5 there)."""
    errs = {}
    peak = {"p_final": np.max(np.abs(ref["p_max"])) if "p_max" in ref else None,
            "ux_final": max(np.max(np.abs(ref["ux"])), np.max(np.abs(ref["uy"])))
            if "ux" in ref else None}
    peak["uy_final"] = peak["ux_final"]
    for k, r in ref.items():
        g = np.asarray(got[k], dtype=np.float64)
        r = np.asarray(r, dtype=np.float64)
        if g.shape != r.shape:
            g = g.reshape(r.shape, order="F")
        scale = peak.get(k) or np.max(np.abs(r))
        errs[k] = float(np.max(np.abs(g - r)) / scale)
    return errs


def spherical_wave(Nx, Ny, dx, c0, Nt, dt, width):
    """p0 = exp(-R^2 / w^2) centred on the axis, zero velocity: the exact
    pressure [(R - ct) g(R - ct) + (R + ct) g(R + ct)] / (2R) at t = (Nt - 1)
    dt (the first step is the initial condition)."""
    x = (np.arange(Nx) - Nx // 2) * dx
    r = np.arange(Ny) * dx
    X, Rr = np.meshgrid(x, r, indexing="ij")
    R = np.sqrt(X ** 2 + Rr ** 2)
    g = lambda s: np.exp(-s ** 2 / width ** 2)  # noqa: E731
    T = (Nt - 1) * dt
    with np.errstate(invalid="ignore", divide="ignore"):
        ex = ((R - c0 * T) * g(R - c0 * T) + (R + c0 * T) * g(R + c0 * T)) / (2 * R)
    ex[R == 0] = (1 - 2 * (c0 * T) ** 2 / width ** 2) * g(c0 * T)
    return g(R), ex, (np.abs(X) < (Nx / 2 - 20) * dx) & (Rr < (Ny - 20) * dx)


def save_and_run_file(cs, tmp_path, lib, record=("p", "p_max", "p_rms", "p_final", "u", "u_final"),
                      matlab_style=False, **opts):
    """Write the case with kspaceFirstOrderAS(..., 'SaveToDisk'), optionally
    re-write it the way MATLAB writes it (chunked, deflated), run it with
    kw_run_file and return the output file's records (h5lite) and the input
    file's variables."""
    import ctypes as C
    import sys
    from pathlib import Path
    import kwhip
    from kwhip import _abi
    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    import h5lite
    kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.setTime(cs["Nt"], cs["dt"])
    fn = tmp_path / "as_in.h5"
    r = kwhip.kspaceFirstOrderAS(kg, dict(cs["medium"]), dict(cs["source"]), dict(cs["sensor"]),
                                 PMLSize=cs["pml"], Smooth=cs["smooth"], SaveToDisk=str(fn), **opts)
    assert r == []
    v = h5lite.read(fn)
    if matlab_style:
        variables = {k: (a, "uint64" if a.dtype == np.uint64 else "single") for k, a in v.items()}
        h5lite.write_input(tmp_path / "as_matlab.h5", variables, h5lite.read_attrs(fn), compression=4)
        fn = tmp_path / "as_matlab.h5"
    flags = 0
    for k in record:
        flags |= _abi.REC[k]
    _abi.call(lib.kw_run_file, str(fn).encode(), str(tmp_path / "as_out.h5").encode(),
              C.c_uint32(flags), C.c_int64(1))
    return h5lite.read(tmp_path / "as_out.h5"), v


def file_records(out, Nx, Ny):
    """The output file's records under the names of the direct call."""
    res = {}
    for k, a in out.items():
        if k in ("Nx", "Ny", "Nz", "Nt", "pml_x_size", "pml_y_size", "pml_z_size", "dt"):
            continue
        a = np.squeeze(np.asarray(a, dtype=np.float64))
        if k.endswith("_final"):
            a = a.reshape(Nx, Ny)
        res[k] = a
    return res
