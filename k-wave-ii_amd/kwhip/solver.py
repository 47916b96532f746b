"""Drop-in kspaceFirstOrder1D / 2D / 3D over the kwhip C ABI.

Mirrors the reference call signature

    sensor_data = kspaceFirstOrder2D(kgrid, medium, source, sensor, 'Name', value, ...)

(+kwave/+legacy/kspaceFirstOrder1D.m:1, kspaceFirstOrder2D.m:1,
kspaceFirstOrder3D.m:1).  ``medium``,
``source`` and ``sensor`` are dicts (or attribute objects) with the
reference's field names; name-value options are given as keyword arguments
or as trailing ('Name', value) pairs.

This module does the reference's host-side precompute -- input checking,
reference sound speed, time array, smoothing (the transform runs on the
device through ``kw_smooth``), staggered density, source scaling, sensor and
source index lists, Cartesian-to-grid mapping -- and then hands the variable
set of the reference's native-offload contract to the device library, which
builds the operators and runs the time loop.  Output selection follows
kspaceFirstOrder3D.m:1360-1383.

Precision: the device computes fields and FFTs in float32 with operators
built in float64 and rounded once, and keeps the statistics (max, min, rms)
in float64.  Time series and final fields are returned as float32,
statistics as float64.
"""

from __future__ import annotations

import ctypes as C
import math
import warnings
from types import SimpleNamespace

import numpy as np

from . import _abi
from .kgrid import kWaveGrid

# options accepted for compatibility that only affect plotting / logging
_VISUAL_OPTIONS = {"PlotSim", "PlotLayout", "PlotPML", "PlotScale", "PlotFreq", "DisplayMask",
                   "LogScale", "RecordMovie", "MovieName", "MovieArgs", "MovieProfile",
                   "CreateLog", "MeshPlot"}
_OPTIONS = {"PMLSize", "PMLAlpha", "PMLInside", "PMLRange", "MultiAxialPMLRatio", "UseFD",
            "Smooth", "DataCast", "DataRecast",
            "CartInterp", "UsekSpace", "UseSG", "ScaleSourceTerms", "SaveToDisk",
            "SaveToDiskExit", "HDFCompressionLevel", "RadialSymmetry",
            "StreamToDisk"} | _VISUAL_OPTIONS
_UNSUPPORTED = set()
# kspaceFirstOrder_setDefaults.m:196
SAVE_TO_DISK_FILENAME_DEF = "kwave_input_data.h5"

RECORD_ALLOWED = ("p", "p_max", "p_min", "p_rms", "p_max_all", "p_min_all", "p_final",
                  "u", "u_max", "u_min", "u_rms", "u_max_all", "u_min_all", "u_final",
                  "u_non_staggered", "I", "I_avg", "u_split_field")


# allowed fields of the input structures (kspaceFirstOrder_inputChecking.m:
# 148-149, 302-350, 729-741), checked as checkFieldNames.m does
_MEDIUM_FIELDS = {"sound_speed", "sound_speed_ref", "density", "alpha_coeff", "alpha_power",
                  "alpha_mode", "alpha_filter", "alpha_sign", "BonA"}
_SENSOR_FIELDS = {"mask", "time_reversal_boundary_data", "frequency_response", "record_mode",
                  "record", "record_start_index"}
_DIRECTIVITY_FIELDS = {"directivity_pattern", "directivity_angle", "directivity_size"}


def _check_fields(obj, allowed, what):
    if obj is None:
        return
    names = obj.keys() if isinstance(obj, dict) else vars(obj).keys()
    for nm in names:
        if nm not in allowed:
            raise ValueError(f"{nm} is not a valid field for the structure {what}.")


def _check_inputs(dim, medium, source, sensor):
    _check_fields(medium, _MEDIUM_FIELDS, "medium")
    src = {"p0", "p", "p_mask", "p_mode", "p_frequency_ref", "u_mask", "u_mode",
           "u_frequency_ref"} | set(["ux", "uy", "uz"][:dim])
    _check_fields(source, src, "source")
    _check_fields(sensor, _SENSOR_FIELDS | (_DIRECTIVITY_FIELDS if dim == 2 else set()), "sensor")

def _get(obj, name, default=None):
    if obj is None:
        return default
    if isinstance(obj, dict):
        v = obj.get(name, default)
    else:
        v = getattr(obj, name, default)
    return default if v is None else v


def _has(obj, name):
    return _get(obj, name) is not None


def _parse_options(args, kwargs):
    opts = dict(kwargs)
    if len(args) % 2:
        raise ValueError("Optional inputs must be entered as Param, Value pairs.")
    for i in range(0, len(args), 2):
        opts[args[i]] = args[i + 1]
    for k in opts:
        if k in _UNSUPPORTED:
            raise NotImplementedError(f"option '{k}' is not supported by kwhip")
        if k not in _OPTIONS:
            raise ValueError(f"Unknown optional input '{k}'.")
    if "MultiAxialPMLRatio" in opts:
        # kspaceFirstOrder_checkOptionalInputs.m:195-203: checked, used by the
        # elastic solvers only
        r = np.asarray(opts["MultiAxialPMLRatio"])
        if r.size != 1 or not np.issubdtype(r.dtype, np.number) or float(r) < 0:
            raise ValueError("Optional input 'MultiAxialPMLRatio' must be a single positive value.")
    return opts


def _check_use_fd(dim, opts):
    """kspaceFirstOrder_checkOptionalInputs.m:440-456: 'UseFD' exists for the
    1-D fluid code only; false keeps the spectral gradients."""
    if "UseFD" not in opts:
        return
    if dim != 1:
        raise ValueError("Optional input 'UseFD' only supported in 1D.")
    v = opts["UseFD"]
    if isinstance(v, (bool, np.bool_)) and not v:
        return
    if v in (2, 4):
        raise NotImplementedError("finite-difference gradients ('UseFD' 2 or 4) are not supported by kwhip")
    raise ValueError("Optional input 'UseFD' must be set to 2, 4, or false.")


# ---------------------------------------------------------------------------
# smoothing window (+kwave/+legacy/getWin.m, Blackman, rotated)
# ---------------------------------------------------------------------------

def _blackman(N: int, a: float = 0.16) -> np.ndarray:
    # getWin.m:230-231 and cosineSeries (:417-424)
    n = np.arange(N, dtype=np.float64)
    coeffs = [(1 - a) / 2, 0.5, a / 2]
    w = coeffs[0] * np.ones(N)
    for i in range(1, 3):
        w = w + (-1) ** i * coeffs[i] * np.cos(i * 2 * math.pi * n / (N - 1))
    return w


def blackman_window(grid_size) -> np.ndarray:
    """abs(getWin(size, 'Blackman', 'Rotation', true, 'Symmetric', rem(size,2)))
    as used by smooth.m: odd sizes symmetric; even sizes built one longer and
    trimmed (getWin.m:188, :300-347)."""
    size = np.asarray(grid_size, dtype=np.int64)
    symmetric = (size % 2).astype(bool)
    N = size + (~symmetric).astype(np.int64)
    L = int(N.max())
    lin = _blackman(L)
    if len(size) == 1:
        # 1-D: no rotation, the window trimmed to the grid (getWin.m:240-256)
        return np.abs(lin[:int(size[0])])
    radius = (L - 1) / 2.0
    ref = np.linspace(-radius, radius, L)
    axes = [np.linspace(-radius, radius, int(n)) for n in N]
    r2 = np.zeros(tuple(int(n) for n in N))
    for ax, v in enumerate(axes):
        sh = [1] * len(N)
        sh[ax] = -1
        r2 = r2 + v.reshape(sh) ** 2
    r = np.sqrt(r2)
    r[r > radius] = radius
    win = np.interp(r, ref, lin)
    sl = tuple(slice(0, int(s)) for s in size)
    return np.abs(win[sl])


def smooth(mat, restore_max=False):
    """+kwave/+legacy/smooth.m: real(ifftn(fftn(mat) .* ifftshift(win))) with
    the rotated Blackman window, optionally restoring the maximum magnitude.

    This is input preparation, done in double precision on the host as the
    reference's MATLAB precompute does before it hands the smoothed p0 (and
    c, rho) to its native engines (kspaceFirstOrder_inputChecking.m:
    1501-1543); the values are rounded to single once, when the device
    copies them.  ``kw_smooth`` in the C ABI is the single-precision device
    version of the same filter for callers of the library."""
    import scipy.fft as sfft

    mat = np.asarray(mat, dtype=np.float64)
    if mat.ndim not in (1, 2, 3):
        raise ValueError("smooth supports 1D, 2D and 3D grids")
    win = blackman_window(mat.shape)
    workers = _abi.host_threads()
    res = np.real(sfft.ifftn(sfft.fftn(mat, workers=workers) * np.fft.ifftshift(win),
                             workers=workers))
    if restore_max:
        res = (np.max(np.abs(mat)) / np.max(np.abs(res))) * res
    return res


# ---------------------------------------------------------------------------
# Cartesian sensor masks (+kwave/+legacy/cart2grid.m, nearest)
# ---------------------------------------------------------------------------

def cart2grid(kgrid: kWaveGrid, cart_data, radial_ny=None):
    """cart2grid.m:59-194: nearest grid point with MATLAB round (half away
    from zero), later points overwrite earlier ones.  Returns the binary mask,
    order_index and reorder_index (1-based).  radial_ny: the axisymmetric
    form (:111-116), y the radius from the first grid row, which lies in
    the first radial_ny rows."""
    cart = np.asarray(cart_data, dtype=np.float64)
    if cart.ndim != 2 or cart.shape[0] != kgrid.dim:
        raise ValueError(f"Cartesian sensor.mask for a {kgrid.dim}D simulation must be given as a {kgrid.dim} by N array.")
    idx = []
    for ax in range(kgrid.dim):
        v = cart[ax] / kgrid.spacings[ax]
        v = np.sign(v) * np.floor(np.abs(v) + 0.5)
        radial = radial_ny is not None and ax == 1
        idx.append(v.astype(np.int64) + (1 if radial else kgrid.shape[ax] // 2 + 1))
        if idx[ax].max() > (radial_ny if radial else kgrid.shape[ax]) or idx[ax].min() < 1:
            raise ValueError("Cartesian points must lie within the grid defined by kgrid.")
    grid = np.zeros(kgrid.shape, dtype=np.int64)
    for p in range(cart.shape[1]):
        grid[tuple(int(idx[ax][p]) - 1 for ax in range(kgrid.dim))] = p + 1
    flat = grid.ravel(order="F")
    reorder_index = flat[flat != 0]
    order_index = np.argsort(reorder_index, kind="stable") + 1
    return (grid != 0), order_index, reorder_index


def optimal_pml_size(grid_size, pml_range=(10, 40), axisymmetric=None):
    """getOptimalPMLSize.m: for each axis the PML size in the inclusive range
    whose grown length N + 2 pml has the smallest largest prime factor (the
    smallest such size).  axisymmetric 'WSWA' / 'WSWS' (:124-133): the
    radial axis grows by one PML and is scored by its expanded length,
    4 (N + pml) or 2 (N + pml) - 2."""
    # MATLAB round: halves away from zero (10.5 -> 11)
    lo, hi = (int(math.copysign(math.floor(abs(float(v)) + 0.5), float(v))) for v in pml_range)
    if lo < 0 or hi < 0:
        raise ValueError("Optional input pml_range must be positive.")
    if hi <= lo:
        raise ValueError("The second value for pml_range must be greater than the first.")

    def largest_prime_factor(n):
        f, last = 2, 1
        while n > 1 and f * f <= n:
            if n % f == 0:
                last = f
                n //= f
            else:
                f += 1
        return max(last, n)

    def grown(ax, N, p):
        if axisymmetric and ax == 1:
            return 4 * (N + p) if axisymmetric == "WSWA" else 2 * (N + p) - 2
        return N + 2 * p

    sizes = []
    for ax, N in enumerate(grid_size):
        best = min(range(lo, hi + 1), key=lambda p: (largest_prime_factor(grown(ax, N, p)), p))
        sizes.append(best)
    return sizes


def cuboid_points(kgrid: kWaveGrid, corners):
    """Sensor points of a cuboid-corners mask: columns [x1, (y1, (z1,)) x2,
    (y2, (z2))], 1-based and inclusive; the points of each cuboid in
    column-major order, cuboid after cuboid (overlaps repeat points)
    (kspaceFirstOrder_inputChecking.m:519-576 checks, :1610-1639 indices).
    Returns the 0-based indices and the integer corners."""
    cor = np.asarray(corners, dtype=np.float64)
    dim, shape = kgrid.dim, kgrid.shape
    if not np.all(np.mod(cor, 1) == 0):
        raise ValueError("sensor.mask cuboid corner indices must be integers.")
    cor = cor.astype(np.int64)
    if np.any(cor[dim:] < cor[:dim]):
        raise ValueError("sensor.mask cuboid corners must be defined as [x1, ..., x2, ...].' "
                         "where x2 => x1, etc.")
    top = np.array(shape, dtype=np.int64)[:, None]
    if np.any(cor < 1) or np.any(cor[:dim] > top) or np.any(cor[dim:] > top):
        raise ValueError("sensor.mask cuboid corners must be within the grid.")
    strides = np.cumprod([1] + list(shape[:-1])).astype(np.int64)
    parts = []
    for q in range(cor.shape[1]):
        # column-major enumeration of the box: x fastest
        axes = [np.arange(cor[a, q] - 1, cor[a + dim, q]) * strides[a] for a in range(dim)]
        grids = np.meshgrid(*axes, indexing="ij")
        parts.append(sum(grids).ravel(order="F"))
    return np.concatenate(parts).astype(np.int64), cor


def _split_cuboids(sd: dict, corners, dim):
    """kspaceFirstOrder_reorderCuboidCorners.m: one record per cuboid, the
    sensor rows reshaped (column-major) to the cuboid, time last; the
    whole-grid outputs in the first record."""
    out = []
    start = 0
    for q in range(corners.shape[1]):
        size = tuple(int(corners[a + dim, q] - corners[a, q] + 1) for a in range(dim))
        n = int(np.prod(size))
        d = {}
        for k, v in sd.items():
            if not (k.endswith("_all") or k.endswith("_final")):
                v = np.asarray(v)
                d[k] = v[start:start + n].reshape(size + v.shape[1:], order="F")
        out.append(d)
        start += n
    out[0].update({k: v for k, v in sd.items() if k.endswith("_all") or k.endswith("_final")})
    return out


def grid_data_linear(kgrid: kWaveGrid, cart_data, radial_ny=None):
    """record.tri / record.bc for 'CartInterp' 'linear'
    (kspaceFirstOrder_createStorageVariables.m:376-398 with gridDataFast2D.m /
    gridDataFast3D.m): for each Cartesian point the simplex of a Delaunay
    triangulation of the grid points that holds it, and the point's
    barycentric coordinates there.

    A regular grid has many Delaunay triangulations (the corners of every
    cell are co-spherical) and MATLAB's choice among them is
    implementation-defined; this one is declared: the Kuhn subdivision of
    each cell.  For a point with cell corner i and local coordinates f, with
    the axes ordered by decreasing f (ties in axis order), the simplex walks
    from i one unit step along each axis in that order, and the weights are
    1 - f[s1], f[s1] - f[s2], ..., f[s_dim] (in 2-D: cells cut along the
    diagonal from (i, j) to (i+1, j+1)).

    Returns tri ([points, dim+1], 1-based column-major grid indices) and bc
    (same shape, float64).  radial_ny: the axisymmetric grid, y the radius
    from the first grid row (kspaceFirstOrder_createStorageVariables.m:
    390-391 shifts y by its smallest value), inside the first radial_ny
    rows."""
    cart = np.asarray(cart_data, dtype=np.float64)
    dim, shape = kgrid.dim, kgrid.shape
    npts = cart.shape[1]
    # grid point n (0-based) of axis a sits at (n - floor(N_a/2)) d_a
    u = np.stack([cart[a] / kgrid.spacings[a] + (0 if radial_ny is not None and a == 1
                                                   else shape[a] // 2) for a in range(dim)])
    top = np.array(shape, dtype=np.float64)[:, None] - 1
    if radial_ny is not None:
        top[1] = radial_ny - 1
    if np.any(u < -1e-9) or np.any(u > top + 1e-9):
        raise ValueError("Cartesian points must lie within the k-space grid defined by kgrid.")
    corner = np.clip(np.floor(u), 0, np.maximum(top - 1, 0)).astype(np.int64)
    f = np.clip(u - corner, 0.0, 1.0)
    order = np.argsort(-f, axis=0, kind="stable")
    levels = np.vstack([np.ones((1, npts)), np.take_along_axis(f, order, axis=0),
                        np.zeros((1, npts))])
    bc = (levels[:-1] - levels[1:]).T
    strides = np.cumprod([1] + list(shape[:-1])).astype(np.int64)
    first = (corner * strides[:, None]).sum(axis=0)
    tri = np.vstack([first, first + np.cumsum(strides[order], axis=0)]).T
    return tri + 1, bc


def check_stability(kgrid: kWaveGrid, c, c_ref, medium):
    """+kwave/+legacy/checkStability.m:60-144: the largest time step for
    which the scheme is stable.  Lossless media: unconditionally stable
    (inf) when c_ref >= max(c), else 2/(c_ref k_max) asin(c_ref/max(c)).
    Absorbing media: the estimate with kappa = 1, refined by the fixed-point
    iteration on kappa(dt) = sinc(c_ref k_max dt/2).  k_max is the largest
    wavenumber magnitude on the grid."""
    vecs = [kgrid.kx_vec, kgrid.ky_vec, kgrid.kz_vec][:kgrid.dim]
    kmax = math.sqrt(sum(float(np.max(np.abs(v))) ** 2 for v in vecs))
    c_max = float(np.max(c))
    alpha = _get(medium, "alpha_coeff")
    if alpha is None or np.all(np.asarray(alpha) == 0):
        if c_ref >= c_max:
            return math.inf
        return 2.0 / (c_ref * kmax) * math.asin(c_ref / c_max)
    mode = _get(medium, "alpha_mode")
    y = 2.0 if mode == "stokes" else float(np.asarray(_get(medium, "alpha_power")))
    a_np = 100.0 * np.asarray(alpha, dtype=np.float64) * (1e-6 / (2 * math.pi)) ** y / \
        (20.0 * math.log10(math.e))
    tau = 0.0 if mode == "no_absorption" else -2.0 * a_np * c ** (y - 1)
    eta = 0.0 if mode == "no_dispersion" else 2.0 * a_np * c ** y * math.tan(math.pi * y / 2)
    tau_min, eta_min = float(np.min(tau)), float(np.min(eta))
    temp1 = c_max * tau_min * kmax ** (y - 1)
    temp2 = 1.0 - eta_min * kmax ** (y - 1)
    dt = (temp1 + math.sqrt(temp1 ** 2 + 4 * temp2)) / (temp2 * kmax * c_max)

    def kappa(t):
        x = c_ref * kmax * t / 2
        return 1.0 if x == 0 else math.sin(x) / x

    def func(t):
        t3 = c_max * tau_min * kappa(t) * kmax ** (y - 1)
        return (t3 + math.sqrt(t3 ** 2 + 4 * temp2)) / (temp2 * kmax * kappa(t) * c_max)

    # the reference iterates until the step changes by at most 1e-12 s (the
    # iteration converges slowly when absorption is weak); the cap only
    # guards against a non-converging input
    dt_old = 0.0
    for _ in range(1000000):
        if abs(dt - dt_old) <= 1e-12:
            break
        dt_old, dt = dt, func(dt)
    return dt


def _find(mask) -> np.ndarray:
    """Column-major linear indices of the nonzero cells of a grid (MATLAB's
    find).  Sparse masks are searched in memory order and the indices
    sorted, which avoids a transposed copy of the whole grid."""
    m = np.asarray(mask)
    if m.ndim > 1 and np.count_nonzero(m) * 8 < m.size:
        lin = np.ravel_multi_index(np.nonzero(m), m.shape, order="F")
        lin.sort()
        return lin
    return np.flatnonzero(np.asarray(m != 0).ravel(order="F"))


# ---------------------------------------------------------------------------
# problem preparation
# ---------------------------------------------------------------------------

class PreparedProblem:
    """A kw_problem plus the arrays it points to and the post-processing
    information; keeps everything alive while the device library uses it."""

    def __init__(self):
        self.keep = []
        self.prob = _abi.Problem()
        self.info = SimpleNamespace()

    def hold(self, a):
        if a is not None:
            self.keep.append(a)
        return a


def _source_series(src, nm, n_points):
    s = np.atleast_2d(np.asarray(_get(src, nm), dtype=np.float64)).copy()
    return s


def _expand(a, sizes, edge):
    """expandMatrix.m: pad by sizes[d] on both sides of axis d, replicating
    the edges (edge=True) or with zeros."""
    a = np.asarray(a)
    pad = [(int(s), int(s)) for s in sizes]
    if edge:
        return np.pad(a, pad, mode="edge")
    return np.pad(a, pad, mode="constant", constant_values=0)


def _pml_outside(kgrid, medium, source, sensor, dim, opts):
    """'PMLInside', false (kspaceFirstOrder_inputChecking.m:1501-1560,
    kspaceFirstOrder_expandGridMatrices.m): the time axis is set on the
    user's grid, p0 is smoothed there, a Cartesian sensor becomes a binary
    mask there; then the grid grows by the PML size on both sides of every
    axis, heterogeneous medium maps are extended by their edge values and
    masks and p0 by zeros.  Returns the expanded inputs, the options for
    them, the slices of the user's grid in the expanded one (the _all and
    _final outputs are cut back to it) and the Cartesian reorder index."""
    pml_size = opts.get("PMLSize", 20 if dim <= 2 else 10)
    sizes = [int(v) for v in np.broadcast_to(np.asarray(pml_size, dtype=np.int64), (dim,))]
    if isinstance(kgrid.Nt, str) or isinstance(kgrid.dt, str):
        if sensor is not None and _has(sensor, "time_reversal_boundary_data"):
            raise ValueError("kgrid.t_array must be explicitly defined for time reversal")
        kgrid.makeTime(np.asarray(_get(medium, "sound_speed"), dtype=np.float64), 0.3)
    args = []
    for n, d, s in zip(kgrid.shape, kgrid.spacings, sizes):
        args += [n + 2 * s, d]
    kg = kWaveGrid(*args)
    kg.setTime(kgrid.Nt, kgrid.dt)
    opts = dict(opts)
    opts["PMLInside"] = True
    sm = np.atleast_1d(np.asarray(opts.get("Smooth", [True, False, False]), dtype=bool))
    med = dict(medium)
    for nm in ("sound_speed", "density", "alpha_coeff", "BonA"):
        if nm in med and np.asarray(med[nm]).size > 1:
            med[nm] = _expand(np.asarray(med[nm], dtype=np.float64), sizes, True)
    src = None
    if source is not None:
        src = dict(source)
        if "p0" in src:
            p0 = np.asarray(src["p0"], dtype=np.float64)
            if p0.shape != kgrid.shape:
                raise ValueError("source.p0 must be the same size as the computational grid.")
            if sm[0]:
                p0 = smooth(p0, True)  # before the expansion
            src["p0"] = _expand(p0, sizes, False)
        for nm in ("p_mask", "u_mask"):
            if nm in src:
                src[nm] = _expand(src[nm], sizes, False)
    # p0 is already smoothed; c and rho are smoothed on the expanded grid
    sm_new = np.array(sm)
    sm_new[0] = False
    opts["Smooth"] = list(sm_new) if sm_new.size > 1 else [False]
    if sm.size == 1 and bool(sm[0]):
        opts["Smooth"] = [False, True, True]
    sens = sensor
    reorder_index = None
    if sensor is not None and _has(sensor, "mask"):
        sens = dict(sensor)
        mask = np.asarray(sens["mask"])
        if mask.shape != kgrid.shape and mask.ndim == 2 and mask.shape[0] == 2 * dim:
            # cuboid corners, checked on the user's grid and shifted by the
            # PML (kspaceFirstOrder_expandGridMatrices.m:99-112)
            cuboid_points(kgrid, mask)
            cor = np.asarray(mask, dtype=np.float64).copy()
            for a in range(dim):
                cor[[a, a + dim]] += sizes[a]
            sens["mask"] = cor
            inside = tuple(slice(s, s + n) for s, n in zip(sizes, kgrid.shape))
            return kg, med, src, sens, opts, inside, None
        if mask.shape != kgrid.shape and mask.ndim == 2 and mask.shape[0] == dim:
            cart = mask
            mask, order_index, reorder_index = cart2grid(kgrid, mask)
            if opts.get("CartInterp", "linear") == "linear" and "time_reversal_boundary_data" not in sens:
                # the points keep their coordinates, which the grown grid
                # (centred as the user's) interpolates: the triangulation is
                # computed after kspaceFirstOrder_expandGridMatrices.m
                inside = tuple(slice(s, s + n) for s, n in zip(sizes, kgrid.shape))
                sens["mask"] = cart
                return kg, med, src, sens, opts, inside, None
            if "time_reversal_boundary_data" in sens:
                d = np.asarray(sens["time_reversal_boundary_data"], dtype=np.float64)
                sens["time_reversal_boundary_data"] = d[np.argsort(order_index, kind="stable")]
        sens["mask"] = _expand(mask, sizes, False)
    inside = tuple(slice(s, s + n) for s, n in zip(sizes, kgrid.shape))
    return kg, med, src, sens, opts, inside, reorder_index


def prepare(kgrid: kWaveGrid, medium, source, sensor, dim: int, opts: dict) -> PreparedProblem:
    """Host precompute of kspaceFirstOrder{2D,3D} up to the native-offload
    variable set (kspaceFirstOrder_inputChecking.m, kspaceFirstOrder3D.m:585-
    607, kspaceFirstOrder_scaleSourceTerms.m)."""
    if kgrid.dim != dim:
        raise ValueError(f"kgrid must be {dim}D for kspaceFirstOrder{dim}D")
    if isinstance(opts.get("PMLSize"), str):
        # kspaceFirstOrder_checkOptionalInputs.m:297-305,484-512
        if opts["PMLSize"] != "auto":
            raise ValueError(f"Optional input 'PMLSize' must be a 1 or {dim} element numerical array, "
                             "or set to 'auto'.")
        if bool(opts.get("PMLInside", True)):
            raise ValueError("'PMLSize' set to 'auto' is only supported with 'PMLInside' set to false.")
        opts = dict(opts, PMLSize=optimal_pml_size(kgrid.shape, opts.get("PMLRange", [10, 40])))
    if not bool(opts.get("PMLInside", True)):
        if medium is None or not _has(medium, "sound_speed"):
            raise ValueError("The field medium.sound_speed must be defined.")
        kg, med, src, sens, o, inside, reorder_index = _pml_outside(
            kgrid, medium, source, sensor, dim, opts)
        pp = prepare(kg, med, src, sens, dim, o)
        pp.info.inside = inside
        if reorder_index is not None:
            pp.info.reorder_index = reorder_index
        return pp
    pp = PreparedProblem()
    pr = pp.prob
    shape = kgrid.shape
    ncell = kgrid.total_grid_points

    _check_inputs(dim, medium, source, sensor)

    # -- medium (inputChecking.m:148-281) -----------------------------------
    if medium is None or not _has(medium, "sound_speed"):
        raise ValueError("The field medium.sound_speed must be defined.")
    c = np.asarray(_get(medium, "sound_speed"), dtype=np.float64)
    if _has(medium, "density"):
        rho = np.asarray(_get(medium, "density"), dtype=np.float64)
    elif c.size == 1:
        rho = np.asarray(1.0)
    else:
        raise ValueError("The field medium.density must be defined.")
    for nm, v in (("sound_speed", c), ("density", rho)):
        if v.size != 1 and v.shape != shape:
            raise ValueError(f"medium.{nm} must be scalar or the same size as the grid")
    absorbing = _has(medium, "alpha_coeff") or _has(medium, "alpha_power")
    alpha_mode = _get(medium, "alpha_mode")
    if alpha_mode is not None and alpha_mode not in ("no_absorption", "no_dispersion", "stokes"):
        raise ValueError("medium.alpha_mode must be set to 'no_absorption', 'no_dispersion', or 'stokes'.")
    if _has(medium, "alpha_filter"):
        # kspaceFirstOrder_inputChecking.m:198-201,240-241
        if alpha_mode == "stokes":
            raise ValueError("Input option medium.alpha_filter is not supported with the "
                             "axisymmetric code or medium.alpha_mode = 'stokes'.")
        if np.asarray(_get(medium, "alpha_filter")).shape != shape:
            raise ValueError("medium.alpha_filter must be the same size as the computational grid.")
    nonlinear = _has(medium, "BonA")
    if nonlinear and not _has(medium, "density"):
        raise ValueError("medium.density must be defined for nonlinear simulations")

    cref_opt = _get(medium, "sound_speed_ref")
    if cref_opt is None or (isinstance(cref_opt, str) and cref_opt == "max"):
        c_ref = float(c.max())
    elif isinstance(cref_opt, str) and cref_opt == "min":
        c_ref = float(c.min())
    elif isinstance(cref_opt, str) and cref_opt == "mean":
        c_ref = float(c.mean())
    elif isinstance(cref_opt, str):
        raise ValueError("Unknown input for medium.sound_speed_ref.")
    else:
        c_ref = float(cref_opt)

    # -- time (inputChecking.m:1227-1281) -----------------------------------
    if isinstance(kgrid.Nt, str) or isinstance(kgrid.dt, str):
        if sensor is not None and _has(sensor, "time_reversal_boundary_data"):
            raise ValueError("kgrid.t_array must be explicitly defined for time reversal")
        kgrid.makeTime(c, 0.3)
    Nt, dt = int(kgrid.Nt), float(kgrid.dt)
    # kspaceFirstOrder_inputChecking.m:1270-1281: a warning, not an error
    if dt > check_stability(kgrid, c, c_ref, medium):
        warnings.warn("WARNING: time step may be too large for a stable simulation.",
                      RuntimeWarning, stacklevel=3)

    # -- options --------------------------------------------------------------
    pml_size = opts.get("PMLSize", 20 if dim <= 2 else 10)
    pml_alpha = opts.get("PMLAlpha", 2.0)
    pml_size = np.broadcast_to(np.asarray(pml_size, dtype=np.int64), (dim,))
    pml_alpha = np.broadcast_to(np.asarray(pml_alpha, dtype=np.float64), (dim,))
    sm = np.atleast_1d(np.asarray(opts.get("Smooth", [True, False, False]), dtype=bool))
    smooth_p0, smooth_c0, smooth_rho0 = bool(sm[0]), bool(sm[(len(sm) + 1) // 2 - 1]), bool(sm[-1])
    cart_interp = opts.get("CartInterp", "linear")
    data_cast = opts.get("DataCast", "off")
    if data_cast not in ("off", "double", "single", "gpuArray-single", "gpuArray-double"):
        raise ValueError(f"unknown DataCast '{data_cast}'")

    # -- sensor ---------------------------------------------------------------
    info = pp.info
    info.dim, info.shape, info.Nt = dim, shape, Nt
    info.record_given = False
    info.reorder_index = None
    info.interp = None
    info.cuboids = None
    info.freq_response = None
    info.dt = dt
    info.inside = None  # user's grid inside the expanded one ('PMLInside', false)
    info.time_rev = False
    info.cartesian = False
    info.use_sensor = sensor is not None and _has(sensor, "mask")
    record = ["p"]
    tr_data = None
    sens_idx = None
    n_sens = 0
    sens_full = False  # the mask is the whole grid (no index list is built)
    if info.use_sensor:
        info.time_rev = _has(sensor, "time_reversal_boundary_data")
        if _has(sensor, "frequency_response") and not info.time_rev:
            fr = np.asarray(_get(sensor, "frequency_response"), dtype=np.float64).ravel()
            if fr.size != 2:
                raise ValueError("sensor.frequency_response must be [center_freq, bandwidth]")
            info.freq_response = fr
        if _has(sensor, "record_mode"):
            # kspaceFirstOrder_inputChecking.m:360-363
            raise ValueError("sensor.record_mode input has been deprecated. Please use the syntax "
                             "sensor.record = {'p_rms', 'p_max', ...} to set recorded fields.")
        if _has(sensor, "record"):
            record = list(_get(sensor, "record"))
            info.record_given = True
            bad = [r for r in record if r not in RECORD_ALLOWED]
            if bad:
                raise NotImplementedError(f"sensor.record entries not supported by kwhip: {bad}")
        mask = np.asarray(_get(sensor, "mask"))
        if mask.shape == shape:
            if np.all(mask):
                sens_full, n_sens = True, ncell
            else:
                sens_idx = _find(mask)
                n_sens = len(sens_idx)
            if info.time_rev:
                tr_data = np.asarray(_get(sensor, "time_reversal_boundary_data"), dtype=np.float64)
        elif mask.ndim == 2 and mask.shape[0] == 2 * dim:
            # cuboid corners (kspaceFirstOrder_inputChecking.m:516-576)
            if info.time_rev:
                raise NotImplementedError("time reversal with a cuboid-corners sensor mask")
            if "u_split_field" in record:
                raise ValueError("The option sensor.record = {'u_split_field'} is only "
                                 "compatible with a binary sensor mask.")
            sens_idx, info.cuboids = cuboid_points(kgrid, mask)
            n_sens = len(sens_idx)
        elif mask.ndim == 2 and mask.shape[0] == dim:
            info.cartesian = True
            if "u_split_field" in record:
                # kspaceFirstOrder_inputChecking.m:1363-1366
                raise ValueError("The option sensor.record = {'u_split_field'} is only "
                                 "compatible with a binary sensor mask.")
            # the nearest-grid mask and its checks apply to 'linear' too
            # (kspaceFirstOrder_inputChecking.m:612-618)
            radial_ny = opts.get("_radial_ny")   # an axisymmetric problem's radial grid
            gmask, order_index, reorder_index = cart2grid(kgrid, mask, radial_ny)
            if cart_interp == "linear" and not info.time_rev:
                # one output row per Cartesian point, in the given order
                # (kspaceFirstOrder_createStorageVariables.m:36-46,63-67)
                info.interp = grid_data_linear(kgrid, mask, radial_ny)
                n_sens = mask.shape[1]
            else:
                sens_idx = _find(gmask)
                n_sens = len(sens_idx)
                info.reorder_index = reorder_index
            if info.time_rev:
                d = np.asarray(_get(sensor, "time_reversal_boundary_data"), dtype=np.float64)
                tr_data = d[np.argsort(order_index, kind="stable")]
        else:
            raise ValueError("sensor.mask must be a binary grid or a dim x N Cartesian array")
        if info.time_rev:
            record = []
            if tr_data.shape != (n_sens, Nt):
                raise ValueError("sensor.time_reversal_boundary_data must be [num_sensor_points x Nt]")
    info.record = record
    # what the device records: the intensities are computed afterwards from
    # p and the non-staggered velocity (kspaceFirstOrder_saveIntensity.m)
    dev_record = set(record) - {"I", "I_avg"}
    if "I" in record or "I_avg" in record:
        dev_record |= {"p", "u_non_staggered"}
    info.dev_record = dev_record
    rec_flags = 0
    for r in dev_record:
        rec_flags |= _abi.REC[r]
    record_start = int(_get(sensor, "record_start_index", 1)) if info.use_sensor else 1
    if info.use_sensor and not info.time_rev and record_start > Nt:
        raise ValueError("sensor.record_start_index must be less than or equal to kgrid.Nt")

    # -- sources (inputChecking.m:748-1014) -----------------------------------
    p0 = None
    if source is not None and _has(source, "p0"):
        p0 = np.asarray(_get(source, "p0"), dtype=np.float64)
        if p0.shape != shape:
            raise ValueError("source.p0 must be the same size as the computational grid.")
        if _has(source, "p"):
            raise ValueError("source.p0 and source.p can't be defined in the same simulation.")

    def time_source(names, mask_name, mode_name):
        present = [nm for nm in names if source is not None and _has(source, nm)]
        if not present:
            return None
        m = np.asarray(_get(source, mask_name))
        if m.shape != shape:
            raise ValueError(f"source.{mask_name} must be the same size as the computational grid.")
        pos = _find(m)
        if len(pos) == 0:
            raise ValueError(f"source.{mask_name} must be a binary grid with at least one element set to 1.")
        mode = _get(source, mode_name, "additive")
        if mode not in _abi.SRC_MODE:
            raise ValueError(f"source.{mode_name} must be set to 'additive', 'additive-no-correction', or 'dirichlet'.")
        # distinct values of the mask (inputChecking.m:850-887), from its
        # nonzero entries plus 0 when any cell is outside the source
        vals = m[np.unravel_index(pos, shape, order="F")]
        uniq_nz = np.unique(vals)
        n_uniq = uniq_nz.size + (1 if len(pos) < m.size else 0)
        labelled = not (n_uniq <= 2 and float(uniq_nz.sum()) == 1)
        sig_index = None
        if labelled:
            sig_index = np.asarray(vals).astype(np.uint64)
        sigs = {nm: np.atleast_2d(np.asarray(_get(source, nm), dtype=np.float64)).copy()
                for nm in present}
        for nm, s in sigs.items():
            if labelled and s.shape[0] != n_uniq - 1:
                raise ValueError("The number of time series must match the number of labelled source elements.")
            if not labelled and s.shape[0] > 1 and s.shape[0] != len(pos):
                raise ValueError("The number of time series must match the number of source elements.")
        return SimpleNamespace(pos=pos, mode=mode, sigs=sigs, sig_index=sig_index)

    src_p = time_source(["p"], "p_mask", "p_mode")
    u_names = ["ux", "uy", "uz"][:dim]
    src_u = time_source(u_names, "u_mask", "u_mode")
    # the k-space source correction at one frequency instead of the k-space
    # filter: signals times cos(2 pi f_ref dt/2), injected without
    # correction (kspaceFirstOrder_inputChecking.m:819-836,922-939,
    # kspaceFirstOrder_scaleSourceTerms.m:42-45,268-282)
    for sx, key in ((src_p, "p_frequency_ref"), (src_u, "u_frequency_ref")):
        if sx is None or not _has(source, key):
            continue
        f_ref = float(_get(source, key))
        if f_ref < 0:
            raise ValueError(f"source.{key} must be nonnegative")
        if f_ref > kgrid.k_max * float(np.min(c)) / (2 * math.pi):
            raise ValueError(f"source.{key} is higher than the maximum frequency supported by the spatial grid.")
        sx.mode = "additive-no-correction"
        w = math.cos(math.pi * f_ref * dt)
        sx.sigs = {nm: s * w for nm, s in sx.sigs.items()}

    # -- smoothing (inputChecking.m:1501-1600) ---------------------------------
    if p0 is not None and smooth_p0:
        p0 = smooth(p0, True)
    if smooth_c0 and c.size > 1:
        c = smooth(c, False)
    if smooth_rho0 and rho.size > 1:
        rho = smooth(rho, False)

    # -- source scaling (kspaceFirstOrder_scaleSourceTerms.m:43-391) ----------
    scale = bool(opts.get("ScaleSourceTerms", True))
    # sound speed at the source points (column-major linear indices)
    c_at = (lambda pos: c[np.unravel_index(pos, c.shape, order="F")]) if c.size > 1 else None
    if src_p is not None and scale:
        s = src_p.sigs["p"]
        if src_p.mode == "dirichlet":
            if c.size == 1:
                s = s / (dim * float(c) ** 2)
            else:
                for i in range(s.shape[0]):  # row i scaled by source point i (reference behaviour)
                    s[i] = s[i] / (dim * c_at(src_p.pos[i]) ** 2)
        else:
            if c.size == 1:
                s = s * (2 * dt / (dim * float(c) * kgrid.dx))
            else:
                for i in range(s.shape[0]):
                    s[i] = s[i] * (2 * dt / (dim * c_at(src_p.pos[i]) * kgrid.dx))
        src_p.sigs["p"] = s
    if src_u is not None and scale and src_u.mode != "dirichlet":
        for ax, nm in enumerate(u_names):
            if nm not in src_u.sigs:
                continue
            s = src_u.sigs[nm]
            dd = kgrid.spacings[ax]
            if c.size == 1:
                s = s * (2 * float(c) * dt / dd)
            else:
                for i in range(s.shape[0]):
                    s[i] = s[i] * (2 * c_at(src_u.pos[i]) * dt / dd)
            src_u.sigs[nm] = s

    # -- staggered density (kspaceFirstOrder3D.m:585-607) ----------------------
    use_sg = bool(opts.get("UseSG", True))
    rho_sg = [None] * 3
    if rho.size > 1 and use_sg:
        # computed on the column-major copy, so every array is contiguous and
        # its flattening free
        rho_f = _abi.fortran_flat(rho).reshape(shape, order="F")
        last = dim - 1
        for ax in range(dim):
            sg = np.empty(shape, dtype=np.float64, order="F")

            def part(z0, z1, ax=ax, sg=sg):
                # slab z0:z1 of the last axis; along the last axis the
                # neighbour plane may lie in the next slab (read only)
                a = rho_f[..., z0:z1]
                o = sg[..., z0:z1]
                o[...] = a
                if ax == last:
                    z1i = min(z1, shape[last] - 1)
                    if z1i > z0:
                        lo = rho_f[..., z0:z1i]
                        o[..., :z1i - z0] = lo + 0.5 * (rho_f[..., z0 + 1:z1i + 1] - lo)
                else:
                    sl_lo = [slice(None)] * dim
                    sl_hi = [slice(None)] * dim
                    sl_lo[ax] = slice(0, -1)
                    sl_hi[ax] = slice(1, None)
                    sl_lo, sl_hi = tuple(sl_lo), tuple(sl_hi)
                    o[sl_lo] = a[sl_lo] + 0.5 * (a[sl_hi] - a[sl_lo])

            _abi.by_slabs(part, shape[last])
            rho_sg[ax] = sg.ravel(order="F")

    # -- fill the kw_problem ---------------------------------------------------
    pr.dim = dim
    pr.Nx, pr.Ny, pr.Nz = kgrid.Nx, kgrid.Ny, (kgrid.Nz if dim == 3 else 1)
    pr.Nt = Nt
    pr.dx, pr.dy, pr.dz = kgrid.dx, kgrid.dy, (kgrid.dz if dim == 3 else 0.0)
    pr.dt = dt
    pr.c_ref = c_ref

    def grid_or_scalar(v):
        if v.size == 1:
            return None, float(v)
        return pp.hold(_abi.fortran_flat(v)), 0.0

    a, s_ = grid_or_scalar(c)
    pr.c0, pr.c0_scalar = _abi.dptr(a), s_
    a, s_ = grid_or_scalar(rho)
    pr.rho0, pr.rho0_scalar = _abi.dptr(a), s_
    for ax in range(3):
        pr.rho0_sg[ax] = _abi.dptr(pp.hold(rho_sg[ax]))
    if nonlinear:
        pr.nonlinear_flag = 1
        a, s_ = grid_or_scalar(np.asarray(_get(medium, "BonA"), dtype=np.float64))
        pr.BonA, pr.BonA_scalar = _abi.dptr(a), s_
    if absorbing:
        if alpha_mode == "stokes":
            pr.absorbing_flag = 2
            pr.alpha_power = 2.0
        else:
            pr.absorbing_flag = 1
            if not _has(medium, "alpha_coeff") or not _has(medium, "alpha_power"):
                raise ValueError("medium.alpha_coeff and medium.alpha_power must be defined")
            y = np.asarray(_get(medium, "alpha_power"), dtype=np.float64)
            if y.size != 1:
                raise ValueError("medium.alpha_power must be scalar.")
            pr.alpha_power = float(y)
            pr.alpha_mode = {"no_absorption": 1, "no_dispersion": 2}.get(alpha_mode, 0)
            if _has(medium, "alpha_filter"):
                af = np.asarray(_get(medium, "alpha_filter"), dtype=np.float64)
                pr.alpha_filter = _abi.dptr(pp.hold(_abi.fortran_flat(af)))
        a, s_ = grid_or_scalar(np.asarray(_get(medium, "alpha_coeff"), dtype=np.float64))
        pr.alpha_coeff, pr.alpha_coeff_scalar = _abi.dptr(a), s_
        if _has(medium, "alpha_sign"):
            sg = np.asarray(_get(medium, "alpha_sign"), dtype=np.float64).ravel()
            if sg.size != 2:
                raise ValueError("medium.alpha_sign must be given as a 2 element numerical array")
            # sign(alpha_sign) multiplies tau and eta, MATLAB's sign(0) = 0
            # included (kspaceFirstOrder_createAbsorptionVariables.m:78-81)
            pr.alpha_sign[0], pr.alpha_sign[1] = float(np.sign(sg[0])), float(np.sign(sg[1]))
            pr.alpha_sign_set = 1
    for ax in range(dim):
        pr.pml_size[ax] = int(pml_size[ax])
        pr.pml_alpha[ax] = float(pml_alpha[ax])
    pr.use_kspace = 1 if bool(opts.get("UsekSpace", True)) else 0
    pr.use_sg = 1 if use_sg else 0

    if p0 is not None:
        pr.p0 = _abi.dptr(pp.hold(_abi.fortran_flat(p0)))

    def fill_source(ts: _abi.TimeSource, src, nm):
        sig = src.sigs[nm]
        ts.flag = sig.shape[1]
        ts.mode = _abi.SRC_MODE[src.mode]
        idx = pp.hold(np.ascontiguousarray(src.pos.astype(np.uint64) + 1))
        ts.num_points = len(idx)
        ts.index = _abi.uptr(idx)
        ts.num_series = sig.shape[0]
        if src.sig_index is not None:
            ts.sig_index = _abi.uptr(pp.hold(np.ascontiguousarray(src.sig_index)))
        ts.input = _abi.dptr(pp.hold(np.ascontiguousarray(sig.ravel(order="F"))))

    if src_p is not None:
        fill_source(pr.p_source, src_p, "p")
    if src_u is not None:
        for ax, nm in enumerate(u_names):
            if nm in src_u.sigs:
                fill_source(pr.u_source[ax], src_u, nm)

    if info.use_sensor:
        full = info.interp is None and (
            sens_full or (n_sens == ncell and np.array_equal(sens_idx, np.arange(ncell))))
        pr.sensor_num_points = n_sens
        if info.interp is not None:
            tri, bc = info.interp
            pr.sensor_tri = _abi.uptr(pp.hold(np.ascontiguousarray(tri.ravel(order="F").astype(np.uint64))))
            pr.sensor_bc = _abi.dptr(pp.hold(np.ascontiguousarray(bc.ravel(order="F"))))
        elif info.cuboids is not None:
            # the corners themselves: the library enumerates their points
            cor = np.ascontiguousarray(info.cuboids.ravel(order="F").astype(np.uint64))
            pr.sensor_mask_corners = _abi.uptr(pp.hold(cor))
            pr.sensor_num_cuboids = info.cuboids.shape[1]
        elif not full:
            pr.sensor_mask_index = _abi.uptr(pp.hold(np.ascontiguousarray(sens_idx.astype(np.uint64) + 1)))
        if _has(sensor, "directivity_angle") and not info.time_rev:
            # kspaceFirstOrder_inputChecking.m:312-343 (time reversal ignores
            # it, :703-705)
            if dim != 2:
                raise ValueError("sensor.directivity_angle is only supported in 2D")
            ang = np.asarray(_get(sensor, "directivity_angle"), dtype=np.float64)
            if ang.shape != shape or info.interp is not None or info.reorder_index is not None \
                    or info.cuboids is not None:
                raise ValueError("sensor.directivity_angle and sensor.mask must be the same size.")
            pattern = _get(sensor, "directivity_pattern", "pressure")
            if pattern not in ("pressure", "gradient"):
                raise ValueError("Unsupported directivity pattern.")
            at = ang.ravel(order="F") if full else ang.ravel(order="F")[sens_idx]
            pr.directivity_angle = _abi.dptr(pp.hold(np.ascontiguousarray(at)))
            pr.directivity_pattern = 0 if pattern == "pressure" else 1
            pr.directivity_size = float(_get(sensor, "directivity_size", 10 * max(kgrid.dx, kgrid.dy)))
        pr.record_flags = rec_flags
        pr.record_start_index = record_start
        if info.time_rev:
            pr.time_reversal_boundary_data = _abi.dptr(pp.hold(np.ascontiguousarray(tr_data.ravel(order="F"))))
    info.n_sensor = n_sens if info.use_sensor else 0
    info.n_rec = (Nt - record_start + 1) if (info.use_sensor and not info.time_rev) else 0
    info.data_recast = bool(opts.get("DataRecast", False))
    info.backend = _abi.active_backend()
    return pp


# ---------------------------------------------------------------------------
# outputs
# ---------------------------------------------------------------------------

def alloc_outputs(pp: PreparedProblem):
    """Allocate the output arrays of a prepared problem and the kw_outputs
    struct that points to them."""
    info = pp.info
    dim, ncell, ns, nr = info.dim, int(np.prod(info.shape)), info.n_sensor, info.n_rec
    out = _abi.Outputs()
    arrays = {}
    rec = set(info.dev_record)
    if info.time_rev:
        rec = {"p_final"}

    def f32(n):
        return np.zeros(n, dtype=np.float32)

    def f64(n):
        return np.zeros(n, dtype=np.float64)

    if "p" in rec:
        arrays["p"] = f32(ns * nr)
        out.p = _abi.fptr(arrays["p"])
    for nm in ("p_max", "p_min", "p_rms"):
        if nm in rec:
            arrays[nm] = f64(ns)
            setattr(out, nm, _abi.dptr(arrays[nm]))
    for nm in ("p_max_all", "p_min_all"):
        if nm in rec:
            arrays[nm] = f64(ncell)
            setattr(out, nm, _abi.dptr(arrays[nm]))
    if "p_final" in rec:
        arrays["p_final"] = f32(ncell)
        out.p_final = _abi.fptr(arrays["p_final"])
    comps = ["ux", "uy", "uz"][:dim]
    for a, u in enumerate(comps):
        if "u" in rec:
            arrays[u] = f32(ns * nr)
            out.u[a] = _abi.fptr(arrays[u])
        for stat in ("max", "min", "rms"):
            if f"u_{stat}" in rec:
                arrays[f"{u}_{stat}"] = f64(ns)
                getattr(out, f"u_{stat}")[a] = _abi.dptr(arrays[f"{u}_{stat}"])
        if "u_non_staggered" in rec:
            arrays[f"{u}_non_staggered"] = f32(ns * nr)
            out.u_non_staggered[a] = _abi.fptr(arrays[f"{u}_non_staggered"])
        if "u_split_field" in rec:
            # kspaceFirstOrder_createStorageVariables.m:329-349
            for part in ("p", "s"):
                arrays[f"{u}_split_{part}"] = f32(ns * nr)
                getattr(out, f"u_split_{part}")[a] = _abi.fptr(arrays[f"{u}_split_{part}"])
        for stat in ("max_all", "min_all"):
            if f"u_{stat}" in rec:
                arrays[f"{u}_{stat}"] = f64(ncell)
                getattr(out, f"u_{stat}")[a] = _abi.dptr(arrays[f"{u}_{stat}"])
        if "u_final" in rec:
            arrays[f"{u}_final"] = f32(ncell)
            out.u_final[a] = _abi.fptr(arrays[f"{u}_final"])
    return out, arrays


def shape_outputs(pp: PreparedProblem, arrays: dict):
    """Reshape to the reference's output layout, reorder Cartesian sensor
    data (kspaceFirstOrder_reorderCartData.m) and select the return value
    (kspaceFirstOrder3D.m:1360-1383)."""
    info = pp.info
    ns, nr, shape = info.n_sensor, info.n_rec, info.shape
    sd = {}
    for k, v in arrays.items():
        if k in ("p", "ux", "uy", "uz") or k.endswith(("_non_staggered", "_split_p", "_split_s")):
            sd[k] = v.reshape((ns, nr), order="F")
        elif k.endswith("_all") or k.endswith("_final"):
            sd[k] = v.reshape(shape, order="F")
            if info.inside is not None:
                # kspaceFirstOrder3D.m:1306-1313, extractSensorData.m:464-569
                sd[k] = sd[k][info.inside]
        else:
            sd[k] = v
    if info.data_recast:
        sd = {k: np.asarray(v, dtype=np.float64) for k, v in sd.items()}
    if info.use_sensor and not info.time_rev and ("I" in info.record or "I_avg" in info.record):
        _save_intensity(sd, info)
    if info.reorder_index is not None and not info.time_rev:
        order = np.argsort(info.reorder_index, kind="stable")
        for k in list(sd):
            if not (k.endswith("_all") or k.endswith("_final")):
                sd[k] = sd[k][order]
    if not info.use_sensor:
        return None
    if info.freq_response is not None and "p" in sd:
        # kspaceFirstOrder3D.m:1346-1348
        sd["p"] = gaussian_filter(sd["p"], 1.0 / info.dt, info.freq_response[0],
                                  info.freq_response[1])
    if info.time_rev:
        return sd["p_final"]
    if info.cuboids is not None:
        # a record per cuboid, also without sensor.record
        # (kspaceFirstOrder2D.m:1109-1130)
        return _split_cuboids(sd, info.cuboids, info.dim)
    if not info.record_given:
        return sd["p"]
    return sd


def gaussian_filter(signal, Fs, freq, bandwidth):
    """+kwave/+legacy/gaussianFilter.m: every row of `signal` filtered in the
    frequency domain by the larger of two unit-height Gaussians at +freq and
    -freq whose full width at half maximum is bandwidth percent of freq
    (gaussian.m)."""
    signal = np.asarray(signal)
    N = signal.shape[1]
    f = (np.arange(N) - N // 2) * (Fs / N)
    sigma2 = (bandwidth / 100.0 * freq / (2.0 * math.sqrt(2.0 * math.log(2.0)))) ** 2
    gain = np.maximum(np.exp(-((f - freq) ** 2) / (2.0 * sigma2)),
                      np.exp(-((f + freq) ** 2) / (2.0 * sigma2)))
    # the gain is centred: move it to FFT order rather than the spectrum
    gain = np.fft.ifftshift(gain)
    out = np.real(np.fft.ifft(np.fft.fft(signal, axis=1) * gain[None, :], axis=1))
    return out.astype(signal.dtype, copy=False)


def fourier_shift(data, shift, axis=None):
    """Shift sampled data by `shift` samples with the Fourier shift theorem
    (+kwave/+legacy/fourierShift.m:73-114: by default along the last dimension, or
    the first for a column)."""
    data = np.asarray(data)
    if axis is None:
        axis = data.ndim - 1
        if data.ndim == 2 and data.shape[1] == 1:
            axis = 0
    N = data.shape[axis]
    if N % 2 == 0:
        k = 2 * np.pi * np.arange(-N // 2, N // 2) / N
    else:
        k = 2 * np.pi * np.arange(-(N - 1) // 2, (N - 1) // 2 + 1) / N
    k[N // 2] = 0.0
    ph = np.fft.ifftshift(np.exp(1j * k * shift))
    sh = [1] * data.ndim
    sh[axis] = N
    out = np.real(np.fft.ifft(ph.reshape(sh) * np.fft.fft(data, axis=axis), axis=axis))
    return out.astype(data.dtype, copy=False)


def _save_intensity(sd: dict, info):
    """Intensity from p and the non-staggered velocity, after the time loop
    (kspaceFirstOrder_saveIntensity.m:44-96): u is moved forward by half a time
    step with a Fourier shift, I = p u, I_avg = its mean over time; the
    velocity and p are dropped again when they were not asked for."""
    comps = ["x", "y", "z"][:info.dim]
    for c in comps:
        u_sgt = fourier_shift(sd[f"u{c}_non_staggered"], 0.5)
        ic = sd["p"] * u_sgt
        sd[f"I{c}"] = ic
        if "I_avg" in info.record:
            sd[f"I{c}_avg"] = np.mean(ic, axis=ic.ndim - 1)
    if "u_non_staggered" not in info.record:
        for c in comps:
            del sd[f"u{c}_non_staggered"]
    if "I" not in info.record:
        for c in comps:
            del sd[f"I{c}"]
    if "p" not in info.record:
        del sd["p"]


def run_prepared(pp: PreparedProblem):
    lib = _abi.load(pp.info.backend)
    out, arrays = alloc_outputs(pp)
    _abi.call(lib.kw_run, C.byref(pp.prob), C.byref(out))
    return shape_outputs(pp, arrays)


def _save_to_disk_options(dim, opts):
    """'SaveToDisk', 'SaveToDiskExit' and 'HDFCompressionLevel'
    (kspaceFirstOrder_checkOptionalInputs.m:125-135,376-414); returns the
    file name or None."""
    for k in ("SaveToDisk", "SaveToDiskExit"):
        if k in opts and dim == 1:
            raise ValueError("Optional input 'SaveToDisk' is not compatible with 1D simulations.")
    if "HDFCompressionLevel" in opts:
        lv = opts["HDFCompressionLevel"]
        if not np.isscalar(lv) or round(lv) != lv or lv > 9 or lv < 0:
            raise ValueError("Optional input 'HDFCompressionLevel' must be an integer between 0 and 9.")
    if "SaveToDiskExit" in opts and not isinstance(opts["SaveToDiskExit"], (bool, np.bool_)):
        raise ValueError("Optional input 'SaveToDiskExit' must be Boolean.")
    v = opts.get("SaveToDisk", False)
    if not isinstance(v, (bool, np.bool_, str)):
        raise ValueError("Optional input 'SaveToDisk' must be Boolean or a String.")
    if v is False or (isinstance(v, (bool, np.bool_)) and not v):
        return None
    return SAVE_TO_DISK_FILENAME_DEF if isinstance(v, (bool, np.bool_)) else v


def save_to_disk(pp: PreparedProblem, filename: str):
    """Write the prepared problem as the native engines' HDF5 input file
    (kspaceFirstOrder_saveToDisk.m: the variable set, names and sizes, in
    single precision with uint64 sizes, flags and indices), through
    kw_save_problem."""
    import os
    info = pp.info
    if info.dim == 3 and (not info.use_sensor or info.cartesian or info.time_rev):
        # kspaceFirstOrder_inputChecking.m:1393-1395
        raise ValueError("The optional input 'SaveToDisk' is currently only compatible with "
                         "forward simulations using a non-zero binary sensor mask.")
    if info.cartesian or not info.use_sensor:
        # kspaceFirstOrder_saveToDisk.m:43-46
        raise ValueError("Optional input 'SaveToDisk' only supported for sensor masks defined as "
                         "a binary matrix or the opposing corners of a rectangle (2D) or cuboid (3D).")
    ext = os.path.splitext(filename)[1]
    if ext == "":
        filename = filename + ".h5"
    elif ext == ".mat":
        raise NotImplementedError("'SaveToDisk' to a MATLAB .mat file is not supported by kwhip")
    elif ext != ".h5":
        raise ValueError("unknown file extension for 'SaveToDisk' filename")
    if os.path.exists(filename):
        os.remove(filename)
    lib = _abi.load(info.backend)
    _abi.call(lib.kw_save_problem, C.byref(pp.prob), filename.encode())
    return filename


def _check_stream_to_disk(dim, sensor, opts):
    """'StreamToDisk' (kspaceFirstOrder_checkOptionalInputs.m:353-374,
    kspaceFirstOrder_inputChecking.m:1369-1374).  The reference buffers the
    recorded pressure for that many steps in host memory, appends each full
    buffer to a temporary file and reads the whole record back at the end
    (kspaceFirstOrder3D.m:1050-1145,1235-1290), so its sensor_data.p is the
    record of a run without it.  Here the record stays in device memory for
    the whole run (one float per sensor point and step), and the option is
    checked and accepted."""
    if "StreamToDisk" not in opts:
        return
    v = opts["StreamToDisk"]
    if dim != 3:
        raise ValueError("Optional input 'StreamToDisk' is currently only compatible with 3D fluid "
                         "simulations.")
    if not (np.size(v) == 1 and (isinstance(v, (bool, np.bool_)) or np.isscalar(v)
                                 and isinstance(v, (int, float, np.number)))):
        raise ValueError("Optional input 'StreamToDisk' must be a single scalar or Boolean value.")
    if v is False or (not isinstance(v, (bool, np.bool_)) and float(v) == 0):
        return
    sen = None if sensor is None else (sensor if isinstance(sensor, dict) else vars(sensor))
    mask = None if sen is None else sen.get("mask")
    if (sen is None or mask is None or not np.any(np.asarray(mask))
            or "time_reversal_boundary_data" in sen):
        raise ValueError("The optional input 'StreamToDisk' is currently only compatible with "
                         "forward simulations using a non-zero sensor mask.")
    if "record" in sen and any(r != "p" for r in sen["record"]):
        raise ValueError("The optional input 'StreamToDisk' is currently only compatible with "
                         "sensor.record = {'p'} (the default).")


def _solve(dim, kgrid, medium, source, sensor, args, kwargs):
    opts = _parse_options(args, kwargs)
    if "RadialSymmetry" in opts:
        raise ValueError("Unknown optional input 'RadialSymmetry'.")
    _check_use_fd(dim, opts)
    _check_stream_to_disk(dim, sensor, opts)
    save = _save_to_disk_options(dim, opts)
    pp = prepare(kgrid, medium, source, sensor, dim, opts)
    if save is not None:
        save_to_disk(pp, save)
        if opts.get("SaveToDiskExit", True):
            # kspaceFirstOrder3D.m:681-684
            return []
    return run_prepared(pp)


def _column_to_1d(obj, names, n):
    """1-D inputs given as MATLAB column vectors (Nx x 1) become (Nx,)."""
    if obj is None:
        return obj
    out = dict(obj) if isinstance(obj, dict) else dict(vars(obj))
    for nm in names:
        if nm in out and out[nm] is not None:
            v = np.asarray(out[nm])
            if v.ndim == 2 and v.shape == (n, 1):
                out[nm] = v[:, 0]
    return out


def kspaceFirstOrder1D(kgrid, medium, source, sensor, *args, **kwargs):  # noqa: N802
    """1D time-domain simulation (reference +kwave/+legacy/kspaceFirstOrder1D.m).
    Fields are arrays of shape (Nx,) (Nx x 1 columns are accepted); a
    Cartesian sensor mask is a 1 x N array of x positions."""
    n = kgrid.Nx
    medium = _column_to_1d(medium, ("sound_speed", "density", "alpha_coeff", "BonA"), n)
    source = _column_to_1d(source, ("p0", "p_mask", "u_mask"), n)
    sensor = _column_to_1d(sensor, ("mask",), n)
    if sensor is not None and "u_split_field" in list(sensor.get("record", [])):
        raise ValueError("sensor.record 'u_split_field' needs a 2D or 3D simulation")
    return _solve(1, kgrid, medium, source, sensor, args, kwargs)


def kspaceFirstOrder2D(kgrid, medium, source, sensor, *args, **kwargs):  # noqa: N802
    """2D time-domain simulation (reference +kwave/+legacy/kspaceFirstOrder2D.m)."""
    return _solve(2, kgrid, medium, source, sensor, args, kwargs)


def kspaceFirstOrder3D(kgrid, medium, source, sensor, *args, **kwargs):  # noqa: N802
    """3D time-domain simulation (reference +kwave/+legacy/kspaceFirstOrder3D.m)."""
    return _solve(3, kgrid, medium, source, sensor, args, kwargs)


# ---------------------------------------------------------------------------
# kspaceFirstOrderAS
# ---------------------------------------------------------------------------

def _as_mirror(f, kind, signed, ws=False):
    """The layout of a radial-grid field on the expanded grid of the
    'WSWA-FFT' symmetry (kspaceFirstOrderAS.m:703-707, 944-952): p-like
    [f, 0, -flip(f(2:end)), -f, 0, flip(f(2:end))], y-staggered [f, flip(f),
    -f, -flip(f)]; unsigned (material grids) every copy keeps its sign and
    the mirror columns of p-like grids repeat their neighbour.  ws: the
    'WSWS-FFT' grid of 2 Ny - 2 points (:720-722, 966-977), p-like
    [f, flip(f(2:end-1))], y-staggered [f(1:end-1), -flip(f(1:end-1))]
    (unsigned: +flip)."""
    f = np.asarray(f, dtype=np.float64)
    nx, ny = f.shape
    if ws:
        if kind == "p":
            return np.concatenate([f, f[:, -2:0:-1]], axis=1)
        return np.concatenate([f[:, :-1], (-1.0 if signed else 1.0) * f[:, -2::-1]], axis=1)
    e = np.zeros((nx, 4 * ny))
    sg = (1, -1, -1, 1) if kind == "p" else (1, 1, -1, -1)
    if not signed:
        sg = (1, 1, 1, 1)
    for b in range(4):
        blk = e[:, b * ny:(b + 1) * ny]
        if kind == "p":
            if b % 2 == 0:
                blk[:] = sg[b] * f
            else:
                blk[:, 1:] = sg[b] * f[:, :0:-1]
                blk[:, 0] = 0.0 if signed else f[:, -1]
        else:
            blk[:] = sg[b] * (f if b % 2 == 0 else f[:, ::-1])
    return e


def _as_points(pos, nx, ny, ws=False):
    """Points of the radial grid (column-major linear indices) on the
    expanded grid of the p-like 'WSWA-FFT' layout: each point (i, j) at
    column j (sign +), 2 Ny - j (-), 2 Ny + j (-) and 4 Ny - j (+), the
    second and fourth copies skipping j = 0, whose mirror column stays 0
    ('WSWS-FFT': column j and, for 0 < j < Ny - 1, 2 Ny - 2 - j, both +).
    Returns the expanded indices in column-major order, the radial point
    each came from and its sign."""
    pos = np.asarray(pos, dtype=np.int64)
    ix, iy = pos % nx, pos // nx
    src = np.arange(pos.size)
    if ws:
        off = (iy >= 1) & (iy <= ny - 2)
        lin = np.concatenate([pos, ix[off] + nx * (2 * ny - 2 - iy[off])])
        row = np.concatenate([src, src[off]])
        order = np.argsort(lin, kind="stable")
        return lin[order], row[order], np.ones(lin.size)
    cols, rows, signs = [iy, 2 * ny + iy], [src, src], [np.ones(pos.size), -np.ones(pos.size)]
    off = iy >= 1
    cols += [2 * ny - iy[off], 4 * ny - iy[off]]
    rows += [src[off], src[off]]
    signs += [-np.ones(off.sum()), np.ones(off.sum())]
    lin = np.concatenate([ix[r] + nx * c for r, c in zip(rows, cols)])
    row = np.concatenate(rows)
    sign = np.concatenate(signs)
    order = np.argsort(lin, kind="stable")
    return lin[order], row[order], sign[order]


def _as_points_u(pos, nx, ny, ws=False):
    """Points of the radial grid's y-staggered layout on the 'WSWA-FFT'
    expanded grid ([f, flip(f), -f, -flip(f)], kspaceFirstOrderAS.m:
    944-948): each point (i, j) at columns j (+), 2 Ny - 1 - j (+),
    2 Ny + j (-) and 4 Ny - 1 - j (-) ('WSWS-FFT', [f(1:end-1),
    -flip(f(1:end-1))], :966-970: j (+) and 2 Ny - 3 - j (-) for j < Ny - 1).
    Same returns as _as_points."""
    pos = np.asarray(pos, dtype=np.int64)
    ix, iy = pos % nx, pos // nx
    src = np.arange(pos.size)
    if ws:
        keep = iy < ny - 1
        lin = np.concatenate([pos[keep], ix[keep] + nx * (2 * ny - 3 - iy[keep])])
        row = np.concatenate([src[keep], src[keep]])
        sign = np.concatenate([np.ones(keep.sum()), -np.ones(keep.sum())])
        order = np.argsort(lin, kind="stable")
        return lin[order], row[order], sign[order]
    cols = [iy, 2 * ny - 1 - iy, 2 * ny + iy, 4 * ny - 1 - iy]
    lin = np.concatenate([ix + nx * c for c in cols])
    row = np.tile(src, 4)
    sign = np.repeat([1.0, 1.0, -1.0, -1.0], pos.size)
    order = np.argsort(lin, kind="stable")
    return lin[order], row[order], sign[order]


def _as_signal_rows(m, sig, pos):
    """Row of the signal matrix per source point: one series, one per
    point, or a labelled mask (kspaceFirstOrder_inputChecking.m:850-887)."""
    vals = m[np.unravel_index(pos, m.shape, order="F")]
    uniq = np.unique(vals)
    n_uniq = uniq.size + (1 if pos.size < m.size else 0)
    if not (n_uniq <= 2 and float(uniq.sum()) == 1):
        if sig.shape[0] != n_uniq - 1:
            raise ValueError("The number of time series must match the number of labelled "
                             "source elements.")
        return vals.astype(np.int64) - 1
    if sig.shape[0] == 1:
        return np.zeros(pos.size, dtype=np.int64)
    if sig.shape[0] == pos.size:
        return np.arange(pos.size)
    raise ValueError("The number of time series must match the number of source elements.")


def _as_velocity_sources(src, med, kgrid, scale, ws=False):
    """source.ux (axial) and source.uy (radial) of an axisymmetric
    simulation on the expanded grid: {axis: (expanded indices, series,
    mode)}, one series per expanded point signed as the field's layout (ux
    p-like, uy y-staggered), scaled first on the radial grid as
    kspaceFirstOrder_scaleSourceTerms.m:284-350 scales them (2 c0 dt / dx,
    2 c0 dt / dy; row i with the sound speed at source point i)."""
    Nx, Ny = kgrid.Nx, kgrid.Ny
    if "uz" in src:
        raise ValueError("source.uz is not defined for axisymmetric simulations.")
    if "u_mask" not in src:
        raise ValueError("source.u_mask must be defined with source.ux or source.uy")
    m = np.asarray(src["u_mask"])
    if m.shape != (Nx, Ny):
        raise ValueError("source.u_mask must be the same size as the computational grid.")
    pos = _find(m)
    if pos.size == 0:
        raise ValueError("source.u_mask must be a binary grid with at least one element set to 1.")
    mode = src.get("u_mode", "additive")
    if mode not in _abi.SRC_MODE:
        raise ValueError("source.u_mode must be set to 'additive', 'additive-no-correction', or "
                         "'dirichlet'.")
    dt = float(kgrid.dt)
    w = 1.0
    if "u_frequency_ref" in src:
        f_ref = float(src["u_frequency_ref"])
        if f_ref < 0:
            raise ValueError("source.u_frequency_ref must be nonnegative")
        mode = "additive-no-correction"
        w = math.cos(math.pi * f_ref * dt)
    c = np.asarray(med["sound_speed"], dtype=np.float64)
    c_at = (lambda i: float(c)) if c.size == 1 else (
        lambda i: float(c.reshape(-1, order="F")[pos[i]]))
    out = {}
    for ax, nm, dd in ((0, "ux", kgrid.dx), (1, "uy", kgrid.dy)):
        if nm not in src:
            continue
        if nm == "uy" and ws and mode != "additive" and np.any(pos // Nx >= Ny - 2):
            # a point source must not set or add to the mirror point that
            # stands in for the last radial column
            raise NotImplementedError("radial velocity sources (source.uy) with 'WSWS' in the "
                                      "'dirichlet' and 'additive-no-correction' modes are supported "
                                      "by kwhip away from the last two radial columns")
        sig = w * np.atleast_2d(np.asarray(src[nm], dtype=np.float64))
        rows = _as_signal_rows(m, sig, pos)
        if scale and mode != "dirichlet":
            for i in range(sig.shape[0]):
                sig[i] *= 2 * c_at(i) * dt / dd
        if nm == "ux":
            epos, row, sign = _as_points(pos, Nx, Ny, ws)
        else:
            epos, row, sign = _as_points_u(pos, Nx, Ny, ws)
        out[ax] = (epos, sign[:, None] * sig[rows[row]], mode)
    return out


def _as_pressure_source(src, med, kgrid, scale, ws=False):
    """source.p of an axisymmetric simulation on the expanded grid: one
    series per expanded point, signed as the p-like layout, scaled first on
    the radial grid as kspaceFirstOrder_scaleSourceTerms.m:52-129 scales
    them for a 2-D grid (row i with the sound speed at source point i)."""
    Nx, Ny = kgrid.Nx, kgrid.Ny
    if "p_mask" not in src:
        raise ValueError("source.p_mask must be defined with source.p")
    m = np.asarray(src["p_mask"])
    if m.shape != (Nx, Ny):
        raise ValueError("source.p_mask must be the same size as the computational grid.")
    pos = _find(m)
    if pos.size == 0:
        raise ValueError("source.p_mask must be a binary grid with at least one element set to 1.")
    mode = src.get("p_mode", "additive")
    if mode not in _abi.SRC_MODE:
        raise ValueError("source.p_mode must be set to 'additive', 'additive-no-correction', or "
                         "'dirichlet'.")
    vals = m[np.unravel_index(pos, m.shape, order="F")]
    uniq = np.unique(vals)
    n_uniq = uniq.size + (1 if pos.size < m.size else 0)
    labelled = not (n_uniq <= 2 and float(uniq.sum()) == 1)
    sig = np.atleast_2d(np.asarray(src["p"], dtype=np.float64)).copy()
    if labelled:
        if sig.shape[0] != n_uniq - 1:
            raise ValueError("The number of time series must match the number of labelled "
                             "source elements.")
        sig_idx = vals.astype(np.int64) - 1
    elif sig.shape[0] == 1:
        sig_idx = np.zeros(pos.size, dtype=np.int64)
    elif sig.shape[0] == pos.size:
        sig_idx = np.arange(pos.size)
    else:
        raise ValueError("The number of time series must match the number of source elements.")
    dt = float(kgrid.dt)
    c = np.asarray(med["sound_speed"], dtype=np.float64)
    if "p_frequency_ref" in src:
        # kspaceFirstOrder_scaleSourceTerms.m:42-45: the correction at one
        # frequency, injected without the k-space filter
        f_ref = float(src["p_frequency_ref"])
        if f_ref < 0:
            raise ValueError("source.p_frequency_ref must be nonnegative")
        mode = "additive-no-correction"
        sig *= math.cos(math.pi * f_ref * dt)
    if scale:
        c_at = (lambda i: float(c)) if c.size == 1 else (
            lambda i: float(c.reshape(-1, order="F")[pos[i]]))
        for i in range(sig.shape[0]):
            if mode == "dirichlet":
                sig[i] /= 2 * c_at(i) ** 2
            else:
                sig[i] *= 2 * dt / (2 * c_at(i) * kgrid.dx)
    epos, row, sign = _as_points(pos, Nx, Ny, ws)
    M = 2 * Ny - 2 if ws else 4 * Ny
    emask = np.zeros(Nx * M, dtype=bool)
    emask[epos] = True
    emask = emask.reshape((Nx, M), order="F")
    return {"p_mask": emask, "p": sign[:, None] * sig[sig_idx[row]], "p_mode": mode}


def _as_pml_outside(kgrid, medium, source, sensor, sym, opts):
    """'PMLInside' false for the axisymmetric code: p0 smoothed on the
    user's grid first (kspaceFirstOrder_inputChecking.m:1501-1535), then
    the grid grows by the PML on both axial sides and on the outer radial
    side only (kspaceFirstOrder_expandGridMatrices.m:38-48,62-70): medium
    maps by their edge values, p0 and masks by zeros; the _final and _all
    outputs are cut back to the user's grid (y1_inside = 1,
    kspaceFirstOrder_inputChecking.m:1737-1749)."""
    pml = opts.get("PMLSize", 20)
    if isinstance(pml, str):
        if pml != "auto":
            raise ValueError("Optional input 'PMLSize' must be a 1 or 2 element numerical array, "
                             "or set to 'auto'.")
        # kspaceFirstOrder_checkOptionalInputs.m:484-512 with the symmetry
        pml = optimal_pml_size((kgrid.Nx, kgrid.Ny), opts.get("PMLRange", [10, 40]), sym[:4])
    px, py = [int(v) for v in np.broadcast_to(np.atleast_1d(pml), (2,))]
    Nx, Ny = kgrid.Nx, kgrid.Ny
    pads = [(px, px), (0, py)]
    med = dict(medium) if isinstance(medium, dict) else dict(vars(medium))
    for nm in ("sound_speed", "density", "alpha_coeff", "BonA"):
        if nm in med and np.size(med[nm]) > 1:
            med[nm] = np.pad(np.asarray(med[nm], dtype=np.float64), pads, mode="edge")
    src = dict(source) if isinstance(source, dict) else dict(vars(source))
    sm = np.atleast_1d(opts.get("Smooth", True))
    if sm.size > 1 and any(sm[1:]):
        raise NotImplementedError("smoothing c0 or rho0 is not supported by kwhip for "
                                  "axisymmetric simulations")
    if "p0" in src:
        p0 = np.asarray(src["p0"], dtype=np.float64)
        if p0.shape != (Nx, Ny):
            raise ValueError("source.p0 must be the same size as the computational grid.")
        if bool(sm[0]):
            p0 = smooth(_as_mirror(p0, "p", True, sym.startswith("WSWS")), True)[:, :Ny]
        src["p0"] = np.pad(p0, pads)
    for nm in ("p_mask", "u_mask"):
        if nm in src:
            src[nm] = np.pad(np.asarray(src[nm]), pads)
    sen = None
    if sensor is not None:
        sen = dict(sensor) if isinstance(sensor, dict) else dict(vars(sensor))
        if np.shape(sen["mask"]) == (Nx, Ny):
            sen["mask"] = np.pad(np.asarray(sen["mask"]), pads)
        # Cartesian points keep their coordinates: the grown grid keeps x
        # centred and y the radius from the axis
    kg = kWaveGrid(Nx + 2 * px, kgrid.dx, Ny + py, kgrid.dy)
    kg.setTime(kgrid.Nt, kgrid.dt)
    o = dict(opts, PMLInside=True, Smooth=False, PMLSize=[px, py], RadialSymmetry=sym)
    pp = _prepare_as(kg, med, src, sen, o)
    if pp is not None:
        # grid outputs come back on the user's part of the radial grid
        pp.info.inside = (slice(px, px + Nx), slice(0, Ny))
    return pp


def kspaceFirstOrderAS(kgrid, medium, source, sensor, *args, **kwargs):  # noqa: N802
    """Axisymmetric time-domain simulation (reference
    +kwave/+legacy/kspaceFirstOrderAS.m): a 2-D grid with x the axial and y
    the radial coordinate, the axis at the first y index.  Runs the 2-D
    kernels on the radially expanded grid of the 'WSWA-FFT' symmetry (the
    default; 'WSWA' gives the same field, the reference's discrete
    trigonometric transforms agreeing with it to machine precision) or of
    'WSWS-FFT' ('WSWS' likewise) with the u_r / r term of the divergence
    added by the library (kw_problem.axisymmetric 1 or 3).
    Supported here: source.p0 or a pressure
    source (every p_mode, laid out on the expanded grid with the signs of
    the p-like layout), velocity sources source.ux (axial) and source.uy
    (radial; with 'WSWS' its 'dirichlet' and 'additive-no-correction'
    modes away from the last two radial columns), lossless or Stokes media
    (with or without B/A),
    binary sensor masks or Cartesian points (x, r), time reversal;
    'SaveToDisk' writes the radial
    grid's input file (axisymmetric_flag 1, 'WSWA'), which kw_run_file and
    kspaceFirstOrder-HIP run; 'PMLInside' false grows the grid by the PML
    outside the axial ends and the outer radius."""
    opts = _parse_options(args, kwargs)
    pp = _prepare_as(kgrid, medium, source, sensor, opts)
    if pp is None:
        return []   # 'SaveToDisk' wrote the input file and exited (kspaceFirstOrderAS.m:587-596)
    return run_prepared(pp)


def _prepare_as(kgrid, medium, source, sensor, opts):
    """The axisymmetric problem prepared for kw_run (or the engine): the
    PreparedProblem on the expanded grid, or None when 'SaveToDisk' wrote
    the input file and exits."""
    opts = dict(opts)
    _check_stream_to_disk(2, sensor, opts)
    sym = opts.pop("RadialSymmetry", "WSWA-FFT")
    if sym not in ("WSWA", "WSWS", "WSWA-FFT", "WSWS-FFT"):
        raise ValueError("Optional input 'RadialSymmetry' must be set to 'WSWA', 'WSWS', "
                         "'WSWA-FFT', 'WSWS-FFT'.")
    save = _save_to_disk_options(2, opts)
    if save is not None and sym != "WSWA":
        # kspaceFirstOrder_inputChecking.m:1402-1411: the file holds 'WSWA'
        if sym.startswith("WSWS"):
            print("  WARNING: Optional input 'RadialSymmetry' changed to 'WSWA' for compatibility "
                  "with 'SaveToDisk'.")
        sym = "WSWA"
    ws = sym in ("WSWS", "WSWS-FFT")
    if getattr(kgrid, "dim", 2) != 2:
        raise ValueError("kspaceFirstOrderAS takes a 2D kWaveGrid (x axial, y radial).")
    if not opts.get("UseSG", True):
        raise ValueError("Optional input 'UseSG' is not supported for axisymmetric simulations.")
    if not opts.get("PMLInside", True):
        return _as_pml_outside(kgrid, medium, source, sensor, sym, opts)
    if isinstance(opts.get("PMLSize"), str):
        raise ValueError("'PMLSize' set to 'auto' is only supported with 'PMLInside' set to false.")
    med = dict(medium) if isinstance(medium, dict) else dict(vars(medium))
    src = dict(source) if isinstance(source, dict) else dict(vars(source))
    sen = None if sensor is None else (dict(sensor) if isinstance(sensor, dict) else dict(vars(sensor)))
    for k in ("alpha_mode", "alpha_filter"):
        if k in med and not (k == "alpha_mode" and med[k] == "stokes"):
            raise ValueError(f"Input option medium.{k} is not supported with the axisymmetric code "
                             "or medium.alpha_mode = 'stokes'.")
    if "alpha_coeff" in med:
        # kspaceFirstOrder_inputChecking.m:173-190: the axisymmetric code is
        # Stokes absorption, alpha_power = 2
        if "alpha_power" in med and float(np.asarray(med["alpha_power"])) != 2:
            print("  WARNING: the axisymmetric code and stokes absorption assume alpha_power = 2, "
                  "user value ignored.")
        med["alpha_mode"] = "stokes"
        med["alpha_power"] = 2.0
    for k in src:
        if k not in ("p0", "p", "p_mask", "p_mode", "p_frequency_ref", "u_mask", "u_mode",
                     "u_frequency_ref", "ux", "uy", "uz"):
            raise ValueError(f"'{k}' is not a valid field for source")
    time_rev = sen is not None and "time_reversal_boundary_data" in sen
    if sen is not None:
        for k in sen:
            if k not in ("mask", "frequency_response", "record_mode", "record",
                         "record_start_index", "time_reversal_boundary_data"):
                raise ValueError(f"'{k}' is not a valid field for sensor")
        nonstag = {"u_non_staggered", "I", "I_avg"} & set(sen.get("record", []))
        if nonstag:
            # kspaceFirstOrder_extractSensorData.m shifts u_r to the
            # non-staggered grid with a periodic FFT over the radial grid
            # alone, not over the symmetric extension these runs step
            raise NotImplementedError("axisymmetric records of the velocity on the non-staggered grid "
                                      f"are not supported by kwhip: {sorted(nonstag)}")
        m = np.asarray(sen.get("mask"))
        if m.shape != (kgrid.Nx, kgrid.Ny):
            if not (m.ndim == 2 and m.shape[0] == 2):
                raise ValueError("sensor.mask must be a binary grid or a 2 x N Cartesian array")
            # Cartesian points (x, r): the nearest radial grid points
            # (cart2grid.m with axisymmetric true); time reversal takes them,
            # its data sorted into the mask's order
            # (kspaceFirstOrder_inputChecking.m:612-633), and records either
            # those points or linear interpolation on the radial grid
            rk = kWaveGrid(kgrid.Nx, kgrid.dx, kgrid.Ny, kgrid.dy)
            gm, order_index, _ = cart2grid(rk, m, kgrid.Ny)
            if time_rev:
                d = np.asarray(sen["time_reversal_boundary_data"], dtype=np.float64)
                sen = dict(sen, mask=gm,
                           time_reversal_boundary_data=d[np.argsort(order_index, kind="stable")])
            else:
                opts["_radial_ny"] = kgrid.Ny
    Nx, Ny = kgrid.Nx, kgrid.Ny
    if ws and Ny < 3:
        raise ValueError("'WSWS' needs at least 3 radial grid points")
    M = 2 * Ny - 2 if ws else 4 * Ny   # expanded radial size
    sm = opts.pop("Smooth", True)
    esrc = {}
    p0 = None
    if "p0" in src:
        if "p" in src:
            raise ValueError("source.p0 and source.p can't be defined in the same simulation.")
        # p0: smoothed on the expanded grid (kspaceFirstOrder_inputChecking.m:
        # 1504-1535), trimmed back, then laid out again
        p0 = np.asarray(src["p0"], dtype=np.float64)
        if bool(np.atleast_1d(sm)[0]):
            p0 = smooth(_as_mirror(p0, "p", True, ws), True)[:, :Ny]
        esrc["p0"] = _as_mirror(p0, "p", True, ws)
    if "p" in src:
        esrc.update(_as_pressure_source(src, med, kgrid, opts.get("ScaleSourceTerms", True), ws))
    usrc = {}
    if any(k in src for k in ("ux", "uy", "uz")):
        usrc = _as_velocity_sources(src, med, kgrid, opts.get("ScaleSourceTerms", True), ws)
    if np.size(np.atleast_1d(sm)) > 1 and any(np.atleast_1d(sm)[1:]):
        raise NotImplementedError("smoothing c0 or rho0 is not supported by kwhip for "
                                  "axisymmetric simulations")
    if save is not None:
        # the radial grid as kspaceFirstOrderAS.m writes it with 'SaveToDisk'
        # (kspaceFirstOrder_saveToDisk.m, axisymmetric_flag 1): a 2-D problem
        # of the radial grid, the sources scaled as a 2-D grid scales them
        rsrc = {k: v for k, v in src.items() if k != "p0"}
        if p0 is not None:
            rsrc["p0"] = p0
        ropts = dict(opts, Smooth=False)
        ropts.pop("SaveToDisk", None)
        ropts.pop("SaveToDiskExit", None)
        ropts.pop("HDFCompressionLevel", None)
        rp = prepare(kgrid, med, rsrc, sen, 2, ropts)
        rp.prob.axisymmetric = 2
        save_to_disk(rp, save)
        if opts.get("SaveToDiskExit", True):
            return None
    kg = kWaveGrid(Nx, kgrid.dx, M, kgrid.dy)
    kg.setTime(kgrid.Nt, kgrid.dt)
    emed = {}
    for k, v in med.items():
        if k in ("sound_speed", "density", "BonA", "alpha_coeff") and np.size(v) > 1:
            emed[k] = _as_mirror(v, "p", False, ws)
        else:
            emed[k] = v
    if "sound_speed_ref" not in emed:
        emed["sound_speed_ref"] = float(np.max(np.asarray(med["sound_speed"])))
    esen = None
    if sen is not None:
        esen = dict(sen)
        mask = np.zeros(Nx * M, dtype=bool)
        if time_rev:
            # the boundary data are enforced on every copy, with its sign
            pos = _find(np.asarray(sen["mask"]))
            epos, row, sign = _as_points(pos, Nx, Ny, ws)
            data = np.atleast_2d(np.asarray(sen["time_reversal_boundary_data"], dtype=np.float64))
            if data.shape[0] != pos.size:
                raise ValueError("The number of time series in sensor.time_reversal_boundary_data "
                                 "must match the number of sensor points.")
            mask[epos] = True
            esen["time_reversal_boundary_data"] = sign[:, None] * data[row]
        mask = mask.reshape((Nx, M), order="F")
        if "_radial_ny" in opts:
            mask = np.asarray(sen["mask"], dtype=np.float64)   # Cartesian (x, r)
        elif not time_rev:
            mask[:, :Ny] = np.asarray(sen["mask"]).astype(bool)
        esen["mask"] = mask
    pml = opts.get("PMLSize", 20)
    pml = [int(v) for v in np.broadcast_to(np.atleast_1d(pml), (2,))]
    if pml[1] > Ny:
        raise ValueError("The size of the PML must be smaller than the size of the grid.")
    opts["PMLSize"] = pml
    opts["Smooth"] = False
    # the pressure signals come scaled with the radial grid's sound speed
    opts["ScaleSourceTerms"] = False
    pp = prepare(kg, emed, esrc, esen, 2, opts)
    pp.prob.axisymmetric = 3 if ws else 1
    for ax, (epos, sig, mode) in usrc.items():
        ts = pp.prob.u_source[ax]
        ts.flag = sig.shape[1]
        ts.mode = _abi.SRC_MODE[mode]
        idx = pp.hold(np.ascontiguousarray(epos.astype(np.uint64) + 1))
        ts.num_points = len(idx)
        ts.index = _abi.uptr(idx)
        ts.num_series = sig.shape[0]
        ts.input = _abi.dptr(pp.hold(np.ascontiguousarray(sig.ravel(order="F"))))
    rho = np.asarray(med["density"], dtype=np.float64)
    if rho.size > 1:
        # kspaceFirstOrderAS.m:440-456 on the radial grid, then laid out
        rsx = rho.copy()
        rsx[:-1, :] = (rho[:-1, :] + rho[1:, :]) / 2
        rsy = rho.copy()
        rsy[:, :-1] = (rho[:, :-1] + rho[:, 1:]) / 2
        ex = np.asfortranarray(_as_mirror(rsx, "p", False, ws)).reshape(-1, order="F")
        ey = np.asfortranarray(_as_mirror(rsy, "u", False, ws)).reshape(-1, order="F")
        ex, ey = np.ascontiguousarray(ex), np.ascontiguousarray(ey)
        pp.keep += [ex, ey]
        pp.prob.rho0_sg[0] = _abi.dptr(ex)
        pp.prob.rho0_sg[1] = _abi.dptr(ey)
    # grid outputs (_all, _final) come back on the radial grid
    pp.info.inside = (slice(None), slice(0, Ny))
    return pp
