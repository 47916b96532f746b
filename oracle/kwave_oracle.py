"""CPU restatement (NumPy + scipy.fft) of k-Wave's legacy first-order k-space
solver time loop, kspaceFirstOrder2D / kspaceFirstOrder3D.

TEST INFRASTRUCTURE ONLY: this module is the oracle.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import
it, as the checker or as the timed CPU baseline.  The product (``kwhip`` and
its HIP library) never imports, links or calls it.

The reference is MATLAB (+kwave/+legacy in ucl-bug/k-wave-ii); MATLAB is not
available, so this file restates the algorithm line by line.  Every function
cites the reference file:line it follows (paths relative to the reference
repository root; ``L/`` abbreviates ``+kwave/+legacy/`` and ``LP/`` abbreviates
``+kwave/+legacy/private/``).

Precision
---------
``data_cast='off'`` (or 'double') runs everything in float64: this is the
"FP64 run of the reference".  ``data_cast='single'`` emulates
``'DataCast','single'``: every operator is built in float64 and rounded once to
float32 (LP/kspaceFirstOrder_dataCast.m:31-268), and all loop arithmetic and
FFTs run in float32 (scipy.fft keeps complex64 for float32 input).

How each part is pinned (tests/test_oracle_*.py)
-------------------------------------------------
* k-vectors, D+/D- shift operators, kappa: spectral derivative of sinusoids
  with +-dx/2 staggering, even and odd N (+kwave/+tests/+unit/
  TestFourierCollocation.m:47-302), brute-force DFT on tiny grids.
* sinc, getPML, db2neper, Blackman window: textbook closed forms and special
  values (TestFourierCollocation.m:229-236, TestSplitFieldPML.m:74-107).
* homogeneous lossless loop (u/rho/p updates, p0 initial condition, kappa):
  exact periodic solution P(k,t) = P0(k) cos(c|k|t) (k-space stepping is
  exact for c = c_ref), the d'Alembert plane wave of TestAcousticSolver.m:
  769-806, 2D plane wave == 3D plane wave, axis-permutation symmetry.
* power-law absorption/dispersion: temporal decay rate and phase speed of
  single Fourier modes vs alpha0*w^y*c and the Kramers-Kronig dispersion
  relation (Treeby & Cox 2010, the model implemented by
  LP/kspaceFirstOrder_createAbsorptionVariables.m).
* heterogeneous c/rho: normal-incidence reflection coefficient
  (Z2-Z1)/(Z2+Z1) of a plane wave at a planar interface.
* velocity on the non-staggered grid and the intensities (LP/
  kspaceFirstOrder_extractSensorData.m:36-57, LP/kspaceFirstOrder_saveIntensity.m,
  L/fourierShift.m): the exact periodic standing wave, whose non-staggered
  velocity and intensity have closed forms; fourierShift against the shifted
  closed form of sampled sinusoids.
* nonlinearity (BonA + nonlinear mass conservation): second-harmonic growth
  of a plane wave vs the Fubini small-distance law beta*w*p0^2*x/(2 rho c^3).
* sources: additive vs additive-no-correction vs dirichlet bookkeeping,
  linearity, 'additive' == explicit source_kappa filtering.
* time reversal: ordering of the boundary enforcement, final assignment.
* sensor statistics: p_max/p_min/p_rms against the recorded time series.
* cart2grid 'nearest': rounding and duplicate handling on hand-made inputs.
* Cartesian 'linear' (grid_data_linear): affine fields are reproduced
  exactly at any point (true of every triangulation of the grid), the
  barycentric coordinates rebuild the point, points on grid nodes reduce to
  the node, and hand-made points fall in the declared simplex.  MATLAB's
  delaunayTriangulation picks one of the many Delaunay triangulations of a
  regular grid in an implementation-defined way; the oracle and the product
  use the declared one (see grid_data_linear), so fields that are not affine
  inside a cell can differ from a MATLAB run by the choice of diagonal.
* velocity_split: single-mode gradient fields are all compressional and
  rotational fields all shear; an irrotational run has zero shear.
* 1-D (kspaceFirstOrder1D): the periodic exact solution, and equality with
  y-invariant 2-D runs for linear problems.
* cuboid corners: per-cuboid records equal the full-grid record there.
* gaussian_filter: gains on sampled tones (closed form).
* directivity_filters: plane-wave closed forms, including the one-bin
  offset of the reference's fftshift for odd sizes.
* alpha_filter: mode selection on a periodic grid (linearity).
* p/u_frequency_ref: a coarse-step tone against the k-space corrected source.
* optimal_pml_size: known answers.
* axisymmetric (kspace_first_order_as, tests/test_as_oracle.py): an on-axis
  Gaussian pulse is the closed-form spherical wave; a planted sign error in
  the u_r / r term is caught; an on-axis Gaussian pressure source ('additive'
  and 'additive-no-correction') gives the field of the 3-D solver on a
  Cartesian grid, and dropping the k-space source correction is caught;
  axial and radial velocity sources likewise; 'dirichlet' holds p at its
  points to the series at every step; time reversal equals a 'dirichlet'
  run of the flipped data one step shifted; 'WSWS-FFT' follows the same
  spherical wave and catches planted layout parities; with the PML outside
  the spherical wave holds on the whole user grid.
"""

from __future__ import annotations

import math
import os
from types import SimpleNamespace

import numpy as np
import scipy.fft as sfft

FFT_WORKERS = int(os.environ.get("KW_ORACLE_WORKERS", os.cpu_count() or 1))


# ---------------------------------------------------------------------------
# Grid (L/kWaveGrid.m)
# ---------------------------------------------------------------------------

def make_dim(N: int, d: float) -> np.ndarray:
    """Wavenumber vector in centred order.  L/kWaveGrid.m:779-797 (makeDim)."""
    if N % 2 == 0:
        nx = np.arange(-N // 2, N // 2, dtype=np.float64) / N
    else:
        nx = np.arange(-(N - 1) // 2, (N - 1) // 2 + 1, dtype=np.float64) / N
    nx[N // 2] = 0.0  # force the middle value to zero (:791)
    return (2.0 * math.pi / d) * nx


class KGrid:
    """Uniform 1D/2D/3D grid.  L/kWaveGrid.m:342-436 (constructor) and
    :494-530 (x_vec, t_array), :612-673 (makeTime).  1-D fields are arrays
    of shape (Nx,)."""

    def __init__(self, Nx, dx, Ny=None, dy=None, Nz=None, dz=None):
        self.Nx, self.dx = int(Nx), float(dx)
        self.Ny, self.dy = (int(Ny), float(dy)) if Ny is not None else (1, 0.0)
        self.Nz, self.dz = (int(Nz), float(dz)) if Nz is not None else (1, 0.0)
        self.dim = 1 + (Ny is not None) + (Nz is not None)
        self.kx_vec = make_dim(self.Nx, self.dx)
        self.ky_vec = make_dim(self.Ny, self.dy) if self.dim >= 2 else np.zeros(1)
        self.kz_vec = make_dim(self.Nz, self.dz) if self.dim == 3 else np.zeros(1)
        self.Nt = None
        self.dt = None

    @property
    def shape(self):
        return ((self.Nx,), (self.Nx, self.Ny), (self.Nx, self.Ny, self.Nz))[self.dim - 1]

    @property
    def N(self):
        return int(np.prod(self.shape))

    @property
    def spacings(self):
        return (self.dx, self.dy, self.dz)[:self.dim]

    def k_vecs(self):
        return [self.kx_vec, self.ky_vec, self.kz_vec][:self.dim]

    def k(self) -> np.ndarray:
        """|k| in centred order, accumulated as in :399-403 / :417-422."""
        ks = self.k_vecs()
        acc = np.zeros(self.shape)
        for ax, kv in enumerate(ks):
            sh = [1] * self.dim
            sh[ax] = -1
            acc = acc + kv.reshape(sh) ** 2
        return np.sqrt(acc)

    @property
    def k_max(self):
        return min(np.max(np.abs(v)) for v in self.k_vecs())

    def vec(self, axis: int) -> np.ndarray:
        """x_vec / y_vec / z_vec (:494-506): Nx*dx*kx_vec*dx/(2 pi)."""
        N = self.shape[axis]
        d = self.spacings[axis]
        return (N * d) * self.k_vecs()[axis] * d / (2.0 * math.pi)

    def set_time(self, Nt: int, dt: float):
        self.Nt, self.dt = int(Nt), float(dt)

    @property
    def t_array(self):
        return np.arange(self.Nt) * self.dt

    def make_time(self, c, cfl: float = 0.3, t_end: float | None = None):
        """makeTime (:612-673)."""
        c = np.asarray(c, dtype=np.float64)
        c_max, c_min = float(c.max()), float(c.min())
        if t_end is None:
            sizes = [n * d for n, d in zip(self.shape, self.spacings)]
            t_end = math.sqrt(sum(s * s for s in sizes)) / c_min
        dt = cfl * min(self.spacings) / c_max
        Nt = math.floor(t_end / dt) + 1
        q = t_end / dt
        rem = t_end - math.trunc(t_end / dt) * dt
        if math.floor(q) != math.ceil(q) and rem == 0:
            Nt += 1
        self.Nt, self.dt = int(Nt), float(dt)
        return self.t_array, self.dt


# ---------------------------------------------------------------------------
# Small operators
# ---------------------------------------------------------------------------

def sinc(x):
    """L/private/sinc.m:40-41: sin(x + pi*z)/(x + z) + z with z = (x == 0)."""
    x = np.asarray(x, dtype=np.float64)
    z = (x == 0).astype(np.float64)
    return np.sin(x + math.pi * z) / (x + z) + z


def db2neper(alpha, y=1.0):
    """L/db2neper.m: 100*alpha*(1e-6/(2 pi))^y/(20*log10(e))."""
    return 100.0 * np.asarray(alpha, dtype=np.float64) * (1e-6 / (2.0 * math.pi)) ** y / (
        20.0 * math.log10(math.e))


def get_pml(N, d, dt, c, pml_size, pml_alpha, staggered):
    """L/getPML.m:52-86 (quartic profile, exp(-alpha*dt/2))."""
    x = np.arange(1, pml_size + 1, dtype=np.float64)
    if staggered:
        left = pml_alpha * (c / d) * (((x + 0.5) - pml_size - 1) / (0 - pml_size)) ** 4
        right = pml_alpha * (c / d) * ((x + 0.5) / pml_size) ** 4
    else:
        left = pml_alpha * (c / d) * ((x - pml_size - 1) / (0 - pml_size)) ** 4
        right = pml_alpha * (c / d) * (x / pml_size) ** 4
    left = np.exp(-left * dt / 2.0)
    right = np.exp(-right * dt / 2.0)
    pml = np.ones(N)
    if pml_size > 0:
        pml[:pml_size] = left
        pml[N - pml_size:] = right
    return pml


def shift_derivative_ops(k_vec, d, use_sg=True):
    """kspaceFirstOrder3D.m:629-642: ifftshift(i k exp(+-i k d/2)) (the
    staggered-grid shift folded into the derivative), or ifftshift(i k)."""
    if use_sg:
        return (np.fft.ifftshift(1j * k_vec * np.exp(1j * k_vec * d / 2)),
                np.fft.ifftshift(1j * k_vec * np.exp(-1j * k_vec * d / 2)))
    return np.fft.ifftshift(1j * k_vec), np.fft.ifftshift(1j * k_vec)


def cosine_series(n, N, coeffs):
    """getWin.m:417-424 (cosineSeries)."""
    s = coeffs[0] * np.ones_like(n, dtype=np.float64)
    for i in range(1, len(coeffs)):
        s = s + (-1) ** i * coeffs[i] * np.cos(i * 2 * math.pi * n / (N - 1))
    return s


def blackman_1d(N, param=0.16):
    """getWin.m:230-231 Blackman (symmetric), a0=(1-a)/2, a1=1/2, a2=a/2."""
    n = np.arange(N, dtype=np.float64)
    return cosine_series(n, N, [(1 - param) / 2, 0.5, param / 2])


def get_win_rotated(grid_size, symmetric):
    """getWin.m:180-347 for type 'Blackman', 'Rotation' true (1-D: the
    window of :240-256 with no rotation).

    N is enlarged by one where the window is not symmetric (:188), the 1-D
    window of length max(N) is rotated radially with linear interpolation
    (interp1) and r clipped to the radius (:285-299 / :318-333), then trimmed
    (:302/:347)."""
    N = np.array(grid_size) + (~np.asarray(symmetric)).astype(int)
    L = int(N.max())
    win_lin = blackman_1d(L)
    if len(N) == 1:
        # 1-D: the window itself, trimmed (:240-256)
        return win_lin[:int(N[0]) - (0 if symmetric[0] else 1)]
    radius = (L - 1) / 2.0
    ll = np.linspace(-radius, radius, L)
    axes = [np.linspace(-radius, radius, int(n)) for n in N]
    grids = np.meshgrid(*axes, indexing="ij")
    r = np.sqrt(sum(g * g for g in grids))
    r[r > radius] = radius
    win = np.interp(r, ll, win_lin)
    sl = tuple(slice(0, int(n) - (0 if s else 1)) for n, s in zip(N, symmetric))
    return win[sl]


def smooth_mat(mat, restore_max=False):
    """L/smooth.m: Blackman window (rotated) applied in k-space, optionally
    restoring the maximum magnitude.  Window symmetry = rem(N, 2)."""
    mat = np.asarray(mat, dtype=np.float64)
    grid_size = mat.shape
    window_symmetry = [bool(n % 2) for n in grid_size]
    win = np.abs(get_win_rotated(grid_size, window_symmetry))
    mat_sm = np.real(sfft.ifftn(sfft.fftn(mat, workers=FFT_WORKERS) * np.fft.ifftshift(win),
                                workers=FFT_WORKERS))
    if restore_max:
        mat_sm = (np.max(np.abs(mat)) / np.max(np.abs(mat_sm))) * mat_sm
    return mat_sm


def matlab_round(x):
    """MATLAB round: half away from zero."""
    x = np.asarray(x, dtype=np.float64)
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def cart2grid(kgrid: KGrid, cart_data, axisymmetric=False):
    """L/cart2grid.m:59-194 (nearest-neighbour mapping; later points
    overwrite earlier ones at the same grid point; axisymmetric: y is the
    radius from the first grid row, :111-116).

    Returns (binary grid mask, order_index, reorder_index) with 1-based
    order/reorder values as in MATLAB."""
    cart = np.asarray(cart_data, dtype=np.float64)
    dim = kgrid.dim
    idx = []
    for ax in range(dim):
        d = kgrid.spacings[ax]
        N = kgrid.shape[ax]
        shift = 1 if (axisymmetric and ax == 1) else math.floor(N / 2) + 1
        v = matlab_round(cart[ax] / d) + shift
        idx.append(v.astype(np.int64))
    for ax in range(dim):
        if idx[ax].max() > kgrid.shape[ax] or idx[ax].min() < 1:
            raise ValueError("Cartesian points must lie within the grid defined by kgrid.")
    grid = np.zeros(kgrid.shape, dtype=np.int64)
    for p in range(cart.shape[1]):
        grid[tuple(int(idx[ax][p]) - 1 for ax in range(dim))] = p + 1
    flat = grid.ravel(order="F")
    reorder_index = flat[flat != 0]
    order_index = np.argsort(reorder_index, kind="stable") + 1
    mask = (grid != 0).astype(np.float64)
    return mask, order_index, reorder_index


def grid_data_linear(kgrid: KGrid, cart_data, axisymmetric=False):
    """Linear interpolation of grid values at Cartesian points
    (LP/kspaceFirstOrder_createStorageVariables.m:376-398 with
    LP/gridDataFast2D.m / gridDataFast3D.m: the enclosing simplex of a
    Delaunay triangulation of the grid points and the barycentric
    coordinates of the point in it; LP/kspaceFirstOrder_extractSensorData.m:
    296-... takes sum(f(tri) .* bc, 2)).

    Every Delaunay triangulation of a regular grid splits each cell into
    simplices of its own corners; the declared one here is the Kuhn
    subdivision.  With the cell's lower corner i and the point's local
    coordinates f in [0, 1]^dim sorted as f[s1] >= f[s2] >= ... (ties by
    axis order), the simplex is v0 = i, v_m = v_(m-1) + e_(s_m), with weights
    1 - f[s1], f[s1] - f[s2], ..., f[s_dim].  In 2-D this cuts every cell
    along its diagonal from (i, j) to (i+1, j+1).

    Grid point n (0-based) of axis a lies at (n - floor(N_a/2)) d_a (the
    axisymmetric grid, LP/kspaceFirstOrder_createStorageVariables.m:390-391:
    y at n d_y, the radius).
    Returns (tri, bc): [npts, dim+1] 0-based column-major linear indices and
    float64 weights."""
    cart = np.asarray(cart_data, dtype=np.float64)
    dim = kgrid.dim
    npts = cart.shape[1]
    corner = np.zeros((dim, npts), dtype=np.int64)
    frac = np.zeros((dim, npts))
    for a in range(dim):
        N = kgrid.shape[a]
        u = cart[a] / kgrid.spacings[a] + (0 if (axisymmetric and a == 1) else math.floor(N / 2))
        if np.any(u < -1e-9) or np.any(u > N - 1 + 1e-9):
            raise ValueError("Cartesian points must lie within the k-space grid defined by kgrid.")
        i0 = np.clip(np.floor(u), 0, max(N - 2, 0)).astype(np.int64)
        corner[a] = i0
        frac[a] = np.clip(u - i0, 0.0, 1.0)
    strides = np.cumprod([1] + list(kgrid.shape[:-1]))
    tri = np.zeros((npts, dim + 1), dtype=np.int64)
    bc = np.zeros((npts, dim + 1))
    for q in range(npts):
        order = sorted(range(dim), key=lambda a: (-frac[a, q], a))
        v = corner[:, q].copy()
        tri[q, 0] = int(np.dot(v, strides))
        prev = 1.0
        for m, a in enumerate(order):
            bc[q, m] = prev - frac[a, q]
            prev = frac[a, q]
            v[a] += 1
            tri[q, m + 1] = int(np.dot(v, strides))
        bc[q, dim] = prev
    return tri, bc


class _Interp:
    """Sensor values at Cartesian points by linear interpolation:
    sum(f(tri) .* bc, 2) in the run's precision, summed over the simplex
    vertices in order (LP/kspaceFirstOrder_extractSensorData.m:296-...)."""

    def __init__(self, tri, bc, rdt):
        self.tri = np.asarray(tri)
        self.bc = np.asarray(bc).astype(rdt)

    def get(self, a):
        flat = np.asarray(a).ravel(order="F")
        acc = flat[self.tri[:, 0]] * self.bc[:, 0]
        for k in range(1, self.tri.shape[1]):
            acc = acc + flat[self.tri[:, k]] * self.bc[:, k]
        return acc

    def __len__(self):
        return self.tri.shape[0]


def directivity_filters(kgrid: KGrid, unique_angles, pattern, size):
    """The k-space filter of each unique sensor angle
    (LP/directionalResponse.m with the wavenumbers of
    LP/kspaceFirstOrder_inputChecking.m:342): 'pressure': sinc(|k_t| size/2)
    with the tangential wavenumber k_t = cos(theta) ky - sin(theta) kx;
    'gradient': |k_n / k| with k_n = sin(theta) ky + cos(theta) kx (0 at
    k = 0).  Built on the centred wavenumber grids and moved to FFT order
    with fftshift, as the reference does (for an odd length this is one bin
    off ifftshift)."""
    kx = kgrid.kx_vec[:, None] * np.ones((1, kgrid.Ny))
    ky = np.ones((kgrid.Nx, 1)) * kgrid.ky_vec[None, :]
    k = np.sqrt(kx ** 2 + ky ** 2)
    out = []
    for th in unique_angles:
        if pattern == "pressure":
            kt = np.abs(math.cos(th) * ky - math.sin(th) * kx)
            d = sinc(kt * size / 2)
        elif pattern == "gradient":
            kn = math.sin(th) * ky + math.cos(th) * kx
            with np.errstate(invalid="ignore", divide="ignore"):
                d = np.abs(kn / k)
            d[k == 0] = 0.0
        else:
            raise ValueError("Unsupported directivity pattern.")
        out.append(np.fft.fftshift(d))
    return out


def gaussian_filter(signal, Fs, freq, bandwidth):
    """L/gaussianFilter.m: each row filtered in the frequency domain by the
    larger of two Gaussians of unit height centred on +freq and -freq, with
    FWHM bandwidth% of freq (variance (bw/100 freq / (2 sqrt(2 ln 2)))^2,
    L/gaussian.m)."""
    signal = np.asarray(signal)
    N = signal.shape[1]
    if N % 2 == 0:
        f = np.arange(-N // 2, N // 2) * Fs / N
    else:
        f = np.arange(-(N - 1) // 2, (N - 1) // 2 + 1) * Fs / N
    var = (bandwidth / 100 * freq / (2 * math.sqrt(2 * math.log(2)))) ** 2
    g = np.maximum(np.exp(-(f - freq) ** 2 / (2 * var)), np.exp(-(f + freq) ** 2 / (2 * var)))
    spec = np.fft.fftshift(np.fft.fft(signal, axis=1), axes=1)
    out = np.real(np.fft.ifft(np.fft.ifftshift(g[None, :] * spec, axes=1), axis=1))
    return out.astype(signal.dtype)


def cuboid_corners_index(kgrid: KGrid, corners):
    """Sensor points of a cuboid-corners mask, [x1, (y1, (z1,)) x2, (y2,
    (z2))] per column, 1-based and inclusive
    (LP/kspaceFirstOrder_inputChecking.m:519-576 checks,
    :1610-1639 indices): the column-major indices of each cuboid in turn,
    concatenated (overlapping cuboids repeat points).  0-based."""
    cor = np.asarray(corners, dtype=np.float64)
    dim = kgrid.dim
    if not np.all(np.mod(cor, 1) == 0):
        raise ValueError("sensor.mask cuboid corner indices must be integers.")
    cor = cor.astype(np.int64)
    if np.any(cor[dim:] - cor[:dim] < 0):
        raise ValueError("sensor.mask cuboid corners must be defined as [x1, ..., x2, ...].' "
                         "where x2 => x1, etc.")
    if np.any(cor < 1) or any(np.any(cor[[a, a + dim]] > kgrid.shape[a]) for a in range(dim)):
        raise ValueError("sensor.mask cuboid corners must be within the grid.")
    idx = []
    for q in range(cor.shape[1]):
        m = np.zeros(kgrid.shape, dtype=bool)
        m[tuple(slice(cor[a, q] - 1, cor[a + dim, q]) for a in range(dim))] = True
        idx.append(find_F(m))
    return np.concatenate(idx), cor


def reorder_cuboid_corners(sd, corners, dim):
    """LP/kspaceFirstOrder_reorderCuboidCorners.m: the sensor rows split per
    cuboid and reshaped (column-major) to the cuboid's size, time last; the
    whole-grid outputs go to the first cuboid.  Returns a list of dicts, one
    per cuboid (the reference's struct array)."""
    out = []
    start = 0
    for q in range(corners.shape[1]):
        size = tuple(int(corners[a + dim, q] - corners[a, q] + 1) for a in range(dim))
        n = int(np.prod(size))
        d = {}
        for k, v in sd.items():
            if k.endswith("_all") or k.endswith("_final"):
                continue
            v = np.asarray(v)
            part = v[start:start + n]
            d[k] = part.reshape(size + v.shape[1:], order="F")
        out.append(d)
        start += n
    for k, v in sd.items():
        if k.endswith("_all") or k.endswith("_final"):
            out[0][k] = v
    return out


def reorder_binary_sensor_data(data, reorder_index):
    """L/reorderBinarySensorData.m: sortrows on the appended reorder column."""
    order = np.argsort(np.asarray(reorder_index), kind="stable")
    return np.asarray(data)[order]


# ---------------------------------------------------------------------------
# Column-major helpers (MATLAB linear indexing)
# ---------------------------------------------------------------------------

def find_F(mask) -> np.ndarray:
    """find(mask ~= 0) as 0-based column-major linear indices."""
    return np.flatnonzero(np.asarray(mask).ravel(order="F") != 0)


class _Idx:
    """Column-major gather/scatter for a fixed index list."""

    def __init__(self, lin, shape):
        self.lin = np.asarray(lin, dtype=np.int64)
        self.sub = np.unravel_index(self.lin, shape, order="F")

    def get(self, a):
        return a[self.sub]

    def set(self, a, v):
        a[self.sub] = v

    def __len__(self):
        return len(self.lin)


def _check_fields(obj, allowed, what):
    """LP/checkFieldNames.m: every field of the structure must be allowed
    (lists of LP/kspaceFirstOrder_inputChecking.m:148-149,302-350,729-741)."""
    if obj is None:
        return
    for nm in (obj.keys() if isinstance(obj, dict) else vars(obj).keys()):
        if nm not in allowed:
            raise ValueError(f"{nm} is not a valid field for the structure {what}.")


def _along(v, ax, dim):
    sh = [1] * dim
    sh[ax] = -1
    return np.asarray(v).reshape(sh)


# ---------------------------------------------------------------------------
# The solver
# ---------------------------------------------------------------------------

RECORD_ALLOWED = {"p", "p_max", "p_min", "p_rms", "p_max_all", "p_min_all", "p_final",
                  "u", "u_max", "u_min", "u_rms", "u_max_all", "u_min_all", "u_final",
                  "u_non_staggered", "I", "I_avg", "u_split_field"}


def fourier_shift(data, shift, dim=None):
    """L/fourierShift.m:73-114: shift sampled data by `shift` samples along
    dimension `dim` (0-based here; default the last one, or the first for a
    column vector) with real(ifft(ifftshift(exp(i k shift)) .* fft(data)))."""
    data = np.asarray(data)
    if dim is None:
        dim = data.ndim - 1
        if data.ndim == 2 and data.shape[1] == 1:
            dim = 0
    N = data.shape[dim]
    if N % 2 == 0:
        k_vec = (2 * np.pi) * (np.arange(-N / 2, N / 2) / N)
    else:
        k_vec = (2 * np.pi) * (np.arange(-(N - 1) / 2, (N - 1) / 2 + 1) / N)
    k_vec[int(np.floor(N / 2))] = 0.0
    shape = [1] * data.ndim
    shape[dim] = N
    ph = np.fft.ifftshift(np.exp(1j * k_vec * shift)).reshape(shape)
    return np.real(np.fft.ifft(ph * np.fft.fft(data, axis=dim), axis=dim))


def save_intensity(sd, dim, record):
    """LP/kspaceFirstOrder_saveIntensity.m:44-96: I = p .* fourierShift(u
    on the non-staggered grid, 1/2) per component, I_avg its mean over the
    last dimension; u_non_staggered and p are removed when not requested."""
    comps = ["x", "y", "z"][:dim]
    for c in comps:
        u_sgt = fourier_shift(sd[f"u{c}_non_staggered"], 0.5)
        ic = (sd["p"] * u_sgt).astype(sd["p"].dtype)
        sd[f"I{c}"] = ic
        if "I_avg" in record:
            sd[f"I{c}_avg"] = np.mean(ic, axis=ic.ndim - 1).astype(ic.dtype)
    if "u_non_staggered" not in record:
        for c in comps:
            del sd[f"u{c}_non_staggered"]
    if "I" not in record:
        for c in comps:
            del sd[f"I{c}"]
    if "p" not in record:
        del sd["p"]
    return sd


def _get(obj, name, default=None):
    if obj is None:
        return default
    if isinstance(obj, dict):
        return obj.get(name, default)
    return getattr(obj, name, default)


def _has(obj, name):
    if obj is None:
        return False
    if isinstance(obj, dict):
        return name in obj and obj[name] is not None
    return getattr(obj, name, None) is not None


def velocity_split(us, kgrid, use_sg=True, rdt=np.float64, workers=None):
    """Compressional and shear parts of the staggered velocity components
    `us` (LP/kspaceFirstOrder_createStorageVariables.m:86-137 for the
    operators, LP/kspaceFirstOrder_extractSensorData.m:124-292): u_k,a =
    shift_a .* fftn(u_a) with shift_a = ifftshift(exp(-i k_a d_a/2)),
    k_norm_a = ifftshift(k_a / |k|) (0 at k = 0); part p of component a is
    real(ifftn(sum_b k_norm_a k_norm_b u_k,b)), part s is
    real(ifftn((1 - k_norm_a^2) u_k,a - sum_{b != a} k_norm_a k_norm_b u_k,b))."""
    dim = kgrid.dim
    shape = kgrid.shape
    cdt = np.complex64 if rdt == np.float32 else np.complex128
    kv = kgrid.k_vecs()
    kc = kgrid.k()
    k_norm = []
    for a in range(dim):
        ka = np.broadcast_to(_along(kv[a], a, dim), shape)
        with np.errstate(invalid="ignore", divide="ignore"):
            kn = ka / kc
        kn[kc == 0] = 0.0
        k_norm.append(np.fft.ifftshift(kn).astype(rdt))
    shift = [_along(np.fft.ifftshift(np.exp(-1j * kv[a] * kgrid.spacings[a] / 2)).astype(cdt)
                    if use_sg else np.ones(shape[a], dtype=cdt), a, dim) for a in range(dim)]
    uk = [shift[a] * sfft.fftn(np.asarray(us[a], dtype=rdt), workers=workers) for a in range(dim)]
    comp, shear = [], []
    for a in range(dim):
        cp = sum(k_norm[a] * k_norm[b] * uk[b] for b in range(dim))
        sh = sum(((1 - k_norm[a] ** 2) if b == a else -k_norm[a] * k_norm[b]) * uk[b]
                 for b in range(dim))
        comp.append(np.real(sfft.ifftn(cp, workers=workers)).astype(rdt))
        shear.append(np.real(sfft.ifftn(sh, workers=workers)).astype(rdt))
    return comp, shear


def expand_matrix(mat, sizes, edge_val=None):
    """L/expandMatrix.m: grow by sizes[d] elements on both sides of every
    axis d; without edge_val the new elements repeat the nearest edge value
    (corners take the corner value), with edge_val they hold it."""
    mat = np.asarray(mat)
    out_shape = tuple(n + 2 * s for n, s in zip(mat.shape, sizes))
    inner = tuple(slice(s, s + n) for s, n in zip(sizes, mat.shape))
    if edge_val is not None:
        out = np.full(out_shape, edge_val, dtype=mat.dtype)
        out[inner] = mat
        return out
    # each new element takes the value of the nearest element of mat
    idx = [np.clip(np.arange(m) - s, 0, n - 1) for m, s, n in zip(out_shape, sizes, mat.shape)]
    return mat[np.ix_(*idx)]


def optimal_pml_size(grid_size, pml_range=(10, 40)):
    """L/getOptimalPMLSize.m: per axis, the PML size in pml_range (inclusive)
    whose grown length N + 2 pml has the smallest largest prime factor (the
    first such size)."""
    # MATLAB round (halves away from zero), not Python's round-half-even
    lo, hi = (int(matlab_round(float(v))) for v in pml_range)
    if lo < 0 or hi < 0:
        raise ValueError("Optional input pml_range must be positive.")
    if hi <= lo:
        raise ValueError("The second value for pml_range must be greater than the first.")

    def largest_factor(n):
        best, f = 1, 2
        while f * f <= n:
            while n % f == 0:
                best, n = f, n // f
            f += 1
        return max(best, n) if n > 1 else best

    out = []
    for N in grid_size:
        facs = [largest_factor(N + 2 * p) for p in range(lo, hi + 1)]
        out.append(lo + int(np.argmin(facs)))
    return out


def _pml_outside(kgrid, medium, source, sensor, *, smooth, cart_interp, **kw):
    """'PMLInside', false: LP/kspaceFirstOrder_inputChecking.m:1501-1560 and
    LP/kspaceFirstOrder_expandGridMatrices.m.  The time axis is set and p0
    smoothed on the user's grid, a Cartesian sensor becomes a binary mask
    there; then the grid grows by the PML size on both sides of every axis
    (heterogeneous medium maps by edge replication, p0 and masks by zeros)
    and the simulation runs on it; the _final and _all grids are cut back to
    the user's grid (kspaceFirstOrder3D.m:1306-1313,
    LP/kspaceFirstOrder_extractSensorData.m:464-569) and Cartesian data
    reordered as usual."""
    dim = kgrid.dim
    pml_size = kw.get("pml_size")
    if pml_size is None:
        pml_size = 20 if dim <= 2 else 10
    sizes = [int(v) for v in np.broadcast_to(np.asarray(pml_size, dtype=np.int64), (dim,))]
    if kgrid.dt is None:
        kgrid.make_time(np.asarray(_get(medium, "sound_speed"), dtype=np.float64), 0.3)
    args = []
    for n, d, s in zip(kgrid.shape, kgrid.spacings, sizes):
        args += [n + 2 * s, d]
    kg = KGrid(*args)
    kg.set_time(kgrid.Nt, kgrid.dt)
    sm = np.atleast_1d(np.asarray(smooth, dtype=bool))
    smooth_p0, smooth_c0, smooth_rho0 = bool(sm[0]), bool(sm[(len(sm) + 1) // 2 - 1]), bool(sm[-1])
    med = {k: _get(medium, k) for k in ("sound_speed", "density", "alpha_coeff", "alpha_power",
                                        "alpha_mode", "alpha_sign", "BonA", "sound_speed_ref")
           if _has(medium, k)}
    for k in ("sound_speed", "density", "alpha_coeff", "BonA"):
        if k in med and np.asarray(med[k]).size > 1:
            med[k] = expand_matrix(np.asarray(med[k], dtype=np.float64), sizes)
    src = None
    if source is not None:
        src = dict(source) if isinstance(source, dict) else dict(vars(source))
        if "p0" in src:
            p0 = np.asarray(src["p0"], dtype=np.float64)
            if smooth_p0:
                p0 = smooth_mat(p0, True)
            src["p0"] = expand_matrix(p0, sizes, 0.0)
        for k in ("p_mask", "u_mask"):
            if k in src:
                src[k] = expand_matrix(np.asarray(src[k]), sizes, 0)
    sens = sensor
    reorder_index = None
    if sensor is not None and _has(sensor, "mask"):
        sens = dict(sensor) if isinstance(sensor, dict) else dict(vars(sensor))
        mask = np.asarray(sens["mask"])
        if mask.shape != kgrid.shape and mask.ndim == 2 and mask.shape[0] == 2 * dim:
            # cuboid corners: checked on the user's grid, then shifted by the
            # PML (LP/kspaceFirstOrder_expandGridMatrices.m:99-112)
            cuboid_corners_index(kgrid, mask)
            shifted = np.asarray(mask, dtype=np.float64).copy()
            for a in range(dim):
                shifted[[a, a + dim]] += sizes[a]
            sens["mask"] = shifted
            out = kspace_first_order(kg, med, src, sens, pml_inside=True,
                                     smooth=(False, smooth_c0, smooth_rho0), cart_interp=cart_interp, **kw)
            return _cut_back(out, sizes, kgrid)
        if mask.shape != kgrid.shape:
            time_rev = "time_reversal_boundary_data" in sens
            mask, order_index, reorder_index = cart2grid(kgrid, mask)
            if cart_interp == "linear" and not time_rev:
                # the points keep their coordinates, which the grown grid
                # (centred as the user's) interpolates
                # (LP/kspaceFirstOrder_createStorageVariables.m:376-398 runs
                # after LP/kspaceFirstOrder_expandGridMatrices.m)
                out = kspace_first_order(kg, med, src, sens, pml_inside=True,
                                         smooth=(False, smooth_c0, smooth_rho0),
                                         cart_interp=cart_interp, **kw)
                return _cut_back(out, sizes, kgrid)
            if time_rev:
                d = np.asarray(sens["time_reversal_boundary_data"], dtype=np.float64)
                sens["time_reversal_boundary_data"] = d[np.argsort(order_index, kind="stable")]
        sens["mask"] = expand_matrix(mask, sizes, 0)
    out = kspace_first_order(kg, med, src, sens, pml_inside=True,
                             smooth=(False, smooth_c0, smooth_rho0), cart_interp=cart_interp, **kw)
    time_rev = sensor is not None and _has(sensor, "time_reversal_boundary_data")
    return _cut_back(out, sizes, kgrid, reorder_index, time_rev)


def _cut_back(out, sizes, kgrid, reorder_index=None, time_rev=False):
    """The _final and _all grids of a run on the grown grid cut back to the
    user's grid; Cartesian sensor rows reordered."""
    inside = tuple(slice(s, s + n) for s, n in zip(sizes, kgrid.shape))
    if out is None:
        return None
    if isinstance(out, list):
        # cuboid corners: the grids sit in the first cuboid's record
        for k in list(out[0]):
            if k.endswith("_all") or k.endswith("_final"):
                out[0][k] = out[0][k][inside]
        return out
    if not isinstance(out, dict):
        if time_rev:
            return out[inside]  # p_final of a time reversal
        return reorder_binary_sensor_data(out, reorder_index) if reorder_index is not None else out
    for k in list(out):
        if k.endswith("_all") or k.endswith("_final"):
            out[k] = out[k][inside]
        elif reorder_index is not None:
            out[k] = reorder_binary_sensor_data(out[k], reorder_index)
    return out


def kspace_first_order(kgrid: KGrid, medium, source, sensor, *, data_cast="off",
                       pml_size=None, pml_alpha=None, pml_inside=True, smooth=(True, False, False),
                       pml_range=(10, 40),
                       cart_interp="linear", use_kspace=True, use_sg=True,
                       scale_source_terms=True, data_recast=False,
                       return_state=False, fft_workers=None, step_times=None):
    """kspaceFirstOrder2D.m:514-1139 / kspaceFirstOrder3D.m:551-1391.

    ``medium``/``source``/``sensor`` are dicts (or namespaces) with the
    reference's field names.  ``step_times`` (a list) receives the wall time
    of every time step (used for the CPU baseline).  Returns ``sensor_data`` selected as in
    kspaceFirstOrder3D.m:1360-1383 (plain p array when sensor.record is
    absent; p_final for time reversal; else a dict)."""
    W = FFT_WORKERS if fft_workers is None else fft_workers
    dim = kgrid.dim
    shape = kgrid.shape
    if isinstance(pml_size, str):
        # LP/kspaceFirstOrder_checkOptionalInputs.m:297-305,484-512
        if pml_size != "auto":
            raise ValueError("Optional input 'PMLSize' must be a 1 or dim element numerical "
                             "array, or set to 'auto'.")
        if pml_inside:
            raise ValueError("'PMLSize' set to 'auto' is only supported with 'PMLInside' set to false.")
        pml_size = optimal_pml_size(shape, pml_range)
    if not pml_inside:
        return _pml_outside(kgrid, medium, source, sensor, data_cast=data_cast,
                            pml_size=pml_size, pml_alpha=pml_alpha, smooth=smooth,
                            cart_interp=cart_interp, use_kspace=use_kspace, use_sg=use_sg,
                            scale_source_terms=scale_source_terms, data_recast=data_recast,
                            fft_workers=fft_workers)
    if data_cast in ("off", "double"):
        rdt, cdt = np.float64, np.complex128
    elif data_cast == "single":
        rdt, cdt = np.float32, np.complex64
    else:
        raise ValueError("data_cast must be 'off', 'double' or 'single'")

    def fftn(a):
        return sfft.fftn(a, workers=W)

    def ifftn_real(a):
        return np.real(sfft.ifftn(a, workers=W))

    _check_fields(medium, {"sound_speed", "sound_speed_ref", "density", "alpha_coeff",
                           "alpha_power", "alpha_mode", "alpha_filter", "alpha_sign", "BonA"},
                  "medium")
    _check_fields(source, {"p0", "p", "p_mask", "p_mode", "p_frequency_ref", "u_mask", "u_mode",
                           "u_frequency_ref"} | set(["ux", "uy", "uz"][:dim]), "source")
    _check_fields(sensor, {"mask", "time_reversal_boundary_data", "frequency_response",
                           "record_mode", "record", "record_start_index"}
                  | ({"directivity_pattern", "directivity_angle", "directivity_size"}
                     if dim == 2 else set()), "sensor")

    # -- medium (LP/kspaceFirstOrder_inputChecking.m:148-281) --------------
    c = np.asarray(_get(medium, "sound_speed"), dtype=np.float64)
    if _has(medium, "density"):
        rho = np.asarray(_get(medium, "density"), dtype=np.float64)
    else:
        if c.size != 1:
            raise ValueError("medium.density must be given for a heterogeneous sound speed")
        rho = np.asarray(1.0)
    absorbing = _has(medium, "alpha_coeff") or _has(medium, "alpha_power")
    alpha_mode = _get(medium, "alpha_mode")
    if absorbing:
        if alpha_mode == "stokes":
            eos = "stokes"
            alpha_power = 2.0
        else:
            eos = "absorbing"
            alpha_power = float(_get(medium, "alpha_power"))
            if not (0 <= alpha_power < 3):
                raise ValueError("medium.alpha_power must be a real number between 0 and 3.")
            if alpha_mode != "no_dispersion" and alpha_power == 1:
                raise ValueError("The power law dispersion term is not valid for alpha_power = 1.")
        alpha_coeff = np.asarray(_get(medium, "alpha_coeff"), dtype=np.float64)
    else:
        eos = "lossless"
    nonlinear = _has(medium, "BonA")
    BonA = np.asarray(_get(medium, "BonA"), dtype=np.float64) if nonlinear else None

    # reference sound speed (LP/kspaceFirstOrder_setSoundSpeedRef.m:30-50)
    cref_opt = _get(medium, "sound_speed_ref")
    if cref_opt is None or (isinstance(cref_opt, str) and cref_opt == "max"):
        c_ref = float(c.max())
    elif isinstance(cref_opt, str) and cref_opt == "min":
        c_ref = float(c.min())
    elif isinstance(cref_opt, str) and cref_opt == "mean":
        c_ref = float(c.mean())
    else:
        c_ref = float(cref_opt)

    # -- time (inputChecking.m:1227-1268) ----------------------------------
    if kgrid.dt is None:
        kgrid.make_time(c, 0.3)
    Nt, dt = kgrid.Nt, kgrid.dt

    # -- sensor (inputChecking.m:291-700) ------------------------------------
    use_sensor = sensor is not None and (_has(sensor, "mask") or isinstance(sensor, dict) and sensor)
    time_rev = use_sensor and _has(sensor, "time_reversal_boundary_data")
    record = set(["p"])
    record_given = use_sensor and _has(sensor, "record")
    if use_sensor and _has(sensor, "record_mode"):
        # LP/kspaceFirstOrder_inputChecking.m:360-363
        raise ValueError("sensor.record_mode input has been deprecated. Please use the syntax "
                         "sensor.record = {'p_rms', 'p_max', ...} to set recorded fields.")
    if record_given:
        record = set(_get(sensor, "record"))
        bad = record - RECORD_ALLOWED
        if bad:
            raise ValueError(f"unknown sensor.record entries: {sorted(bad)}")
    if time_rev:
        record = set()
    record_start_index = int(_get(sensor, "record_start_index", 1)) if use_sensor else 1
    if use_sensor and not time_rev and record_start_index > Nt:
        raise ValueError("sensor.record_start_index must be less than or equal to kgrid.Nt")
    reorder_index = None
    tr_data = None
    interp = None
    cuboids = None
    if use_sensor:
        mask = np.asarray(_get(sensor, "mask"))
        if mask.shape == shape:
            sensor_mask = mask
        elif mask.ndim == 2 and mask.shape[0] == 2 * dim:
            if time_rev:
                raise NotImplementedError("time reversal with a cuboid-corners sensor mask")
            if "u_split_field" in record:
                raise ValueError("The option sensor.record = {'u_split_field'} is only "
                                 "compatible with a binary sensor mask.")
            cub_idx, cuboids = cuboid_corners_index(kgrid, mask)
        elif mask.ndim == 2 and mask.shape[0] == dim:
            if "u_split_field" in record:
                # LP/kspaceFirstOrder_inputChecking.m:1363-1366
                raise ValueError("The option sensor.record = {'u_split_field'} is only "
                                 "compatible with a binary sensor mask.")
            # the nearest-grid mask is computed (and its checks applied) for
            # 'linear' too (LP/kspaceFirstOrder_inputChecking.m:612-618)
            sensor_mask, order_index, reorder_index = cart2grid(kgrid, mask)
            if cart_interp == "linear" and not time_rev:
                # LP/kspaceFirstOrder_createStorageVariables.m:36-46,376-398:
                # no reordering, one output row per Cartesian point
                interp = grid_data_linear(kgrid, mask)
                reorder_index = None
            if time_rev:
                d = np.asarray(_get(sensor, "time_reversal_boundary_data"), dtype=np.float64)
                tr_data = d[np.argsort(order_index, kind="stable")]
        else:
            raise ValueError("sensor.mask must be a binary grid or a dim x N Cartesian array")
        if time_rev and tr_data is None:
            tr_data = np.asarray(_get(sensor, "time_reversal_boundary_data"), dtype=np.float64)
        if interp is not None:
            sensor_idx = _Interp(*interp, rdt)
        elif cuboids is not None:
            sensor_idx = _Idx(cub_idx, shape)
        else:
            sensor_idx = _Idx(find_F(sensor_mask), shape)
    if time_rev:
        # fliplr (kspaceFirstOrder3D.m:721-728)
        tr_data = tr_data[:, ::-1]

    # -- sources (inputChecking.m:748-1014) ----------------------------------
    p0 = None
    if _has(source, "p0"):
        p0 = np.asarray(_get(source, "p0"), dtype=np.float64)
        if p0.shape != shape:
            raise ValueError("source.p0 must be the same size as the computational grid.")
        if _has(source, "p"):
            raise ValueError("source.p0 and source.p can't be defined in the same simulation.")

    def parse_time_source(names, mask_name, mode_name):
        sigs = {}
        for nm in names:
            if _has(source, nm):
                s = np.atleast_2d(np.asarray(_get(source, nm), dtype=np.float64)).copy()
                sigs[nm] = s
        if not sigs:
            return None
        m = np.asarray(_get(source, mask_name))
        if m.shape != shape:
            raise ValueError(f"source.{mask_name} must be the same size as the computational grid.")
        if np.sum(m != 0) == 0:
            raise ValueError(f"source.{mask_name} must have at least one element set to 1.")
        mode = _get(source, mode_name, "additive")
        if mode not in ("additive", "dirichlet", "additive-no-correction"):
            raise ValueError(f"source.{mode_name} must be 'additive', 'additive-no-correction' or 'dirichlet'")
        pos = find_F(m)
        uniq = np.unique(m)
        if uniq.size <= 2 and uniq.sum() == 1:
            labelled = False
            sig_index = None
            for s in sigs.values():
                if s.shape[0] > 1 and s.shape[0] != len(pos):
                    raise ValueError("number of time series must match the number of source elements")
        else:
            labelled = True
            sig_index = np.asarray(m).ravel(order="F")[pos].astype(np.int64) - 1
        return SimpleNamespace(sigs=sigs, pos=_Idx(pos, shape), sig_index=sig_index,
                               labelled=labelled, mode=mode,
                               flag={k: v.shape[1] for k, v in sigs.items()})

    src_p = parse_time_source(["p"], "p_mask", "p_mode")
    u_names = ["ux", "uy", "uz"][:dim]
    src_u = parse_time_source(u_names, "u_mask", "u_mode")
    # k-space source correction as a function of frequency
    # (LP/kspaceFirstOrder_inputChecking.m:819-836,922-939,
    # LP/kspaceFirstOrder_scaleSourceTerms.m:42-45,268-282): the signals
    # times cos(2 pi f_ref dt / 2), injected without the k-space correction
    for srcx, key in ((src_p, "p_frequency_ref"), (src_u, "u_frequency_ref")):
        if srcx is not None and _has(source, key):
            f_ref = float(_get(source, key))
            if f_ref < 0:
                raise ValueError(f"source.{key} must be nonnegative")
            if f_ref > kgrid.k_max * float(np.min(c)) / (2 * math.pi):
                raise ValueError(f"source.{key} is higher than the maximum frequency supported "
                                 "by the spatial grid.")
            srcx.mode = "additive-no-correction"
            for nm in srcx.sigs:
                srcx.sigs[nm] = srcx.sigs[nm] * math.cos(2 * math.pi * f_ref * dt / 2)

    # -- smoothing (inputChecking.m:1501-1543 for p0 with the maximum
    # restored, :1578-1600 for c and rho; 'Smooth' holds 1-3 flags assigned
    # as p0 = v(1), c = v(ceil((end+1)/2)), rho = v(end),
    # checkOptionalInputs.m:436-438).  c_ref and the time step were taken
    # from the unsmoothed sound speed above, as in the reference.
    sm = np.atleast_1d(np.asarray(smooth, dtype=bool))
    smooth_p0, smooth_c0, smooth_rho0 = bool(sm[0]), bool(sm[(len(sm) + 1) // 2 - 1]), bool(sm[-1])
    if p0 is not None and smooth_p0:
        p0 = smooth_mat(p0, True)
    if smooth_c0 and c.size > 1:
        c = smooth_mat(c)
    if smooth_rho0 and rho.size > 1:
        rho = smooth_mat(rho)

    # -- PML settings (setDefaults.m:151-170) --------------------------------
    if pml_size is None:
        pml_size = 20 if dim <= 2 else 10
    if pml_alpha is None:
        pml_alpha = 2.0
    pml_sizes = list(np.broadcast_to(np.asarray(pml_size), (dim,)).astype(int))
    pml_alphas = list(np.broadcast_to(np.asarray(pml_alpha, dtype=np.float64), (dim,)))

    # -- absorption (LP/kspaceFirstOrder_createAbsorptionVariables.m:29-100)
    k_c = kgrid.k()
    if eos == "absorbing":
        a_np = db2neper(alpha_coeff, alpha_power)
        with np.errstate(divide="ignore"):
            if alpha_mode != "no_absorption":
                nabla1 = k_c ** (alpha_power - 2)
                nabla1[np.isinf(nabla1)] = 0
                nabla1 = np.fft.ifftshift(nabla1)
                tau = -2.0 * a_np * c ** (alpha_power - 1)
            else:
                nabla1, tau = np.asarray(0.0), np.asarray(0.0)
            if alpha_mode != "no_dispersion":
                nabla2 = k_c ** (alpha_power - 1)
                nabla2[np.isinf(nabla2)] = 0
                nabla2 = np.fft.ifftshift(nabla2)
                eta = 2.0 * a_np * c ** alpha_power * math.tan(math.pi * alpha_power / 2)
            else:
                nabla2, eta = np.asarray(0.0), np.asarray(0.0)
        if _has(medium, "alpha_filter"):
            af = np.asarray(_get(medium, "alpha_filter"), dtype=np.float64)
            nabla1 = np.fft.ifftshift(np.fft.fftshift(nabla1) * af)
            nabla2 = np.fft.ifftshift(np.fft.fftshift(nabla2) * af)
        if _has(medium, "alpha_sign"):
            sgn = np.asarray(_get(medium, "alpha_sign"), dtype=np.float64)
            tau = np.sign(sgn[0]) * tau
            eta = np.sign(sgn[1]) * eta
    elif eos == "stokes":
        a_np = db2neper(alpha_coeff, 2)
        tau = -2.0 * a_np * c
        if _has(medium, "alpha_sign"):
            tau = np.sign(np.asarray(_get(medium, "alpha_sign"))[0]) * tau

    # -- pseudonyms (inputChecking.m:1682-1688) ------------------------------
    c0 = c
    rho0 = rho

    # -- source scaling (LP/kspaceFirstOrder_scaleSourceTerms.m:43-391) ------
    dx = kgrid.dx
    if src_p is not None and scale_source_terms:
        s = src_p.sigs["p"]
        if src_p.mode == "dirichlet":
            if c0.size == 1:
                s = s / (dim * c0 ** 2)
            else:
                for i in range(s.shape[0]):   # quirk: row i uses source point i
                    s[i] = s[i] / (dim * c0.ravel(order="F")[src_p.pos.lin[i]] ** 2)
        else:
            if c0.size == 1:
                s = s * (2 * dt / (dim * c0 * dx))
            else:
                for i in range(s.shape[0]):
                    s[i] = s[i] * (2 * dt / (dim * c0.ravel(order="F")[src_p.pos.lin[i]] * dx))
        src_p.sigs["p"] = s
    if src_u is not None and scale_source_terms and src_u.mode != "dirichlet":
        for ax, nm in enumerate(u_names):
            if nm not in src_u.sigs:
                continue
            s = src_u.sigs[nm]
            d = kgrid.spacings[ax]
            if c0.size == 1:
                s = s * (2 * c0 * dt / d)
            else:
                for i in range(s.shape[0]):
                    s[i] = s[i] * (2 * c0.ravel(order="F")[src_u.pos.lin[i]] * dt / d)
            src_u.sigs[nm] = s

    # -- staggered density (kspaceFirstOrder3D.m:585-607) --------------------
    rho0_sg = []
    for ax in range(dim):
        if rho0.ndim == dim and use_sg:
            r = np.moveaxis(rho0, ax, 0)
            sg = r.copy()
            sg[:-1] = r[:-1] + 0.5 * (r[1:] - r[:-1])  # interpn '*linear' at x+dx/2
            rho0_sg.append(np.moveaxis(sg, 0, ax))   # last plane: NaN -> original
        else:
            rho0_sg.append(rho0)
    rho0_sg_inv = [1.0 / r for r in rho0_sg]

    # -- PML and derivative operators (kspaceFirstOrder3D.m:619-661) ---------
    pml = [get_pml(shape[a], kgrid.spacings[a], dt, c_ref, pml_sizes[a], pml_alphas[a], False)
           for a in range(dim)]
    pml_sg = [get_pml(shape[a], kgrid.spacings[a], dt, c_ref, pml_sizes[a], pml_alphas[a], use_sg)
              for a in range(dim)]
    kv = kgrid.k_vecs()
    ddk_pos, ddk_neg = [], []
    for a in range(dim):
        dp, dn = shift_derivative_ops(kv[a], kgrid.spacings[a], use_sg)
        ddk_pos.append(dp)
        ddk_neg.append(dn)
    additive = ((src_p is not None and src_p.mode == "additive") or
                (src_u is not None and src_u.mode == "additive"))
    if use_kspace:
        kappa = np.fft.ifftshift(sinc(c_ref * k_c * dt / 2))
        source_kappa = np.fft.ifftshift(np.cos(c_ref * k_c * dt / 2)) if additive else None
    else:
        kappa = np.asarray(1.0)
        source_kappa = np.asarray(1.0)

    # -- data cast (LP/kspaceFirstOrder_dataCast.m) --------------------------
    def R(a):
        return np.asarray(a, dtype=np.float64).astype(rdt)

    def C(a):
        return np.asarray(a).astype(cdt)

    dt_c = rdt(dt)
    kappa = R(kappa)
    c0 = R(c0)
    rho0 = R(rho0)
    ddk_pos = [_along(C(v), a, dim) for a, v in enumerate(ddk_pos)]
    ddk_neg = [_along(C(v), a, dim) for a, v in enumerate(ddk_neg)]
    pml = [_along(R(v), a, dim) for a, v in enumerate(pml)]
    pml_sg = [_along(R(v), a, dim) for a, v in enumerate(pml_sg)]
    rho0_sg_inv = [R(v) for v in rho0_sg_inv]
    if eos == "absorbing":
        nabla1, nabla2, eta, tau = R(nabla1), R(nabla2), R(eta), R(tau)
    elif eos == "stokes":
        tau = R(tau)
    if nonlinear:
        BonA = R(BonA)
    if p0 is not None:
        p0 = R(p0)
    if source_kappa is not None:
        # not in the cast list, but double .* single is computed in single
        # in MATLAB, which is the same as rounding the operator once
        source_kappa = R(source_kappa)
    if src_p is not None:
        src_p.sigs = {k: R(v) for k, v in src_p.sigs.items()}
    if src_u is not None:
        src_u.sigs = {k: R(v) for k, v in src_u.sigs.items()}

    # -- fields (kspaceFirstOrder3D.m:695-702) ------------------------------
    p = np.zeros(shape, dtype=rdt)
    rhos = [np.zeros(shape, dtype=rdt) for _ in range(dim)]
    us = [np.zeros(shape, dtype=rdt) for _ in range(dim)]
    p_k = np.zeros(shape, dtype=cdt)

    # -- storage (LP/kspaceFirstOrder_createStorageVariables.m:57-400) -------
    sd = {}
    n_rec = Nt - record_start_index + 1
    if use_sensor and not time_rev:
        ns = len(sensor_idx)
        intensity = bool(record & {"I", "I_avg"})
        # p and u_non_staggered are also recorded for the intensities
        # (LP/kspaceFirstOrder_createStorageVariables.m:148,309)
        if "p" in record or intensity:
            sd["p"] = np.zeros((ns, n_rec), dtype=rdt)
        for nm in ("p_max", "p_min", "p_rms"):
            if nm in record:
                sd[nm] = np.zeros(ns, dtype=rdt)
        ucomp = ["ux", "uy", "uz"][:dim]
        if "u" in record:
            for u in ucomp:
                sd[u] = np.zeros((ns, n_rec), dtype=rdt)
        if "u_non_staggered" in record or intensity:
            for u in ucomp:
                sd[u + "_non_staggered"] = np.zeros((ns, n_rec), dtype=rdt)
        if "u_split_field" in record:
            # LP/kspaceFirstOrder_createStorageVariables.m:329-349
            for u in ucomp:
                sd[u + "_split_p"] = np.zeros((ns, n_rec), dtype=rdt)
                sd[u + "_split_s"] = np.zeros((ns, n_rec), dtype=rdt)
        for stat in ("max", "min", "rms"):
            if "u_" + stat in record:
                for u in ucomp:
                    sd[f"{u}_{stat}"] = np.zeros(ns, dtype=rdt)
    shift_neg = [_along(C(np.fft.ifftshift(np.exp(-1j * kv[a] * kgrid.spacings[a] / 2))
                         if use_sg else np.ones(shape[a])), a, dim) for a in range(dim)]
    split_field = "u_split_field" in record

    def tsig(src, name, t):
        """source.X(sig_index, t) for one time column."""
        s = src.sigs[name]
        col = s[:, t]
        if src.labelled:
            return col[src.sig_index]
        if s.shape[0] == 1:
            return np.full(len(src.pos), col[0], dtype=rdt)
        return col

    def apply_source(field, src, name, t):
        if src.mode == "dirichlet":
            src.pos.set(field, tsig(src, name, t))
            return field
        if src.mode == "additive":
            sm = np.zeros(shape, dtype=rdt)
            src.pos.set(sm, tsig(src, name, t))
            sm = ifftn_real(source_kappa * fftn(sm)).astype(rdt)
            return field + sm
        field = field.copy()
        src.pos.set(field, src.pos.get(field) + tsig(src, name, t))
        return field

    def extract(file_index):
        """LP/kspaceFirstOrder_extractSensorData.m:41-292 (binary mask)."""
        fi = file_index  # 1-based
        pv = sensor_idx.get(p)
        if "p" in sd:
            sd["p"][:, fi - 1] = pv
            if dir_filters is not None:
                # kspaceFirstOrder2D.m:945-947, LP/directionalResponse.m: the
                # time series (only) from p_k filtered per sensor angle
                for filt, sel in dir_filters:
                    pd = np.real(sfft.ifftn(p_k * filt, workers=W)).astype(rdt)
                    sd["p"][sel, fi - 1] = sensor_idx.get(pd)[sel]
        if "p_max" in record:
            sd["p_max"] = pv.copy() if fi == 1 else np.maximum(sd["p_max"], pv)
        if "p_min" in record:
            sd["p_min"] = pv.copy() if fi == 1 else np.minimum(sd["p_min"], pv)
        if "p_rms" in record:
            sd["p_rms"] = np.sqrt((sd["p_rms"] ** 2 * rdt(fi - 1) + pv ** 2) / rdt(fi)).astype(rdt)
        ucomp = ["ux", "uy", "uz"][:dim]
        if "u" in record:
            for a, u in enumerate(ucomp):
                sd[u][:, fi - 1] = sensor_idx.get(us[a])
        if u_ns_needed:
            for a, u in enumerate(ucomp):
                sh = np.real(sfft.ifft(shift_neg[a] * sfft.fft(us[a], axis=a, workers=W),
                                       axis=a, workers=W)).astype(rdt)
                sd[u + "_non_staggered"][:, fi - 1] = sensor_idx.get(sh)
        if split_field:
            comp, shear = velocity_split(us, kgrid, use_sg, rdt, W)
            for a, u in enumerate(ucomp):
                sd[u + "_split_p"][:, fi - 1] = sensor_idx.get(comp[a])
                sd[u + "_split_s"][:, fi - 1] = sensor_idx.get(shear[a])
        for stat in ("max", "min"):
            if "u_" + stat in record:
                for a, u in enumerate(ucomp):
                    v = sensor_idx.get(us[a])
                    key = f"{u}_{stat}"
                    if fi == 1:
                        sd[key] = v.copy()
                    else:
                        sd[key] = np.maximum(sd[key], v) if stat == "max" else np.minimum(sd[key], v)
        if "u_rms" in record:
            for a, u in enumerate(ucomp):
                v = sensor_idx.get(us[a])
                key = f"{u}_rms"
                sd[key] = np.sqrt((sd[key] ** 2 * rdt(fi - 1) + v ** 2) / rdt(fi)).astype(rdt)
        # whole-grid statistics (:459-569)
        for nm, fld, op in (("p_max_all", [p], np.maximum), ("p_min_all", [p], np.minimum)):
            if nm in record:
                sd[nm] = fld[0].copy() if fi == 1 else op(sd[nm], fld[0])
        for stat, op in (("max", np.maximum), ("min", np.minimum)):
            if f"u_{stat}_all" in record:
                for a, u in enumerate(ucomp):
                    key = f"{u}_{stat}_all"
                    sd[key] = us[a].copy() if fi == 1 else op(sd[key], us[a])

    u_ns_needed = bool(record & {"u_non_staggered", "I", "I_avg"})
    dim_c = rdt(dim)
    dir_filters = None
    if use_sensor and not time_rev and _has(sensor, "directivity_angle"):
        # LP/kspaceFirstOrder_inputChecking.m:312-343 (2-D, binary masks)
        if dim != 2:
            raise ValueError("sensor.directivity_angle is only supported in 2D")
        ang = np.asarray(_get(sensor, "directivity_angle"), dtype=np.float64)
        if ang.shape != shape or interp is not None or reorder_index is not None or cuboids is not None:
            raise ValueError("sensor.directivity_angle and sensor.mask must be the same size.")
        pattern = _get(sensor, "directivity_pattern", "pressure")
        dsize = float(_get(sensor, "directivity_size", 10 * max(kgrid.dx, kgrid.dy)))
        at_sensors = sensor_idx.get(ang)
        uniq = np.unique(at_sensors)
        dir_filters = [(f.astype(rdt), at_sensors == th)
                       for f, th in zip(directivity_filters(kgrid, uniq, pattern, dsize), uniq)]

    # -- time loop (kspaceFirstOrder3D.m:769-1208) ---------------------------
    index_end = Nt - 1 if time_rev else Nt
    import time as _time
    for t_index in range(1, index_end + 1):
        t = t_index - 1
        _t0 = _time.perf_counter()
        if time_rev:
            # :772-788 (2D :715-729)
            sensor_idx.set(p, tr_data[:, t].astype(rdt))
            p_k = fftn(p)
            if dim == 3:
                rmod = p / (rdt(3) * c0 ** 2)
            elif dim == 2:
                rmod = rdt(0.5) * p / (c0 ** 2)
            else:
                rmod = p / (c0 ** 2)   # kspaceFirstOrder1D.m:607
            rmod = np.broadcast_to(rmod, shape)
            for a in range(dim):
                sensor_idx.set(rhos[a], sensor_idx.get(rmod))
        # u update (:792-803)
        kp = kappa * p_k
        for a in range(dim):
            g = ifftn_real(ddk_pos[a] * kp).astype(rdt)
            us[a] = (pml_sg[a] * (pml_sg[a] * us[a] - dt_c * rho0_sg_inv[a] * g)).astype(rdt)
        # velocity sources (:806-886)
        if src_u is not None:
            for a, nm in enumerate(u_names):
                if nm in src_u.sigs and src_u.flag[nm] >= t_index:
                    us[a] = apply_source(us[a], src_u, nm, t).astype(rdt)
        # divergence (:907-909)
        du = [ifftn_real(ddk_neg[a] * (kappa * fftn(us[a]))).astype(rdt) for a in range(dim)]
        # density update (:912-927)
        if not nonlinear:
            for a in range(dim):
                rhos[a] = (pml[a] * (pml[a] * rhos[a] - dt_c * rho0 * du[a])).astype(rdt)
        else:
            rpr = rdt(2) * sum(rhos[1:], rhos[0]) + rho0
            for a in range(dim):
                rhos[a] = (pml[a] * (pml[a] * rhos[a] - dt_c * rpr * du[a])).astype(rdt)
        # pressure source (:930-962)
        if src_p is not None and src_p.flag["p"] >= t_index:
            if src_p.mode == "additive":
                sm = np.zeros(shape, dtype=rdt)
                src_p.pos.set(sm, tsig(src_p, "p", t))
                sm = ifftn_real(source_kappa * fftn(sm)).astype(rdt)
                rhos = [(r + sm).astype(rdt) for r in rhos]
            else:
                for a in range(dim):
                    rhos[a] = apply_source(rhos[a], src_p, "p", t)
        # equation of state (:965-1019)
        rsum = sum(rhos[1:], rhos[0])
        if eos == "lossless":
            inner = rsum
        elif eos == "absorbing":
            dsum = sum(du[1:], du[0])
            inner = (rsum + tau * ifftn_real(nabla1 * fftn(rho0 * dsum)).astype(rdt)
                     - eta * ifftn_real(nabla2 * fftn(rsum)).astype(rdt))
        else:  # stokes
            dsum = sum(du[1:], du[0])
            inner = rsum + tau * rho0 * dsum
        if nonlinear:
            inner = inner + BonA * rsum ** 2 / (rdt(2) * rho0)
        p = (c0 ** 2 * inner).astype(rdt)
        p = np.broadcast_to(p, shape).copy()
        # initial condition (:1023-1037)
        if t_index == 1 and p0 is not None:
            p = p0.copy()
            for a in range(dim):
                rhos[a] = np.broadcast_to(p0 / (dim_c * c0 ** 2), shape).astype(rdt).copy()
            pk0 = fftn(p)
            for a in range(dim):
                g = ifftn_real(ddk_pos[a] * (kappa * pk0)).astype(rdt)
                us[a] = np.broadcast_to(dt_c * rho0_sg_inv[a] * g / rdt(2), shape).astype(rdt).copy()
        p_k = fftn(p)
        # sensor data (:1045-1147)
        if use_sensor and not time_rev and t_index >= record_start_index:
            extract(t_index - record_start_index + 1)
        if step_times is not None:
            step_times.append(_time.perf_counter() - _t0)

    if time_rev:
        sensor_idx.set(p, tr_data[:, index_end].astype(rdt))  # :1211-1213

    # -- post (:1300-1383) ---------------------------------------------------
    if "p_rms" in sd:
        sd["p_rms"] = sd["p_rms"].astype(rdt)
    if time_rev or "p_final" in record:
        sd["p_final"] = p.copy()
    if "u_final" in record:
        for a, u in enumerate(["ux", "uy", "uz"][:dim]):
            sd[u + "_final"] = us[a].copy()
    # dataRecast, then saveIntensity, then reorderCartData (:1327-1343)
    if data_recast:
        sd = {k: v.astype(np.float64) for k, v in sd.items()}
    if use_sensor and not time_rev and record & {"I", "I_avg"}:
        sd = save_intensity(sd, dim, record)
    if reorder_index is not None and use_sensor and not time_rev:
        for key in list(sd.keys()):
            if key.endswith("_all") or key.endswith("_final"):
                continue
            sd[key] = reorder_binary_sensor_data(sd[key], reorder_index)
    if use_sensor and not time_rev and _has(sensor, "frequency_response") and "p" in sd:
        # kspaceFirstOrder3D.m:1346-1348
        fr = np.asarray(_get(sensor, "frequency_response"), dtype=np.float64).ravel()
        sd["p"] = gaussian_filter(sd["p"], 1.0 / dt, fr[0], fr[1])
    if return_state:
        return sd, SimpleNamespace(p=p, rho=rhos, u=us)
    if not use_sensor:
        return None
    if time_rev:
        return sd["p_final"]
    if cuboids is not None:
        # kspaceFirstOrder2D.m:1109-1130: a struct array even without
        # sensor.record
        return reorder_cuboid_corners(sd, cuboids, dim)
    if not record_given:
        return sd["p"]
    return sd


def rel_linf(a, b):
    """Relative L-inf: max|a-b| / max|b| (the gate used throughout)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


# ---------------------------------------------------------------------------
# Axisymmetric solver (kspaceFirstOrderAS.m), 'WSWA-FFT' radial symmetry
# ---------------------------------------------------------------------------

def as_mirror_p(f):
    """kspaceFirstOrderAS.m:703-707: a p-like field of the radial grid on the
    expanded grid, [f, 0, -flip(f(2:end)), -f, 0, flip(f(2:end))] along y."""
    nx, ny = f.shape
    e = np.zeros((nx, 4 * ny), dtype=f.dtype)
    e[:, 0:ny] = f
    e[:, ny + 1:2 * ny] = -f[:, :0:-1]
    e[:, 2 * ny:3 * ny] = -f
    e[:, 3 * ny + 1:4 * ny] = f[:, :0:-1]
    return e


def as_mirror_u(f, signs=(1, 1, -1, -1)):
    """kspaceFirstOrderAS.m:944-948: a y-staggered field, [f, flip(f), -f,
    -flip(f)] (signs per copy; the u_r / r field takes (1, -1, -1, 1))."""
    nx, ny = f.shape
    e = np.zeros((nx, 4 * ny), dtype=f.dtype)
    for b, sg in enumerate(signs):
        e[:, b * ny:(b + 1) * ny] = sg * (f if b % 2 == 0 else f[:, ::-1])
    return e


def as_mirror_p_ws(f):
    """kspaceFirstOrderAS.m:720-722 ('WSWS-FFT'): a p-like field on the
    expanded grid of 2 Ny - 2 points, [f, flip(f(2:end-1))] along y."""
    return np.concatenate([f, f[:, -2:0:-1]], axis=1)


def as_mirror_u_ws(f, sign=-1):
    """kspaceFirstOrderAS.m:966-977 ('WSWS-FFT'): a y-staggered field
    without its last point, [f(1:end-1), sign flip(f(1:end-1))]; sign -1
    for u_r (HAHA), +1 for u_r / r (HSHS)."""
    return np.concatenate([f[:, :-1], sign * f[:, -2::-1]], axis=1)


def _as_pml_outside(kgrid, medium, source, sensor, *, pml_size, smooth, radial_symmetry, **kw):
    """'PMLInside' false for the axisymmetric code: p0 smoothed on the
    user's grid (kspaceFirstOrder_inputChecking.m:1501-1535), then the grid
    grows by the PML on both axial sides and on the outer radial side only
    (kspaceFirstOrder_expandGridMatrices.m:38-48,62-70), medium maps by
    their edge values, p0 and masks by zeros; the _final outputs are cut
    back to the user's grid (x1_inside.., y1_inside = 1,
    kspaceFirstOrder_inputChecking.m:1737-1749)."""
    px, py = (pml_size, pml_size) if np.ndim(pml_size) == 0 else pml_size
    Nx, Ny = kgrid.Nx, kgrid.Ny
    sizes = [(px, px), (0, py)]
    grow = lambda a, edge: np.pad(np.asarray(a), sizes, mode="edge" if edge else "constant")  # noqa: E731
    med = dict(medium)
    for nm in ("sound_speed", "density", "alpha_coeff", "BonA"):
        if nm in med and np.size(med[nm]) > 1:
            med[nm] = grow(np.asarray(med[nm], dtype=np.float64), True)
    src = dict(source)
    if "p0" in src:
        p0 = np.asarray(src["p0"], dtype=np.float64)
        if smooth:
            mirror = as_mirror_p_ws if radial_symmetry == "WSWS-FFT" else as_mirror_p
            p0 = smooth_mat(mirror(p0), restore_max=True)[:, :Ny]
        src["p0"] = grow(p0, False)
    for nm in ("p_mask", "u_mask"):
        if nm in src:
            src[nm] = grow(src[nm], False)
    sen = dict(sensor)
    sen["mask"] = grow(np.asarray(sensor["mask"]), False)
    kg = KGrid(Nx + 2 * px, kgrid.dx, Ny + py, kgrid.dy)
    kg.set_time(kgrid.Nt, kgrid.dt)
    out = kspace_first_order_as(kg, med, src, sen, pml_size=(px, py), smooth=False,
                                radial_symmetry=radial_symmetry, **kw)
    return {k: (v[px:px + Nx, :Ny] if k.endswith(("_final", "_all")) else v) for k, v in out.items()}


def kspace_first_order_as(kgrid: KGrid, medium, source, sensor, *, data_cast="off",
                          pml_size=20, pml_alpha=2.0, smooth=True, radial_symmetry="WSWA-FFT",
                          pml_inside=True, cart_interp="linear"):
    """kspaceFirstOrderAS.m:355-1435 with 'RadialSymmetry' 'WSWA-FFT' (or
    'WSWS-FFT', the expanded grid of 2 Ny - 2 points) for the
    cases the library runs: an initial pressure or a pressure source (every
    p_mode, :1055-1128, and p_frequency_ref), velocity sources ux and uy
    (every u_mode, :788-936), lossless or Stokes media
    (linear or with B/A; the axisymmetric code takes medium.alpha_coeff as
    Stokes absorption with alpha_power 2, kspaceFirstOrder_inputChecking.m:
    173-190, kspaceFirstOrderAS.m:1134-1163), a binary sensor mask recording
    'p', 'p_max', 'p_min',
    'p_rms', 'p_final', 'u' and 'u_final' (x axial, y radial from the
    axis), or time reversal of sensor.time_reversal_boundary_data (:640-650,
    690-702, 1343-1345).  Returns the sensor data dict (p as [points x Nt]);
    {'p_final': p} for time reversal."""
    if not pml_inside:
        return _as_pml_outside(kgrid, medium, source, sensor, data_cast=data_cast,
                               pml_size=pml_size, pml_alpha=pml_alpha, smooth=smooth,
                               radial_symmetry=radial_symmetry)
    rdt = np.float32 if data_cast == "single" else np.float64
    cdt = np.complex64 if data_cast == "single" else np.complex128
    Nx, Ny = kgrid.Nx, kgrid.Ny
    dx, dy, dt, Nt = kgrid.dx, kgrid.dy, kgrid.dt, kgrid.Nt
    px, py = (pml_size, pml_size) if np.ndim(pml_size) == 0 else pml_size
    ax, ay = (pml_alpha, pml_alpha) if np.ndim(pml_alpha) == 0 else pml_alpha
    c0 = np.asarray(_get(medium, "sound_speed"), dtype=np.float64)
    rho0 = np.asarray(_get(medium, "density"), dtype=np.float64)
    bona = _get(medium, "BonA")
    nonlinear = bona is not None
    c_ref = float(_get(medium, "sound_speed_ref", np.max(c0)))
    stokes = _has(medium, "alpha_coeff")
    if stokes:
        tau = (-2.0 * db2neper(_get(medium, "alpha_coeff"), 2) * c0)
    p0 = source.get("p0")
    p0 = None if p0 is None else np.asarray(p0, dtype=np.float64)
    ws = radial_symmetry == "WSWS-FFT"
    if not ws and radial_symmetry != "WSWA-FFT":
        raise ValueError(radial_symmetry)
    mirror_p = as_mirror_p_ws if ws else as_mirror_p
    if ws:
        mirror_u = lambda f: as_mirror_u_ws(f, -1)  # noqa: E731
        mirror_u_on_y = lambda f: as_mirror_u_ws(f, 1)  # noqa: E731
    else:
        mirror_u = as_mirror_u
        mirror_u_on_y = lambda f: as_mirror_u(f, (1, -1, -1, 1))  # noqa: E731
    M = 2 * Ny - 2 if ws else 4 * Ny
    if p0 is not None and smooth:
        # kspaceFirstOrder_inputChecking.m:1504-1535: smoothed on the
        # expanded grid (WS origin symmetry), trimmed back
        p0 = smooth_mat(mirror_p(p0), restore_max=True)[:, :Ny]
    # pressure source (kspaceFirstOrderAS.m:1055-1128) with the scaling of
    # scaleSourceTerms.m for a 2-D grid (row i scaled with c0 at source
    # point i, the reference's quirk)
    psrc = None
    if "p_mask" in source:
        pmask = np.asarray(source["p_mask"])
        pos = np.flatnonzero(pmask.reshape(-1, order="F"))
        sig = np.atleast_2d(np.asarray(source["p"], dtype=np.float64)).copy()
        mode = source.get("p_mode", "additive")
        if "p_frequency_ref" in source:
            # kspaceFirstOrder_inputChecking.m:819-836 and
            # kspaceFirstOrder_scaleSourceTerms.m:42-45
            sig = sig * np.cos(2 * np.pi * float(source["p_frequency_ref"]) * dt / 2)
            mode = "additive-no-correction"
        if sig.shape[0] == 1:
            sig_idx = np.zeros(pos.size, dtype=int)
        elif pmask.max() > 1:
            sig_idx = pmask.reshape(-1, order="F")[pos].astype(int) - 1
        else:
            sig_idx = np.arange(pos.size)
        cflat = np.broadcast_to(c0, (Nx, Ny)).reshape(-1, order="F")
        for i in range(sig.shape[0]):
            ci = cflat[pos[i]]
            sig[i] = sig[i] / (2 * ci ** 2) if mode == "dirichlet" else sig[i] * (2 * dt / (2 * ci * dx))
        psrc = (pos, sig, sig_idx, mode)
    # velocity sources (kspaceFirstOrderAS.m:788-936): one mask and mode for
    # ux (axial) and uy (radial), scaled as scaleSourceTerms.m:286-350 does
    # (2 c0 dt / dx and 2 c0 dt / dy, row i with c0 at source point i)
    usrc = {}
    if "u_mask" in source:
        umask = np.asarray(source["u_mask"])
        upos = np.flatnonzero(umask.reshape(-1, order="F"))
        umode = source.get("u_mode", "additive")
        uwt = 1.0
        if "u_frequency_ref" in source:
            uwt = np.cos(2 * np.pi * float(source["u_frequency_ref"]) * dt / 2)
            umode = "additive-no-correction"
        cflat = np.broadcast_to(c0, (Nx, Ny)).reshape(-1, order="F")
        for nm, dd in (("ux", dx), ("uy", dy)):
            if nm not in source:
                continue
            sig = uwt * np.atleast_2d(np.asarray(source[nm], dtype=np.float64))
            if sig.shape[0] == 1:
                sidx = np.zeros(upos.size, dtype=int)
            elif umask.max() > 1:
                sidx = umask.reshape(-1, order="F")[upos].astype(int) - 1
            else:
                sidx = np.arange(upos.size)
            if umode != "dirichlet":
                sig = np.stack([sig[i] * (2 * cflat[upos[i]] * dt / dd) for i in range(sig.shape[0])])
            usrc[nm] = (sig, sidx)
    tr = sensor.get("time_reversal_boundary_data") if isinstance(sensor, dict) else None
    # operators (kspaceFirstOrderAS.m:463-565)
    pml_x = get_pml(Nx, dx, dt, c_ref, px, ax, False)
    pml_x_sg = get_pml(Nx, dx, dt, c_ref, px, ax, True)
    pml_y = get_pml(Ny, dy, dt, c_ref, py, ay, False)
    pml_y_sg = get_pml(Ny, dy, dt, c_ref, py, ay, True)
    pml_y[:py] = 1.0  # getPML.m: axisymmetric, no PML at the axis
    pml_y_sg[:py] = 1.0
    kx = make_dim(Nx, dx)
    ky = make_dim(M, dy)
    k = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    kappa = np.fft.ifftshift(sinc(c_ref * k * dt / 2)).astype(rdt)
    ddx_pos = np.fft.ifftshift(1j * kx * np.exp(1j * kx * dx / 2))[:, None].astype(cdt)
    ddx_neg = np.fft.ifftshift(1j * kx * np.exp(-1j * kx * dx / 2))[:, None].astype(cdt)
    ddy_k = np.fft.ifftshift(1j * ky)[None, :].astype(cdt)
    y_pos = np.fft.ifftshift(np.exp(1j * ky * dy / 2))[None, :].astype(cdt)
    y_neg = np.fft.ifftshift(np.exp(-1j * ky * dy / 2))[None, :].astype(cdt)
    y_sg = (np.arange(Ny) * dy + dy / 2)[None, :]
    source_kappa = np.fft.ifftshift(np.cos(c_ref * k * dt / 2)).astype(rdt)
    if rho0.ndim == 2:
        # interpn '*linear' at x + dx/2 and y + dy/2, NaN past the last point
        # replaced by rho0 (kspaceFirstOrderAS.m:440-456)
        rsx = rho0.copy()
        rsx[:-1, :] = (rho0[:-1, :] + rho0[1:, :]) / 2
        rsy = rho0.copy()
        rsy[:, :-1] = (rho0[:, :-1] + rho0[:, 1:]) / 2
    else:
        rsx = rsy = rho0
    rsx_inv = (1.0 / rsx).astype(rdt)
    rsy_inv = (1.0 / rsy).astype(rdt)
    c2 = (c0 ** 2).astype(rdt)
    rho0c = np.asarray(rho0, dtype=rdt)
    ba = np.asarray(bona, dtype=rdt) if nonlinear else None
    tau_c = np.asarray(tau, dtype=rdt) if stokes else None
    pml_x, pml_x_sg = pml_x.astype(rdt)[:, None], pml_x_sg.astype(rdt)[:, None]
    pml_y, pml_y_sg = pml_y.astype(rdt)[None, :], pml_y_sg.astype(rdt)[None, :]
    inv_y = (1.0 / y_sg).astype(rdt)
    dtc = rdt(dt)

    mask = np.asarray(sensor["mask"])
    interp = reorder = None
    if mask.shape != (Nx, Ny):
        # Cartesian points (x, r) (kspaceFirstOrder_inputChecking.m:590-633,
        # cart2grid.m with axisymmetric true); time reversal takes the
        # nearest grid points with its data sorted into their order
        gm, order_index, reorder_index = cart2grid(kgrid, mask, axisymmetric=True)
        if tr is not None:
            tr = np.asarray(tr, dtype=np.float64)[np.argsort(order_index, kind="stable")]
        elif cart_interp == "linear":
            interp = grid_data_linear(kgrid, mask, axisymmetric=True)
        else:
            reorder = reorder_index
        mask = gm
    mask = np.asarray(mask, dtype=bool)
    idx = np.flatnonzero(mask.reshape(-1, order="F"))
    if interp is not None:
        # one row per Cartesian point (extractSensorData.m: sum(f(tri) .* bc, 2))
        interp = _Interp(interp[0], interp[1], rdt)
        idx = np.arange(len(interp))
    record = list(sensor.get("record", ["p"]))
    start = int(sensor.get("record_start_index", 1))
    out = {}
    nrec = Nt - start + 1
    if "p" in record:
        out["p"] = np.zeros((idx.size, nrec))
    if "u" in record:
        out["ux"] = np.zeros((idx.size, nrec))
        out["uy"] = np.zeros((idx.size, nrec))
    pmax = np.full(idx.size, -np.inf)
    pmin = np.full(idx.size, np.inf)
    prms = np.zeros(idx.size)

    fft2 = lambda a: sfft.fft2(a, workers=FFT_WORKERS)  # noqa: E731
    ifft2 = lambda a: np.real(sfft.ifft2(a, workers=FFT_WORKERS)).astype(rdt)  # noqa: E731
    p = np.zeros((Nx, Ny), dtype=rdt)
    ux = np.zeros_like(p)
    uy = np.zeros_like(p)
    rx = np.zeros_like(p)
    ry = np.zeros_like(p)
    p0c = None if p0 is None else p0.astype(rdt)
    if tr is not None:
        trd = np.asarray(tr, dtype=np.float64)[:, ::-1]   # fliplr (:640-650)
        Nt_loop = Nt - 1
    else:
        Nt_loop = Nt
    c2_flat = np.broadcast_to(c2, (Nx, Ny))

    def flat_set(a, idx_, vals):
        f = a.reshape(-1, order="F").copy()
        f[idx_] = vals
        return f.reshape(a.shape, order="F").astype(rdt)

    def flat_add(a, idx_, vals):
        f = a.reshape(-1, order="F").copy()
        f[idx_] = f[idx_] + vals
        return f.reshape(a.shape, order="F").astype(rdt)

    def grad(pp):
        pk = kappa * fft2(mirror_p(pp))
        return ifft2(ddx_pos * pk)[:, :Ny], ifft2(ddy_k * y_pos * pk)[:, :Ny]

    def pick(f):
        if interp is not None:
            return interp.get(f).astype(np.float64)
        return np.reshape(f, -1, order="F")[idx].astype(np.float64)

    for t in range(1, Nt_loop + 1):
        if tr is not None:
            p = flat_set(p, idx, trd[:, t - 1])
            rmod = (0.5 * p / c2_flat).reshape(-1, order="F")[idx]
            rx = flat_set(rx, idx, rmod)
            ry = flat_set(ry, idx, rmod)
        dpdx, dpdy = grad(p)
        ux = pml_x_sg * (pml_x_sg * ux - dtc * rsx_inv * dpdx)
        uy = pml_y_sg * (pml_y_sg * uy - dtc * rsy_inv * dpdy)
        for nm, layout in (("ux", mirror_p), ("uy", mirror_u)):
            if nm not in usrc or t > usrc[nm][0].shape[1]:
                continue
            v = usrc[nm][0][usrc[nm][1], t - 1]
            f = ux if nm == "ux" else uy
            if umode == "dirichlet":
                f = flat_set(f, upos, v)
            elif umode == "additive":
                sm_ = np.zeros(Nx * Ny)
                sm_[upos] = v
                sm_ = sm_.reshape((Nx, Ny), order="F")
                f = (f + ifft2(source_kappa * fft2(layout(sm_)))[:, :Ny]).astype(rdt)
            else:
                f = flat_add(f, upos, v)
            if nm == "ux":
                ux = f
            else:
                uy = f
        duxdx = ifft2(ddx_neg * kappa * fft2(mirror_p(ux)))[:, :Ny]
        u_on_y = mirror_u_on_y(inv_y * uy)
        duydy = ifft2(kappa * y_neg * (ddy_k * fft2(mirror_u(uy)) + fft2(u_on_y)))[:, :Ny]
        if nonlinear:
            k_ = 2 * (rx + ry) + rho0c
        else:
            k_ = rho0c
        rx = pml_x * (pml_x * rx - dtc * k_ * duxdx)
        ry = pml_y * (pml_y * ry - dtc * k_ * duydy)
        if psrc is not None and t <= psrc[1].shape[1]:
            pos, sig, sig_idx, mode = psrc
            v = sig[sig_idx, t - 1]
            if mode == "dirichlet":
                rx = flat_set(rx, pos, v)
                ry = flat_set(ry, pos, v)
            elif mode == "additive":
                sm_ = np.zeros((Nx, Ny)).reshape(-1, order="F")
                sm_[pos] = v
                sm_ = sm_.reshape((Nx, Ny), order="F")
                add = ifft2(source_kappa * fft2(mirror_p(sm_)))[:, :Ny]
                rx = (rx + add).astype(rdt)
                ry = (ry + add).astype(rdt)
            else:
                rx = flat_add(rx, pos, v)
                ry = flat_add(ry, pos, v)
        st = tau_c * rho0c * (duxdx + duydy) if stokes else 0.0
        if nonlinear:
            p = c2 * (rx + ry + st + ba * (rx + ry) ** 2 / (2 * rho0c))
        else:
            p = c2 * (rx + ry + st)
        if t == 1 and p0c is not None:
            p = p0c.copy()
            rx = (p0c / (2 * c2)).astype(rdt)
            ry = rx.copy()
            dpdx, dpdy = grad(p)
            ux = dtc * rsx_inv * dpdx / 2
            uy = dtc * rsy_inv * dpdy / 2
        if tr is None and t >= start:
            j = t - start
            pv = pick(p)
            if "p" in out:
                out["p"][:, j] = pv
            if "ux" in out:
                out["ux"][:, j] = pick(ux)
                out["uy"][:, j] = pick(uy)
            pmax = np.maximum(pmax, pv)
            pmin = np.minimum(pmin, pv)
            prms += pv ** 2
    if tr is not None:
        # :1343-1345: the final data column, then p_final
        p = flat_set(p, idx, trd[:, Nt - 1])
        return {"p_final": p.astype(np.float64)}
    res = {}
    if "p" in record:
        res["p"] = out["p"]
    if "p_max" in record:
        res["p_max"] = pmax
    if "p_min" in record:
        res["p_min"] = pmin
    if "p_rms" in record:
        res["p_rms"] = np.sqrt(prms / nrec)
    if "p_final" in record:
        res["p_final"] = p.astype(np.float64)
    if "u" in record:
        res["ux"], res["uy"] = out["ux"], out["uy"]
    if "u_final" in record:
        res["ux_final"], res["uy_final"] = ux.astype(np.float64), uy.astype(np.float64)
    if reorder is not None:
        # LP/kspaceFirstOrder_reorderCartData.m: rows back in the order of
        # the Cartesian points
        for k in ("p", "p_max", "p_min", "p_rms", "ux", "uy"):
            if k in res:
                res[k] = reorder_binary_sensor_data(res[k], reorder)
    return res
