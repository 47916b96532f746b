"""Pins the oracle's operators against closed forms and library routines
(CPU only).

Each check is chosen so that a plausible mistake fails it: a wrong sign of
the staggered shift, a transposed window coefficient, an off-by-one in the
PML profile, MATLAB's round-half-away-from-zero, the Nyquist bin.
Mirrors +kwave/+tests/+unit/TestFourierCollocation.m:47-302 (spectral
derivatives of sinusoids with +-dx/2 staggering, sinc) and
TestSplitFieldPML.m:74-107 (PML is identity inside the domain).
"""

import math

import numpy as np
import pytest

from oracle import kwave_oracle as ko


@pytest.mark.parametrize("N", [8, 9, 16, 27, 64, 75])
def test_make_dim_matches_fftfreq(N):
    dx = 1e-4
    k = ko.make_dim(N, dx)
    ref = np.fft.fftshift(np.fft.fftfreq(N, d=dx)) * 2 * math.pi
    assert np.allclose(k, ref, rtol=0, atol=1e-9 * np.max(np.abs(ref)))
    assert k[N // 2] == 0.0


def test_sinc_closed_form():
    x = np.array([0.0, 0.1, 1.0, math.pi, 2.5, -3.0])
    got = ko.sinc(x)
    assert got[0] == pytest.approx(1.0, abs=1e-15)
    assert got[3] == pytest.approx(0.0, abs=1e-15)
    assert np.allclose(got[1:], np.sinc(x[1:] / math.pi), rtol=1e-14, atol=1e-15)


def test_db2neper_definition():
    # alpha0 dB/(MHz^y cm) at f MHz is alpha0 f^y dB/cm = alpha0 f^y 100/(20 log10 e) Np/m,
    # and the Np coefficient multiplies omega^y
    for a0, y, f in ((0.75, 1.5, 1.0), (0.5, 1.3, 2.5), (2.0, 2.0, 0.3)):
        a_np = ko.db2neper(a0, y)
        np_per_m = a0 * f ** y * 100 / (20 * math.log10(math.e))
        assert a_np * (2 * math.pi * f * 1e6) ** y == pytest.approx(np_per_m, rel=1e-12)


@pytest.mark.parametrize("staggered", [False, True])
def test_pml_profile(staggered):
    N, n, dx, dt, c, a = 64, 10, 1e-4, 2e-8, 1500.0, 2.0
    pml = ko.get_pml(N, dx, dt, c, n, a, staggered)
    assert np.all(pml[n:N - n] == 1.0)            # identity inside
    assert np.all(pml[:n] < 1.0) and np.all(pml[N - n:] < 1.0)
    # quartic profile: left x=1 -> full strength; right x=n -> full strength
    off = 0.5 if staggered else 0.0
    x = np.arange(1, n + 1) + off
    right = np.exp(-a * (c / dx) * (x / n) ** 4 * dt / 2)
    left = np.exp(-a * (c / dx) * ((x - n - 1) / -n) ** 4 * dt / 2)
    assert np.allclose(pml[N - n:], right, rtol=1e-14)
    assert np.allclose(pml[:n], left, rtol=1e-14)
    if not staggered:
        assert np.allclose(pml[:n], pml[N - n:][::-1], rtol=1e-14)   # symmetric
    assert np.all(np.diff(pml[:n]) > 0)             # absorption grows towards the edge


@pytest.mark.parametrize("N", [8, 17, 64])
def test_blackman_coefficients(N):
    assert np.allclose(ko.blackman_1d(N), np.blackman(N), rtol=0, atol=1e-14)


def _blackman_closed(n, N):
    # the textbook Blackman window 0.42 - 0.5 cos(2 pi n/(N-1)) + 0.08 cos(4 pi n/(N-1))
    return 0.42 - 0.5 * math.cos(2 * math.pi * n / (N - 1)) + 0.08 * math.cos(4 * math.pi * n / (N - 1))


def _window_fns():
    from kwhip.solver import blackman_window
    return [("oracle", lambda shape: np.abs(ko.get_win_rotated(shape, [bool(n % 2) for n in shape]))),
            ("product", blackman_window)]


@pytest.mark.parametrize("which", [0, 1])
def test_rotated_window_closed_forms(which):
    """getWin.m:260-347 ('Rotation' true) as smooth.m calls it, for the oracle
    and for the product's host copy: on an odd grid the central row and
    column are the 1-D Blackman window; the window depends on the radius only
    (transposes and mirrors agree); off the axes it is the linear
    interpolation of the 1-D window at the radius (hand-computed at r = 5,
    an integer radius on a diagonal, and at r = sqrt 2); beyond the radius
    it is clipped to the window's end value; an even grid is the window of
    the next odd size with its last row and column dropped."""
    name, win = _window_fns()[which]
    N, c = 17, 8
    w = win((N, N))
    b = np.array([_blackman_closed(n, N) for n in range(N)])
    assert np.allclose(w[:, c], b, rtol=0, atol=1e-15), name
    assert np.allclose(w[c, :], b, rtol=0, atol=1e-15), name
    assert np.allclose(w, w.T, rtol=0, atol=1e-15) and np.allclose(w, w[::-1, :], rtol=0, atol=1e-15)
    assert w[c + 3, c + 4] == pytest.approx(b[c + 5], abs=1e-15)   # r = 5
    r = math.sqrt(2.0)
    assert w[c + 1, c + 1] == pytest.approx(b[c + 1] + (r - 1) * (b[c + 2] - b[c + 1]), abs=1e-15)
    assert w[0, 0] == pytest.approx(abs(b[N - 1]), abs=1e-15)      # clipped to the radius
    w3 = win((N, N, N))
    assert np.allclose(w3[c, c, :], b, rtol=0, atol=1e-15)
    assert w3[c + 2, c + 3, c + 6] == pytest.approx(b[c + 7], abs=1e-15)  # r = 7
    we = win((16, 16))
    assert we.shape == (16, 16) and np.allclose(we, w[:16, :16], rtol=0, atol=1e-15)


@pytest.mark.parametrize("shape", [(32, 32), (31, 33), (16, 18, 15)])
def test_smooth_preserves_constants_and_restores_max(shape):
    c = np.full(shape, 3.5)
    assert np.allclose(ko.smooth_mat(c), 3.5, rtol=1e-12)
    rng = np.random.default_rng(0)
    f = rng.standard_normal(shape)
    sm = ko.smooth_mat(f, True)
    assert np.max(np.abs(sm)) == pytest.approx(np.max(np.abs(f)), rel=1e-12)
    # a low-pass filter: the high-frequency energy drops
    F, S = np.abs(np.fft.fftn(f)), np.abs(np.fft.fftn(ko.smooth_mat(f)))
    assert S.ravel()[0] == pytest.approx(F.ravel()[0], rel=1e-12)
    assert np.sum(S ** 2) < 0.6 * np.sum(F ** 2)


@pytest.mark.parametrize("N", [32, 33])
@pytest.mark.parametrize("use_sg", [True, False])
def test_spectral_derivative_of_sinusoid(N, use_sg):
    """D+ gives the derivative at x + dx/2, D- at x - dx/2 (TestFourierCollocation)."""
    dx = 1e-3
    L = N * dx
    k = ko.make_dim(N, dx)
    x = np.arange(N) * dx
    for m in (1, 3, 7):
        f = np.sin(2 * math.pi * m * x / L)
        dp, dn = ko.shift_derivative_ops(k, dx, use_sg)
        gp = np.real(np.fft.ifft(dp * np.fft.fft(f)))
        gn = np.real(np.fft.ifft(dn * np.fft.fft(f)))
        w = 2 * math.pi * m / L
        sh = dx / 2 if use_sg else 0.0
        assert np.allclose(gp, w * np.cos(w * (x + sh)), atol=1e-9 * w)
        assert np.allclose(gn, w * np.cos(w * (x - sh)), atol=1e-9 * w)


def test_derivative_of_nyquist_mode_even_grid():
    """The Nyquist mode of an even grid: D+ applied to (-1)^n is real-valued
    (Hermitian symmetry kept), and D+ then D- equals -k_nyq^2."""
    N, dx = 16, 1e-3
    k = ko.make_dim(N, dx)
    dp, dn = ko.shift_derivative_ops(k, dx, True)
    f = (-1.0) ** np.arange(N)
    g = np.fft.ifft(dp * np.fft.fft(f))
    assert np.max(np.abs(np.imag(g))) < 1e-10 * np.max(np.abs(g))
    h = np.real(np.fft.ifft(dn * np.fft.fft(np.real(g))))
    assert np.allclose(h, -(math.pi / dx) ** 2 * f, rtol=1e-12)


def test_brute_force_dft_derivative_tiny_grid():
    """The FFT-based operator equals an explicit DFT-matrix evaluation."""
    N, dx = 8, 0.5e-3
    k = ko.make_dim(N, dx)
    dp, _ = ko.shift_derivative_ops(k, dx, True)
    F = np.exp(-2j * math.pi * np.outer(np.arange(N), np.arange(N)) / N)
    Fi = np.conj(F) / N
    D = Fi @ np.diag(dp) @ F
    rng = np.random.default_rng(1)
    f = rng.standard_normal(N)
    assert np.allclose(np.real(D @ f), np.real(np.fft.ifft(dp * np.fft.fft(f))), atol=1e-9)


def test_cart2grid_rounding_and_duplicates():
    kg = ko.KGrid(10, 1.0, 10, 1.0)
    # grid index = round(x/dx) + floor(N/2) + 1 with MATLAB round (half away from 0)
    pts = np.array([[0.0, 0.5, -0.5, 2.4, -2.5, 0.49],
                    [0.0, 0.0, 0.0, 1.0, 1.0, 0.0]])
    mask, order, reorder = ko.cart2grid(kg, pts)
    want = {(5, 5): 6, (6, 5): 2, (4, 5): 3, (7, 6): 4, (2, 6): 5}   # 0-based; point 1 overwritten by 6
    got = {tuple(int(v) for v in ix): None for ix in np.argwhere(mask)}
    assert set(got) == set(want)
    flat = np.zeros((10, 10), dtype=int)
    for (i, j), p in want.items():
        flat[i, j] = p
    assert list(reorder) == list(flat.ravel(order="F")[flat.ravel(order="F") != 0])
    assert list(order) == list(np.argsort(reorder, kind="stable") + 1)
    with pytest.raises(ValueError):
        ko.cart2grid(kg, np.array([[20.0], [0.0]]))


@pytest.mark.parametrize("shape", [(10, 9), (7, 8, 6)])
def test_grid_data_linear_reproduces_affine_fields(shape):
    """Linear interpolation over any triangulation of the grid reproduces an
    affine field exactly at every point inside the grid; the barycentric
    weights are non-negative, sum to one and rebuild the point from the
    simplex vertices."""
    dim = len(shape)
    d = [0.5 + 0.25 * a for a in range(dim)]
    args = []
    for n, dd in zip(shape, d):
        args += [n, dd]
    kg = ko.KGrid(*args)
    rng = np.random.default_rng(3)
    lo = [-math.floor(n / 2) * dd for n, dd in zip(shape, d)]
    hi = [(n - 1 - math.floor(n / 2)) * dd for n, dd in zip(shape, d)]
    pts = np.stack([rng.uniform(l, h, 200) for l, h in zip(lo, hi)])
    pts[:, :2 ** dim] = np.array(np.meshgrid(*[[l, h] for l, h in zip(lo, hi)],
                                             indexing="ij")).reshape(dim, -1)  # the corners
    tri, bc = ko.grid_data_linear(kg, pts)
    assert np.all(bc >= -1e-12) and np.allclose(bc.sum(axis=1), 1.0, atol=1e-12)
    coords = np.meshgrid(*[(np.arange(n) - math.floor(n / 2)) * dd for n, dd in zip(shape, d)],
                         indexing="ij")
    for a in range(dim):
        xa = coords[a].ravel(order="F")
        assert np.allclose((xa[tri] * bc).sum(axis=1), pts[a], atol=1e-12)
    coef = [1.3, -0.7, 2.1][:dim]
    f = 0.4 + sum(c * x for c, x in zip(coef, coords))
    interp = ko._Interp(tri, bc, np.float64)
    assert np.allclose(interp.get(f), 0.4 + sum(c * x for c, x in zip(coef, pts)), atol=1e-12)


def test_grid_data_linear_declared_triangulation():
    """The declared (Kuhn) triangulation on hand-made points: in 2-D cells
    are cut along the diagonal from (i, j) to (i+1, j+1); a grid node takes
    weight 1; points outside the grid raise the reference's error."""
    kg = ko.KGrid(6, 1.0, 6, 1.0)   # nodes at -3 .. 2 on both axes
    # (0.25, 0.75) from node (3, 3) (0-based): below the diagonal side fy > fx,
    # simplex (3,3) -> (3,4) -> (4,4) with weights 1-0.75, 0.75-0.25, 0.25
    tri, bc = ko.grid_data_linear(kg, np.array([[0.25], [0.75]]))
    assert list(tri[0]) == [3 + 6 * 3, 3 + 6 * 4, 4 + 6 * 4]
    assert np.allclose(bc[0], [0.25, 0.5, 0.25])
    tri, bc = ko.grid_data_linear(kg, np.array([[0.75], [0.25]]))
    assert list(tri[0]) == [3 + 6 * 3, 4 + 6 * 3, 4 + 6 * 4]
    assert np.allclose(bc[0], [0.25, 0.5, 0.25])
    tri, bc = ko.grid_data_linear(kg, np.array([[1.0, 2.0], [-3.0, 2.0]]))
    flat = (tri * (bc > 0.5)).sum(axis=1)
    assert list(flat) == [4 + 6 * 0, 5 + 6 * 5] and np.allclose(bc.max(axis=1), 1.0)
    kg3 = ko.KGrid(4, 1.0, 4, 1.0, 4, 1.0)
    tri, bc = ko.grid_data_linear(kg3, np.array([[0.5], [0.25], [0.125]]))
    # f = (0.5, 0.25, 0.125) from node (2, 2, 2): x, then y, then z
    assert list(tri[0]) == [42, 43, 47, 63] and np.allclose(bc[0], [0.5, 0.25, 0.125, 0.125])
    with pytest.raises(ValueError, match="k-space grid"):
        ko.grid_data_linear(kg, np.array([[2.2], [0.0]]))


def test_find_column_major():
    m = np.zeros((3, 4))
    m[2, 0] = m[0, 1] = m[1, 3] = 1
    assert list(ko.find_F(m)) == [2, 3, 10]


@pytest.mark.parametrize("N", [16, 17])
def test_fourier_shift_closed_form(N):
    """fourierShift.m moves a periodic band-limited signal forward by a
    fraction of a sample: x_n = cos(2 pi f n / N + phi) shifted by s is
    cos(2 pi f (n + s) / N + phi) (f below Nyquist); along the last axis of
    a [points x time] array, or the first of a column.  The oracle and the
    host post-processing of kwhip restate it separately; both are held to
    the closed form."""
    from kwhip.solver import fourier_shift as fs_product

    n = np.arange(N)
    rows = np.stack([np.cos(2 * np.pi * f * n / N + 0.3 * f) for f in (1, 3, 5)])
    for s in (0.5, -0.25):
        want = np.stack([np.cos(2 * np.pi * f * (n + s) / N + 0.3 * f) for f in (1, 3, 5)])
        for fs in (ko.fourier_shift, fs_product):
            assert np.max(np.abs(fs(rows, s) - want)) < 1e-12
            col = rows[1][:, None]
            assert np.max(np.abs(fs(col, s)[:, 0] - want[1])) < 1e-12


def test_expand_matrix_known_answer():
    """expandMatrix.m: without an edge value the new elements repeat the
    nearest edge element (corners take the corner value); with one they
    hold it."""
    m = np.arange(6.0).reshape(2, 3)
    want_edge = np.array([[0, 0, 0, 1, 2, 2, 2],
                          [0, 0, 0, 1, 2, 2, 2],
                          [3, 3, 3, 4, 5, 5, 5],
                          [3, 3, 3, 4, 5, 5, 5]], dtype=float)
    assert np.array_equal(ko.expand_matrix(m, [1, 2]), want_edge)
    want_zero = np.zeros((4, 5))
    want_zero[1:3, 1:4] = m
    assert np.array_equal(ko.expand_matrix(m, [1, 1], 0), want_zero)
    m3 = np.arange(24.0).reshape(2, 3, 4)
    assert np.array_equal(ko.expand_matrix(m3, [1, 2, 1]), np.pad(m3, [(1, 1), (2, 2), (1, 1)], mode="edge"))


@pytest.mark.parametrize("shape", [(16, 12), (12, 10, 8), (15, 9)])
def test_velocity_split_of_gradient_and_rotational_fields(shape):
    """velocity_split on single-mode fields sampled at the staggered points
    (x_a + d_a/2 for component a): the gradient of phi = sin(k.x) is purely
    compressional (part p = the field at the grid points, part s = 0), a
    rotational field (k_y f, -k_x f, 0) with f = sin(k.x) purely shear; the
    parts sum to the non-staggered velocity."""
    dim = len(shape)
    d = [1e-3 * (1 + 0.1 * a) for a in range(dim)]
    args = []
    for n, dd in zip(shape, d):
        args += [n, dd]
    kg = ko.KGrid(*args)
    m = [1, 2, 1][:dim]
    kvec = [2 * np.pi * m[a] / (shape[a] * d[a]) for a in range(dim)]
    grids = np.meshgrid(*[np.arange(n) * dd for n, dd in zip(shape, d)], indexing="ij")

    def at(shift_axis):
        # phase k.x at the grid points moved by half a cell along one axis
        return sum(kvec[a] * (grids[a] + (d[a] / 2 if a == shift_axis else 0.0)) for a in range(dim))

    grad = [kvec[a] * np.cos(at(a)) for a in range(dim)]
    grad_ns = [kvec[a] * np.cos(at(-1)) for a in range(dim)]
    comp, shear = ko.velocity_split(grad, kg)
    for a in range(dim):
        assert np.max(np.abs(comp[a] - grad_ns[a])) < 1e-10 * max(kvec)
        assert np.max(np.abs(shear[a])) < 1e-10 * max(kvec)
    rot = [kvec[1] * np.sin(at(0)), -kvec[0] * np.sin(at(1))] + ([np.zeros(shape)] if dim == 3 else [])
    rot_ns = [kvec[1] * np.sin(at(-1)), -kvec[0] * np.sin(at(-1))] + ([np.zeros(shape)] if dim == 3 else [])
    comp, shear = ko.velocity_split(rot, kg)
    for a in range(dim):
        assert np.max(np.abs(comp[a])) < 1e-10 * max(kvec)
        assert np.max(np.abs(shear[a] - rot_ns[a])) < 1e-10 * max(kvec)


@pytest.mark.parametrize("N", [64, 63])
def test_gaussian_filter_gain_on_tones(N):
    """gaussianFilter on sampled tones at FFT bin frequencies: the gain is
    the larger of the two Gaussians at +-freq, exp(-(f - freq)^2 / (2 s^2))
    with FWHM = bandwidth% of freq, i.e. 1 at the centre frequency and 1/2
    at half the bandwidth away from it."""
    Fs = 64e6
    df = Fs / N
    freq = 10 * df
    bw = 40.0
    s2 = (bw / 100 * freq / (2 * math.sqrt(2 * math.log(2)))) ** 2
    t = np.arange(N) / Fs
    for m, ph in ((10, 0.3), (12, 1.1), (7, -0.4), (13, 0.0)):
        f = m * df
        sig = np.cos(2 * np.pi * f * t + ph)[None, :]
        out = ko.gaussian_filter(sig, Fs, freq, bw)
        gain = math.exp(-(f - freq) ** 2 / (2 * s2))
        assert np.allclose(out, gain * sig, atol=1e-12), m
    # FWHM: a tone bw/2 % away from the centre keeps half its amplitude
    half = math.exp(-((bw / 200 * freq) ** 2) / (2 * s2))
    assert abs(half - 0.5) < 1e-12


def test_input_field_names_are_checked():
    """checkFieldNames: a field outside the allowed list of each structure
    raises the reference's error; sensor directivity fields exist in 2-D
    only; sensor.record_mode is refused as deprecated."""
    kg = ko.KGrid(24, 1e-4)
    kg.set_time(5, 2e-8)
    med = {"sound_speed": 1500.0, "density": 1000.0}
    p0 = np.zeros(24)
    p0[10] = 1
    sens = {"mask": np.ones(24)}
    with pytest.raises(ValueError, match="speed is not a valid field for the structure medium"):
        ko.kspace_first_order(kg, dict(med, speed=1.0), {"p0": p0}, sens, smooth=(False,))
    with pytest.raises(ValueError, match="uy is not a valid field for the structure source"):
        ko.kspace_first_order(kg, med, {"p0": p0, "uy": np.zeros(5)}, sens, smooth=(False,))
    with pytest.raises(ValueError, match="directivity_angle is not a valid field for the structure sensor"):
        ko.kspace_first_order(kg, med, {"p0": p0}, dict(sens, directivity_angle=np.zeros(24)),
                              smooth=(False,))
    with pytest.raises(ValueError, match="record_mode input has been deprecated"):
        ko.kspace_first_order(kg, med, {"p0": p0}, dict(sens, record_mode="statistics"),
                              smooth=(False,))


def test_optimal_pml_size_known_answers():
    """getOptimalPMLSize: N + 2 pml with the smallest largest prime factor
    in the range, the first such size: 100 + 28 = 128 = 2^7 -> 14 (120 =
    2^3 3 5 at 10 loses); 101: 121 = 11^2, 123 = 3 41, 125 = 5^3 -> 12 (odd
    lengths cannot reach 3 before 243); 97: 117, 119, 121, 123, 125 -> 14;
    101 in 13..20: 127, 129 = 3 43, 131, 133 = 7 19, 135 = 3^3 5 -> 17."""
    assert ko.optimal_pml_size([100, 101, 97]) == [14, 12, 14]
    assert ko.optimal_pml_size([101], (13, 20)) == [17]
    # 'PMLRange' is rounded as MATLAB's round: 12.5 -> 13 (Python's
    # round-half-even would give 12 and the answer 12)
    assert ko.optimal_pml_size([101], (12.5, 20)) == [17]
    assert ko.optimal_pml_size([101], (12.4, 20)) == [12]
    with pytest.raises(ValueError):
        ko.optimal_pml_size([64], (20, 10))
