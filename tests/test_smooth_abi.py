"""kw_smooth (the device smoothing filter of the C ABI, single precision)
against the host smoothing of the Python layer (double precision, as the
reference's smooth.m): out = real(ifftn(fftn(in) .* ifftshift(win))) with the
rotated Blackman window, on 1-D, 2-D and 3-D grids of even, odd and mixed
sizes."""

import ctypes as C

import numpy as np
import pytest

SHAPES = [(96,), (75,), (32, 30), (27, 25), (16, 18, 20), (15, 9, 25)]


def device_smooth(lib, mat):
    from kwhip import _abi
    from kwhip.solver import blackman_window
    mat = np.asarray(mat, dtype=np.float64)
    N = (C.c_int64 * 3)(*(list(mat.shape) + [1] * (3 - mat.ndim)))
    src = _abi.fortran_flat(mat)
    win = _abi.fortran_flat(blackman_window(mat.shape))
    out = np.empty_like(src)
    _abi.call(lib.kw_smooth, C.c_int32(mat.ndim), N, _abi.dptr(src), _abi.dptr(win), _abi.dptr(out))
    return out.reshape(mat.shape, order="F")


def _check(lib, shape):
    from kwhip.solver import smooth
    rng = np.random.default_rng(len(shape) * 100 + shape[0])
    f = rng.standard_normal(shape)
    ref = smooth(f, False)
    got = device_smooth(lib, f)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err <= 1e-6, (shape, err)


@pytest.mark.parametrize("shape", SHAPES)
def test_kw_smooth_emu(shape):
    from kwhip import _abi
    _check(_abi.load("emu"), shape)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES + [(128, 96), (64, 64, 64)])
def test_kw_smooth_hip(shape, hip_lib):
    _check(hip_lib, shape)
