"""CPU emulator parity: the kernel bodies of the HIP library, built for the
lane-per-thread CPU emulator (every global access bounds-checked, LDS
poisoned with NaN, barrier counts checked), against the oracle on the same
inputs as the GPU parity tests.  Same gates: relative L-inf <= 1e-5 against
the oracle in single precision and <= 1e-4 against double precision.

This runs the host planner and every pass kind without a GPU, including the
fixed-length FFT variants (axis lengths 64, 128, 256, 384) on thin grids.
"""

import numpy as np
import pytest

import cases

TOL32 = 1e-5
TOL64 = 1e-4


@pytest.fixture(autouse=True)
def _few_pencils(monkeypatch):
    # fewer pencils per emulated block: same code paths, faster emulation
    monkeypatch.setenv("KW_MAX_PENCILS", "4")


def _check(cs):
    got = cases.run_kwhip(cs, "emu")
    e32 = cases.compare(got, cases.run_oracle(cs, "single"))
    e64 = cases.compare(got, cases.run_oracle(cs, "off"))
    for k in e32:
        assert e32[k] <= TOL32, (k, e32[k])
        assert e64[k] <= TOL64, (k, e64[k])


@pytest.mark.parametrize("name", cases.CASES_1D)
def test_emu_parity_1d(name):
    _check(cases.case(name, 1))


@pytest.mark.parametrize("name", cases.CASES_2D)
def test_emu_parity_2d(name):
    _check(cases.case(name, 2))


@pytest.mark.parametrize("name", cases.CASES_3D)
def test_emu_parity_3d(name):
    _check(cases.case(name, 3))


@pytest.mark.parametrize("dim,name,shape", [
    (2, "p0_hetero_absorbing", (64, 30)),
    (2, "psrc_many", (30, 64)),
    (2, "time_reversal", (128, 24)),
    (2, "stokes", (16, 256)),
    (2, "psrc_label", (384, 16)),
    (3, "p0_nonlinear_absorbing", (12, 10, 64)),
    (3, "usrc_dir", (128, 16, 16)),
    (3, "no_dispersion", (8, 8, 256)),
    (3, "psrc_dir", (16, 384, 16)),
    # lengths that are not multiples of 8 elements per thread, odd lengths,
    # and the long lengths built without strided-pass variants
    (2, "p0_hetero_absorbing", (75, 30)),
    (3, "psrc_add", (16, 81, 16)),
    (2, "time_reversal", (135, 24)),
    (3, "p0_nonlinear_absorbing", (12, 10, 135)),
    (2, "psrc_label", (270, 16)),
    (2, "stokes", (16, 270)),
    # 1-D grids on fixed-length x kernels
    (1, "p0_hetero_absorbing", (75,)),
    (1, "psrc_many", (128,)),
    (1, "time_reversal", (270,)),
    # lengths with a factor 7 (radix-7 stages of the runtime engine)
    (1, "p0_hetero_absorbing", (98,)),
    (2, "time_reversal", (49, 30)),
    (3, "p0_nonlinear_absorbing", (14, 21, 28)),
    # the velocity split's last-axis pass on fixed-length kernels
    (2, "u_split", (30, 64)),
    (3, "u_split", (12, 16, 64)),
    # lengths with a prime factor above 7 (Bluestein's chirp-z transform on
    # every axis kind): 22 = 2 x 11, 26 = 2 x 13, 34 = 2 x 17, 31 prime
    (1, "p0_hetero_absorbing", (22,)),
    (1, "time_reversal", (31,)),
    (2, "p0_hetero_absorbing", (22, 26)),
    (2, "psrc_add", (34, 20)),
    (2, "time_reversal", (26, 31)),
    (3, "p0_nonlinear_absorbing", (22, 26, 17)),
    (3, "psrc_dir", (13, 20, 34)),
    (3, "u_split", (22, 13, 16)),
    # the longest x axes: 512 threads per pencil
    (2, "p0_lossless", (4096, 10)),
    (2, "p0_lossless", (2039, 10)),
])
def test_emu_parity_fixed_lengths(dim, name, shape):
    _check(cases.case(name, dim, shape))


def test_emu_generic_plans_match_fixed(monkeypatch):
    """The runtime-plan kernels and the fixed-length kernels agree."""
    cs = cases.case("p0_hetero_absorbing", 3, (64, 12, 10))
    a = cases.run_kwhip(cs, "emu")
    monkeypatch.setenv("KW_NO_FIXED", "1")
    b = cases.run_kwhip(cs, "emu")
    for k in a:
        ref = np.max(np.abs(a[k]))
        assert np.max(np.abs(np.asarray(a[k]) - np.asarray(b[k]))) <= 1e-5 * ref
