"""GPU parity: the HIP library against the CPU oracle on the same inputs.

Gate 1: relative L-inf <= 1e-5 against the oracle in single precision (the
reference's 'DataCast','single').  Gate 2: <= 1e-4 against the oracle in
double precision.  Every output field of each case is checked.
"""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-4


def _check(cs, hip_lib):
    print(f"[case] dim={cs['dim']} shape={cs['shape']}", flush=True)
    got = cases.run_kwhip(cs, "hip")
    e32 = cases.compare(got, cases.run_oracle(cs, "single"))
    e64 = cases.compare(got, cases.run_oracle(cs, "off"))
    for k in e32:
        assert e32[k] <= TOL32, (k, e32[k])
        assert e64[k] <= TOL64, (k, e64[k])


@pytest.mark.parametrize("name", cases.CASES_1D)
def test_parity_1d(name, hip_lib):
    _check(cases.case(name, 1), hip_lib)


@pytest.mark.parametrize("name,shape", [("p0_hetero_absorbing", (75,)),
                                        ("psrc_many", (128,)),
                                        ("time_reversal", (1024,)),
                                        ("usrc_add", (2000,))])
def test_parity_1d_sizes(name, shape, hip_lib):
    _check(cases.case(name, 1, shape), hip_lib)


@pytest.mark.parametrize("name", cases.CASES_2D)
def test_parity_2d(name, hip_lib):
    _check(cases.case(name, 2), hip_lib)


@pytest.mark.parametrize("name", cases.CASES_3D)
def test_parity_3d(name, hip_lib):
    _check(cases.case(name, 3), hip_lib)


@pytest.mark.parametrize("name,shape", [("p0_hetero_absorbing", (27, 25)),
                                        ("psrc_many", (45, 50)),
                                        ("time_reversal", (64, 48)),
                                        ("p0_hetero_absorbing", (56, 42)),
                                        ("u_split", (28, 35)),
                                        ("psrc_add", (112, 98))])
def test_parity_2d_sizes(name, shape, hip_lib):
    _check(cases.case(name, 2, shape), hip_lib)


@pytest.mark.parametrize("name,shape", [("p0_hetero_absorbing", (15, 9, 25)),
                                        ("psrc_add", (24, 30, 32)),
                                        ("p0_nonlinear_absorbing", (32, 32, 32)),
                                        ("p0_nonlinear_absorbing", (14, 21, 28)),
                                        ("psrc_dir", (21, 16, 14))])
def test_parity_3d_sizes(name, shape, hip_lib):
    _check(cases.case(name, 3, shape), hip_lib)


@pytest.mark.parametrize("dim,name,shape", [
    # lengths with a prime factor above 7: Bluestein's chirp-z transform
    (1, "p0_hetero_absorbing", (22,)),
    (2, "p0_hetero_absorbing", (22, 26)),
    (2, "time_reversal", (34, 31)),
    (3, "p0_nonlinear_absorbing", (22, 26, 34)),
    (3, "psrc_dir", (13, 20, 34)),
    (3, "u_split", (22, 13, 16)),
    # long x axes: 4096 (512 threads per pencil) and the prime 2039
    # (chirp-z length 4096)
    (2, "p0_hetero_absorbing", (4096, 64)),
    (2, "psrc_add", (2039, 24)),
])
def test_parity_any_length(dim, name, shape, hip_lib):
    _check(cases.case(name, dim, shape), hip_lib)
