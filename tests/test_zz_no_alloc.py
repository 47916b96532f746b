"""kw_run_noalloc (the entry point that allocates no device memory) against
kw_run through the C ABI: the same kernels run on buffers carved out of the
library's static scratch array instead of hipMalloc'd ones, so the outputs
match kw_run's.  Two calls in a row, the second on a smaller problem (the
array keeps the first call's contents), and a problem larger than the
scratch array, which must be refused with KW_ENOMEM."""

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

TOL = 1e-6  # same kernels and operators: the outputs agree to rounding


def test_noalloc_matches_kw_run(hip_lib):
    big = cases.case("p0_nonlinear_absorbing", 3, (32, 32, 32))
    small = cases.case("psrc_many", 2)
    for label, cs in (("3-D 32^3 nonlinear absorbing p0", big), ("2-D psrc_many", small)):
        print(f"[noalloc] {label}: shape {cs['shape']}, Nt {cs['Nt']}", flush=True)
        pp = cases.prepared(cs)
        want = cases.run_entry(hip_lib, "kw_run", pp)
        got = cases.run_entry(hip_lib, "kw_run_noalloc", pp)
        assert set(got) == set(want)
        for k in want:
            g = np.asarray(got[k], dtype=np.float64)
            w = np.asarray(want[k], dtype=np.float64)
            assert g.shape == w.shape, (label, k)
            err = np.max(np.abs(g - w)) / max(np.max(np.abs(w)), 1e-300)
            print(f"[noalloc] {label} {k}: rel L-inf vs kw_run {err:.2e}", flush=True)
            assert err <= TOL, (label, k, err)


def test_noalloc_refuses_oversize(hip_lib):
    from kwhip import _abi
    N = 192
    cs = {"dim": 3, "shape": (N, N, N), "dx": 1e-4, "Nt": 2, "dt": 2e-8,
          "medium": {"sound_speed": 1500.0, "density": 1000.0},
          "source": {"p0": np.zeros((N, N, N))},
          "sensor": {"mask": np.ones((N, N, N)), "record": ["p_final"]},
          "opts": {"PMLSize": 10, "Smooth": False}}
    print("[noalloc] 3-D 192^3 (needs more than the scratch array)", flush=True)
    pp = cases.prepared(cs)
    with pytest.raises(_abi.KWError) as ei:
        cases.run_entry(hip_lib, "kw_run_noalloc", pp)
    assert ei.value.code == _abi.KW_ENOMEM
    assert "scratch" in str(ei.value)
