"""GPU parity over the full length of the benchmark recipes, against oracle
runs committed as fixtures (tests/golden/make_longrun.py writes them; the
oracle needs minutes to tens of minutes per run).

* The 256^3 benchmark recipe at 64^3 and 128^3 for all 1000 steps: the
  additive tone burst is on for the first ~530 steps and off afterwards
  (kspaceFirstOrder3D.m:930: flags.source_p >= t_index), so both regimes
  and the FP32 drift over the whole run are covered.  p_max, p_rms and
  p_final on a subsampled grid: relative L-inf <= 1e-5 against the oracle in
  single precision, <= 1e-4 against double precision, normalised by the
  largest magnitude over the whole grid; for p_final that is the run's peak
  pressure (the largest p_max).  After 1000 steps the wave has left the grid
  and p_final is about 3e-6 of the peak, at the single-precision rounding
  floor of a field that carried the peak: the oracle's own single- and
  double-precision runs differ by 8e-4 of p_final's own largest value, so
  that error is printed but not gated.
* The nonlinear precision-study recipe (B/A = 6, 1 MPa bowl source) at
  96^3 and 192^3 for 2000 steps, sensor p on the 64 x 64 focal plane (the
  192^3 oracle runs took 44 min in single and 84 min in double precision;
  the full 384^3 recipe is out of the oracle's reach): p_max and p_rms
  at every plane point and the p series on a subsampled set of points and
  steps, relative L-inf and relative L2 stated against both oracles.
"""

import sys
import zlib
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))

TOL32 = 1e-5
TOL64 = 1e-4


def _fixture(name):
    path = GOLDEN / name
    if not path.exists():
        pytest.fail(f"missing fixture {path}: run tests/golden/make_longrun.py")
    return np.load(path)


@pytest.mark.parametrize("N", [64, 128])
def test_disc_source_recipe_full_length(N, hip_lib):
    import kwhip
    import make_longrun as ml

    fx = _fixture(f"disc_source_long_{N}.npz")
    kg, med, src, sens, opts = ml.disc_inputs(N)
    assert int(fx["crc"]) == ml.input_crc(med, src)
    n_on = src["p"].shape[1]
    print(f"[longrun] disc-source recipe {N}^3, {kg.Nt} steps, source on for {n_on}", flush=True)
    assert n_on < kg.Nt
    got = kwhip.kspaceFirstOrder3D(kg, med, src, sens, **opts)
    s = int(fx["stride"])
    for tag, tol in (("fp32", TOL32), ("fp64", TOL64)):
        for k in ("p_max", "p_rms", "p_final"):
            g = ml.as_grid(got[k], N)[::s, ::s, ::s]
            r = fx[f"{tag}_{k}"].astype(np.float64)
            own = np.max(np.abs(g - r)) / float(fx[f"{tag}_{k}_absmax"])
            scale = "p_max" if k == "p_final" else k
            err = np.max(np.abs(g - r)) / float(fx[f"{tag}_{scale}_absmax"])
            print(f"[longrun] {N}^3 {k} vs {tag} oracle: rel L-inf {err:.2e} "
                  f"(of its own largest value {own:.2e})", flush=True)
            assert err <= tol, (N, tag, k, err)


@pytest.mark.parametrize("N", [96, 192])
def test_nonlinear_bowl_recipe(N, hip_lib):
    import kwhip
    import make_longrun as ml

    fx = _fixture(f"nonlinear_bowl_{N}.npz")
    kg, med, src, sens, opts = ml.bowl_inputs(N)
    assert int(fx["crc"]) == ml.input_crc(med, src)
    print(f"[longrun] nonlinear bowl recipe {N}^3, {kg.Nt} steps, bowl of {int(src['p_mask'].sum())} points",
          flush=True)
    got = kwhip.kspaceFirstOrder3D(kg, med, src, sens, **opts)
    gp = ml.bowl_p_sub(got["p"]).astype(np.float64)
    for tag, tol in (("fp32", TOL32), ("fp64", TOL64)):
        r = fx[f"{tag}_p"].astype(np.float64)
        linf = np.max(np.abs(gp - r)) / float(fx[f"{tag}_p_absmax"])
        l2 = np.sqrt(np.sum((gp - r) ** 2) / np.sum(r ** 2))
        print(f"[longrun] nonlinear bowl {N}^3 p vs {tag} oracle: rel L-inf {linf:.2e}, rel L2 {l2:.2e}", flush=True)
        assert linf <= tol, (N, tag, "p", linf)
        for k in ("p_max", "p_rms"):
            g = np.asarray(got[k], dtype=np.float64).ravel()
            r = fx[f"{tag}_{k}"].astype(np.float64)
            err = np.max(np.abs(g - r)) / np.max(np.abs(r))
            print(f"[longrun] nonlinear bowl {N}^3 {k} vs {tag} oracle: rel L-inf {err:.2e}", flush=True)
            assert err <= tol, (N, tag, k, err)
