"""The tile autotuner on the GPU: for a small 2-D and 3-D grid (where the
built-in tiles leave most of the device idle) kw_tiles measures and returns
tiles that the launch limits allow, and a simulation under the tuned tiles
gives bit-identical outputs to one under the built-in tiles (KW_TUNE=0);
the step times under both are printed."""

import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(shape, Nt=300):
    import kwhip
    kg = kwhip.kWaveGrid(*sum(([n, 1e-4] for n in shape), []))
    kg.setTime(Nt, 2e-8)
    p0 = np.zeros(shape)
    p0[tuple(n // 2 for n in shape)] = 1.0
    mask = np.zeros(shape, dtype=bool)
    mask[4] = True
    med = {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 0.75, "alpha_power": 1.5}
    return kg, med, {"p0": p0}, {"mask": mask, "record": ["p", "p_max", "p_final"]}


@pytest.mark.parametrize("shape", [(256, 256), (64, 64, 64)])
def test_tuned_tiles_bounds_and_identical_results(monkeypatch, hip_lib, shape):
    import kwhip
    kg, med, src, sens = _problem(shape)
    dim = len(shape)
    fn = kwhip.kspaceFirstOrder3D if dim == 3 else kwhip.kspaceFirstOrder2D
    opts = {"PMLSize": 10, "Smooth": False}
    for k in ("KW_PENCILS_X", "KW_PENCILS_Y", "KW_PENCILS_L", "KW_MAX_PENCILS", "KW_TUNE_FILE"):
        monkeypatch.delenv(k, raising=False)

    monkeypatch.setenv("KW_TUNE", "0")
    rule = kwhip.launch_tiles(kg, med, src, sens, dim, **opts)
    fn(kg, med, src, sens, **opts)
    t0 = time.perf_counter()
    ref = fn(kg, med, src, sens, **opts)
    t_rule = time.perf_counter() - t0

    monkeypatch.delenv("KW_TUNE")
    tiles = kwhip.launch_tiles(kg, med, src, sens, dim, **opts)
    fn(kg, med, src, sens, **opts)
    t0 = time.perf_counter()
    got = fn(kg, med, src, sens, **opts)
    t_tuned = time.perf_counter() - t0
    print(f"[tiles] {shape}: built-in {rule} {t_rule * 1e3:.1f} ms, tuned {tiles} "
          f"{t_tuned * 1e3:.1f} ms (300 steps incl. host work)", flush=True)

    gx = -(-shape[0] // 8)
    gl = -(-shape[-1] // 8)
    assert 1 <= tiles[0] <= 512 // gx
    assert 1 <= tiles[2] <= 1024 // gl
    assert (tiles[1] == 0) == (dim == 2)
    for k in ref:
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), k
