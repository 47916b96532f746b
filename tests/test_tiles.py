"""Launch tiles (pencils per block of the x, strided y and last-axis pass
families) and the tile autotuner (kw_tiles in include/kwhip.h), on the CPU
lane emulator.

* Tiles change which block handles a pencil, never the arithmetic on it:
  outputs are bit-identical under different tiles.
* The autotuner (forced on the emulator with KW_TUNE=1) returns tiles that
  the launch limits allow, writes its choice to KW_TUNE_FILE, and a later
  process takes the file's tiles instead of measuring.
* Uncapped, the autotuner steps every candidate up to 64 pencils per block
  with the emulator's bounds check on every global access: each candidate
  runs in a simulation created with its own tiles (the spectral tables are
  laid out by the last-axis tile at creation, so re-planning a live
  simulation would read past them).
"""

import os
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import cases

ROOT = Path(__file__).resolve().parents[1]


def _run_with(monkeypatch, cs, tiles):
    for k, v in zip(("KW_PENCILS_X", "KW_PENCILS_Y", "KW_PENCILS_L"), tiles):
        if v is None:
            monkeypatch.delenv(k, raising=False)
        else:
            monkeypatch.setenv(k, str(v))
    return cases.run_kwhip(cs, "emu")


@pytest.mark.parametrize("name,dim", [("p0_hetero_absorbing", 2), ("psrc_add", 3)])
def test_tiles_do_not_change_results(monkeypatch, name, dim):
    monkeypatch.setenv("KW_MAX_PENCILS", "4")
    monkeypatch.setenv("KW_TUNE", "0")
    cs = cases.case(name, dim=dim)
    ref = _run_with(monkeypatch, cs, (None, None, None))
    for tiles in ((1, 1, 1), (3, 2, 3)):
        got = _run_with(monkeypatch, cs, tiles)
        assert set(got) == set(ref)
        for k in ref:
            assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), (tiles, k)


_CHILD = r"""
import sys
sys.path[:0] = [{root!r}, {pkg!r}]
import numpy as np
import kwhip
from kwhip import _abi
shape = {shape!r}
kg = kwhip.kWaveGrid(*sum(([n, 1e-4] for n in shape), []))
kg.setTime(8, 2e-8)
p0 = np.zeros(shape); p0[tuple(n // 2 for n in shape)] = 1.0
with _abi._backend_override("emu"):
    t = kwhip.launch_tiles(kg, {{"sound_speed": 1500.0, "density": 1000.0}}, {{"p0": p0}},
                           {{"mask": np.ones(shape, dtype=bool)}}, len(shape), PMLSize=2,
                           Smooth=False)
print(" ".join(map(str, t)))
"""


def _tiles_in_child(shape, env):
    code = _CHILD.format(root=str(ROOT), pkg=str(ROOT / "k-wave-ii_amd"), shape=shape)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    return tuple(int(v) for v in r.stdout.split())


@pytest.mark.parametrize("shape", [(24, 20), (12, 10, 8)])
def test_autotuner_bounds_and_file(tmp_path, shape):
    path = tmp_path / "tiles.txt"
    env = dict(os.environ, KW_TUNE="1", KW_TUNE_FILE=str(path), KW_MAX_PENCILS="4")
    for k in ("KW_PENCILS_X", "KW_PENCILS_Y", "KW_PENCILS_L"):
        env.pop(k, None)
    tx, ty, tl = _tiles_in_child(shape, env)
    # 8 elements per thread: ceil(N / 8) threads per pencil; x blocks <= 512
    # threads, y and last-axis blocks <= 1024; KW_MAX_PENCILS caps all three
    gx = -(-shape[0] // 8)
    assert 1 <= tx <= min(4, 512 // gx)
    assert 1 <= tl <= 4
    if len(shape) == 3:
        assert 1 <= ty <= 4
    else:
        assert ty == 0
    lines = path.read_text().split("\n")
    rows = [list(map(int, ln.split())) for ln in lines if ln.strip()]
    key = [len(shape), shape[0], shape[1], shape[2] if len(shape) == 3 else 1, 0, 0]
    assert [r[:6] for r in rows] == [key]
    assert rows[0][6] == tx and rows[0][8] == tl
    # a later process reads the file instead of measuring
    path.write_text(" ".join(map(str, key + [1, 2 if len(shape) == 3 else 1, 3])) + "\n")
    assert _tiles_in_child(shape, env) == (1, 2 if len(shape) == 3 else 0, 3)


@pytest.mark.parametrize("shape", [(12, 10, 8), (24, 20)])
def test_autotuner_uncapped_candidates_stay_in_bounds(shape):
    env = dict(os.environ, KW_TUNE="1")
    for k in ("KW_PENCILS_X", "KW_PENCILS_Y", "KW_PENCILS_L", "KW_MAX_PENCILS", "KW_TUNE_FILE"):
        env.pop(k, None)
    tx, ty, tl = _tiles_in_child(shape, env)
    assert 1 <= tx <= 64 and 1 <= tl <= 64
    assert (ty == 0) == (len(shape) == 2) and ty <= 64
