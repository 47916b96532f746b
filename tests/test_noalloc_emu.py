"""kw_run_noalloc on the CPU lane emulator (no GPU): the simulation's
buffers come from the library's scratch block, the outputs equal kw_run's
bit for bit, a second, smaller call after a first one still does, and a
problem larger than the scratch block is refused with KW_ENOMEM."""

import numpy as np
import pytest

import cases


@pytest.fixture(scope="module")
def emu_lib():
    from kwhip import _abi
    return _abi.load("emu")


def test_noalloc_equals_kw_run_on_emulator(emu_lib):
    for cs in (cases.case("psrc_many", 2), cases.case("p0_lossless", 2, (24, 20))):
        pp = cases.prepared(cs)
        want = cases.run_entry(emu_lib, "kw_run", pp)
        got = cases.run_entry(emu_lib, "kw_run_noalloc", pp)
        assert set(got) == set(want)
        for k in want:
            assert np.array_equal(np.asarray(got[k]), np.asarray(want[k])), (cs["shape"], k)


def test_noalloc_refuses_oversize_on_emulator(emu_lib):
    from kwhip import _abi
    N = 192
    cs = {"dim": 3, "shape": (N, N, N), "dx": 1e-4, "Nt": 2, "dt": 2e-8,
          "medium": {"sound_speed": 1500.0, "density": 1000.0},
          "source": {"p0": np.zeros((N, N, N))},
          "sensor": {"mask": np.ones((N, N, N)), "record": ["p_final"]},
          "opts": {"PMLSize": 10, "Smooth": False}}
    pp = cases.prepared(cs)
    with pytest.raises(_abi.KWError) as ei:
        cases.run_entry(emu_lib, "kw_run_noalloc", pp)
    assert ei.value.code == _abi.KW_ENOMEM
    assert "scratch" in str(ei.value)
