import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT, ROOT / "k-wave-ii_amd", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))
os.environ.setdefault("HSA_ENABLE_IPC_MODE_LEGACY", "0")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long CPU test")


def pytest_sessionstart(session):
    """Tests load the libraries that are built; only when none is built is
    the build run first."""
    lib = ROOT / "k-wave-ii_amd" / "kwhip" / "_lib"
    if not (lib / "libkwhip.so").exists() or not (lib / "libkwhip_emu.so").exists():
        import __graft_entry__

        __graft_entry__.build()


@pytest.fixture(scope="session")
def hip_lib():
    """The product library on the GPU; fails loudly if it is missing."""
    import kwhip

    lib = kwhip.load("hip")
    assert lib.kw_backend().decode() == "hip-gfx950"
    return lib
