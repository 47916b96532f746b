"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the oracle in double precision): the oracle still reproduces them, and
the emulator build and the HIP library match them at the double-precision
gate (relative L-inf <= 1e-4)."""

from pathlib import Path

import numpy as np
import pytest

import cases

GOLDEN_DIR = Path(__file__).resolve().parent / "golden"
import sys  # noqa: E402

sys.path.insert(0, str(GOLDEN_DIR))
import make_golden  # noqa: E402

TOL64 = 1e-4
IDS = [make_golden.fixture_name(*g) for g in make_golden.GOLDEN]


def _load(dim, name, shape):
    path = GOLDEN_DIR / make_golden.fixture_name(dim, name, shape)
    z = np.load(path, allow_pickle=False)
    cs = cases.case(name, dim, shape)
    assert int(z["crc"]) == make_golden.input_crc(cs), "case generator changed; regenerate fixtures"
    assert tuple(z["shape"]) == tuple(cs["shape"])
    return cs, z


def _errors(out, z):
    if not isinstance(out, dict):
        out = {"p": out}
    keys = sorted(k[4:] for k in z.files if k.startswith("out_"))
    assert sorted(out) == keys
    errs = {}
    for k in keys:
        want = z["out_" + k].astype(np.float64)
        got = np.asarray(out[k], dtype=np.float64).ravel(order="F")[::int(z["stride_" + k])]
        errs[k] = np.max(np.abs(got - want)) / np.max(np.abs(want))
    return errs


@pytest.mark.parametrize("g", make_golden.GOLDEN, ids=IDS)
def test_oracle_reproduces_golden(g):
    cs, z = _load(*g)
    errs = _errors(cases.run_oracle(cs, "off"), z)
    assert max(errs.values()) < 1e-6, errs


@pytest.mark.parametrize("g", make_golden.GOLDEN, ids=IDS)
def test_emulator_matches_golden(g, monkeypatch):
    monkeypatch.setenv("KW_MAX_PENCILS", "4")
    cs, z = _load(*g)
    errs = _errors(cases.run_kwhip(cs, "emu"), z)
    assert max(errs.values()) <= TOL64, errs


@pytest.mark.gpu
@pytest.mark.parametrize("g", make_golden.GOLDEN, ids=IDS)
def test_hip_matches_golden(g, hip_lib):
    cs, z = _load(*g)
    print(f"[golden] {IDS[make_golden.GOLDEN.index(g)]}", flush=True)
    errs = _errors(cases.run_kwhip(cs, "hip"), z)
    assert max(errs.values()) <= TOL64, errs
