"""The kspaceFirstOrder-HIP command line on the GPU: an input file written
the way the reference's MATLAB side writes it (chunked and deflated,
tests/golden/h5lite.py) is run with the flags kspaceFirstOrder3DC.m passes,
and the output file's records are checked against the oracle in single
precision (relative L-inf <= 1e-5) and double precision (<= 1e-4)."""

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import h5lite  # noqa: E402

TOL32, TOL64 = 1e-5, 1e-4


@pytest.mark.parametrize("name,dim,shape", [("psrc_add", 3, (32, 24, 20)),
                                            ("p0_hetero_absorbing", 2, (64, 48))])
def test_command_line_run(hip_lib, tmp_path, name, dim, shape):
    from kwhip import _abi
    cs = cases.case(name, dim, shape)
    record = ["p", "p_max", "p_rms", "p_final"]
    cs["sensor"] = dict(cs["sensor"], record=record)
    print(f"[cli] {name} {shape}", flush=True)
    pp = cases.prepared(cs)
    _abi.call(hip_lib.kw_save_problem, C.byref(pp.prob), str(tmp_path / "in.h5").encode())
    v = h5lite.read(tmp_path / "in.h5")
    variables = {k: (a, "uint64" if a.dtype == np.uint64 else "single") for k, a in v.items()}
    h5lite.write_input(tmp_path / "matlab.h5", variables, h5lite.read_attrs(tmp_path / "in.h5"),
                       compression=4)
    exe = _abi.LIB_DIR / "kspaceFirstOrder-HIP"
    r = subprocess.run([str(exe), "-i", str(tmp_path / "matlab.h5"), "-o", str(tmp_path / "out.h5"),
                        "--p_raw", "--p_max", "--p_rms", "--p_final"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = h5lite.read(tmp_path / "out.h5")
    for cast, tol in (("single", TOL32), ("off", TOL64)):
        ref = cases.run_oracle(cs, cast)
        for k in record:
            got = np.squeeze(out[k]).astype(np.float64)
            want = np.asarray(ref[k], dtype=np.float64)
            if k == "p_final":
                want = want.reshape(shape)
            want = np.squeeze(want)
            assert got.shape == want.shape, (k, got.shape, want.shape)
            err = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
            print(f"[cli] {name} {k} vs oracle ({cast}): {err:.2e}", flush=True)
            assert err <= tol, (k, cast, err)
