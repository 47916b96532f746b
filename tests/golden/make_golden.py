"""Writes the golden fixtures of this directory: seeded parity cases run
through the oracle in double precision ('DataCast' off).

The reference (MATLAB) commits no golden vectors, so these fixtures pin the
oracle's outputs as data: tests/test_golden.py checks that the oracle still
reproduces them and that the emulator and the HIP library match them.  Each
fixture stores the case name, dimension, grid shape, a CRC of the case's
inputs (so a changed case generator is detected) and the recorded outputs,
rounded to float32; outputs above MAX_VALUES values keep every k-th entry
of their flattened (column-major) form, with k stored as ``stride_<name>``,
to keep the files small.

Run from the repository root:  python tests/golden/make_golden.py
(--missing writes only the fixtures that do not exist yet)
"""

import sys
import zlib
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path[:0] = [str(ROOT), str(ROOT / "tests")]

import cases  # noqa: E402

GOLDEN = [
    (2, "p0_hetero_absorbing", None),
    (2, "p0_nonlinear_absorbing", None),
    (2, "psrc_add", None),
    (2, "time_reversal", None),
    (3, "p0_lossless", None),
    (3, "psrc_label", None),
    (3, "usrc_many", None),
    (3, "stokes", (16, 12, 20)),
    (1, "p0_hetero_absorbing", None),
    (1, "psrc_add", None),
    (1, "time_reversal", None),
    (2, "directivity_pressure", None),
    (2, "cartesian_linear", None),
    (3, "u_split", None),
]


MAX_VALUES = 4096


def input_crc(cs):
    crc = 0
    for part in ("medium", "source", "sensor"):
        for k in sorted(cs[part]):
            v = cs[part][k]
            if isinstance(v, (list, tuple)) and v and isinstance(v[0], str):
                b = ",".join(v).encode()
            elif isinstance(v, str):
                b = v.encode()
            else:
                b = np.ascontiguousarray(np.asarray(v, dtype=np.float64)).tobytes()
            crc = zlib.crc32(k.encode() + b, crc)
    return crc


def fixture_name(dim, name, shape):
    tag = "" if shape is None else "_" + "x".join(map(str, shape))
    return f"{name}_{dim}d{tag}.npz"


def main():
    missing_only = "--missing" in sys.argv[1:]
    for dim, name, shape in GOLDEN:
        if missing_only and (HERE / fixture_name(dim, name, shape)).exists():
            continue
        cs = cases.case(name, dim, shape)
        out = cases.run_oracle(cs, "off")
        if not isinstance(out, dict):
            out = {"p": out}
        arrays = {}
        for k, v in out.items():
            flat = np.asarray(v, dtype=np.float32).ravel(order="F")
            stride = -(-flat.size // MAX_VALUES)
            arrays[f"out_{k}"] = flat[::stride]
            arrays[f"stride_{k}"] = np.int64(stride)
        np.savez_compressed(HERE / fixture_name(dim, name, shape), name=name, dim=dim,
                            shape=np.asarray(cs["shape"]), crc=np.uint32(input_crc(cs)), **arrays)
        print(fixture_name(dim, name, shape), {k: v.shape for k, v in arrays.items()})


if __name__ == "__main__":
    main()
