"""The native engines' file contract (kw_save_problem, kw_run_file and the
kspaceFirstOrder-HIP command line) on the CPU lane emulator, checked with an
independent HDF5 reader and writer (tests/golden/h5lite.py):

* the input file holds the variable set, names, types, MATLAB sizes and
  attributes of kspaceFirstOrder_saveToDisk.m / writeMatrix.m /
  writeAttributes.m, and the values of the prepared problem;
* running that file gives the outputs of the same problem run through
  kw_run, under the record names kspaceFirstOrder3DC.m:571-777 reads, for
  binary and cuboid-corner sensors;
* the same variables re-written the way MATLAB writes them (chunked
  storage, a version-1 B-tree chunk index, deflate) give the same outputs;
* the 'SaveToDisk' / 'SaveToDiskExit' options of kspaceFirstOrder2D/3D.
"""

import ctypes as C
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import cases

GOLDEN = Path(__file__).resolve().parent / "golden"
sys.path.insert(0, str(GOLDEN))
import h5lite  # noqa: E402

TOL = 1e-5  # relative L-inf, the single-precision parity bar

REC_NAMES = ["p", "p_max", "p_min", "p_rms", "p_max_all", "p_min_all", "p_final", "u", "u_max",
             "u_min", "u_rms", "u_max_all", "u_min_all", "u_final", "u_non_staggered"]


@pytest.fixture(scope="module")
def emu():
    from kwhip import _abi
    return _abi.load("emu")


def _save(lib, cs, path):
    from kwhip import _abi
    pp = cases.prepared(cs)
    _abi.call(lib.kw_save_problem, C.byref(pp.prob), str(path).encode())
    return pp


def _run_file(lib, inp, out, record, start=1):
    from kwhip import _abi
    flags = 0
    for r in record:
        flags |= _abi.REC[r]
    _abi.call(lib.kw_run_file, str(inp).encode(), str(out).encode(), C.c_uint32(flags),
              C.c_int64(start))


def test_input_file_variables(emu, tmp_path):
    cs = cases.case("psrc_many", 2)
    pp = _save(emu, cs, tmp_path / "in.h5")
    v = h5lite.read(tmp_path / "in.h5")
    Nx, Ny = cs["shape"]
    # uint64 sizes and flags, as MATLAB [1 1 1] arrays
    for nm, val in (("Nx", Nx), ("Ny", Ny), ("Nz", 1), ("Nt", cs["Nt"]), ("p0_source_flag", 0),
                    ("sensor_mask_type", 0), ("absorbing_flag", 0), ("elastic_flag", 0),
                    ("p_source_mode", 2), ("p_source_many", 1), ("pml_z_size", 0)):
        assert v[nm].dtype == np.uint64 and v[nm].shape == (1, 1, 1), nm
        assert int(v[nm].ravel()[0]) == val, nm
    src = np.atleast_2d(cs["source"]["p"])
    assert int(v["p_source_flag"].ravel()[0]) == src.shape[1]
    # 1-based column-major indices of the mask, as a column
    idx = np.flatnonzero(np.asarray(cs["source"]["p_mask"]).ravel(order="F")) + 1
    assert v["p_source_index"].shape == (idx.size, 1, 1)
    assert np.array_equal(v["p_source_index"].ravel(), idx)
    # scaled signals [points x time] in single precision: the prepared ones
    sig = np.ctypeslib.as_array(pp.prob.p_source.input, shape=(src.shape[1] * idx.size,))
    assert v["p_source_input"].dtype == np.float32 and v["p_source_input"].shape == (idx.size, src.shape[1], 1)
    assert np.array_equal(v["p_source_input"].ravel(order="F"), sig.astype(np.float32))
    # reals: no 'z' variables in 2-D
    for nm in ("dz", "pml_z_alpha", "rho0_sgz"):
        assert nm not in v
    assert np.isclose(float(v["dt"].ravel()[0]), np.float32(cs["dt"]))
    assert np.isclose(float(v["c_ref"].ravel()[0]), np.float32(pp.prob.c_ref))
    # attributes (writeMatrix.m, writeAttributes.m, getH5Literals.m)
    root = h5lite.read_attrs(tmp_path / "in.h5")
    assert root["file_type"] == "input" and root["major_version"] == "1" and root["minor_version"] == "2"
    assert h5lite.read_attrs(tmp_path / "in.h5", "Nx") == {"domain_type": "real", "data_type": "long"}
    assert h5lite.read_attrs(tmp_path / "in.h5", "dt") == {"domain_type": "real", "data_type": "float"}


CONTRACT_CASES = [
    ("p0_hetero_absorbing", 2, None, ["p", "p_max", "p_min", "p_rms", "p_final", "p_max_all"]),
    ("psrc_add", 3, None, ["p", "p_rms", "p_final"]),
    ("usrc_add", 2, None, ["p", "u", "u_max", "u_final", "u_non_staggered"]),
    ("p0_nonlinear_absorbing", 2, (27, 32), ["p_max", "p_min_all"]),
    ("stokes", 2, None, ["p", "p_final"]),
]


def _expected(cs, record, backend="emu"):
    cs = dict(cs)
    cs["sensor"] = dict(cs["sensor"], record=record)
    return cases.run_kwhip(cs, backend)


def _compare_outputs(out, want, record, dim, shape):
    comps = ["x", "y", "z"][:dim]
    for r in record:
        if r.startswith("u"):
            names = [("u" + c + r[1:]) for c in comps]
        else:
            names = [r]
        for nm in names:
            got = np.squeeze(out[nm])
            w = np.asarray(want[nm])
            if nm.endswith(("_final", "_all")):
                w = w.reshape(shape)
            w32 = np.squeeze(w.astype(np.float32))
            assert got.shape == w32.shape, (nm, got.shape, w32.shape)
            # the file holds the medium, dt and sources in single precision
            # (the contract's 'single' reals), so the operators differ from
            # those built from the double inputs by single-precision rounding
            err = float(np.max(np.abs(got.astype(np.float64) - w32)) / max(np.max(np.abs(w32)), 1e-30))
            assert err <= TOL, (nm, err)


@pytest.mark.parametrize("name,dim,shape,record", CONTRACT_CASES)
def test_run_file_equals_kw_run(emu, tmp_path, name, dim, shape, record):
    cs = cases.case(name, dim, shape)
    _save(emu, cs, tmp_path / "in.h5")
    # the first recorded step goes over as the -s option, not in the file
    _run_file(emu, tmp_path / "in.h5", tmp_path / "out.h5", record,
              int(cs["sensor"].get("record_start_index", 1)))
    out = h5lite.read(tmp_path / "out.h5")
    want = _expected(cs, record)
    _compare_outputs(out, want, record, dim, cs["shape"])
    assert h5lite.read_attrs(tmp_path / "out.h5")["file_type"] == "output"
    n = list(cs["shape"]) + [1] * (3 - dim)
    assert [int(out[k].ravel()[0]) for k in ("Nx", "Ny", "Nz")] == n


@pytest.mark.parametrize("compression", [0, 4])
def test_matlab_style_input(emu, tmp_path, compression):
    """The variables re-written as MATLAB writes them (chunked, deflated)
    run to the same outputs."""
    cs = cases.case("p0_hetero_absorbing", 3, (16, 12, 10))
    record = ["p", "p_max", "p_final"]
    _save(emu, cs, tmp_path / "in.h5")
    v = h5lite.read(tmp_path / "in.h5")
    variables = {k: (a, "uint64" if a.dtype == np.uint64 else "single") for k, a in v.items()}
    attrs = h5lite.read_attrs(tmp_path / "in.h5")
    h5lite.write_input(tmp_path / "matlab.h5", variables, attrs, compression=compression)
    # the chunked layout really is what the file holds (writeMatrix.m shapes)
    w = h5lite.read(tmp_path / "matlab.h5")
    assert all(np.array_equal(w[k], v[k]) for k in v)
    _run_file(emu, tmp_path / "in.h5", tmp_path / "a.h5", record)
    _run_file(emu, tmp_path / "matlab.h5", tmp_path / "b.h5", record)
    a, b = h5lite.read(tmp_path / "a.h5"), h5lite.read(tmp_path / "b.h5")
    for k in record:
        assert np.array_equal(a[k], b[k]), k
    _compare_outputs(b, _expected(cs, record), record, 3, cs["shape"])


def test_cuboid_corner_groups(emu, tmp_path):
    cs = cases.case("cuboid_corners", 3)
    record = ["p", "p_max"]
    _save(emu, cs, tmp_path / "in.h5")
    v = h5lite.read(tmp_path / "in.h5")
    assert int(v["sensor_mask_type"].ravel()[0]) == 1
    corners = np.asarray(cs["sensor"]["mask"])
    assert v["sensor_mask_corners"].shape == (6, corners.shape[1], 1)
    _run_file(emu, tmp_path / "in.h5", tmp_path / "out.h5", record)
    out = h5lite.read(tmp_path / "out.h5")
    want = _expected(cs, record)
    assert isinstance(want, list) and sorted(out["p"]) == [str(i + 1) for i in range(len(want))]
    for q, wq in enumerate(want):
        g = out["p"][str(q + 1)]
        ext = [int(corners[3 + a, q] - corners[a, q] + 1) for a in range(3)]
        assert g.shape == tuple(ext) + (np.asarray(wq["p"]).shape[-1],)
        # records per cuboid of shape [nx ny nz (Nt)], as the reference
        # reshapes them (reorderCuboidCorners.m)
        for got, w in ((g, np.asarray(wq["p"])),
                       (out["p_max"][str(q + 1)], np.asarray(wq["p_max"]).reshape(ext, order="F"))):
            assert got.shape == w.shape
            assert np.max(np.abs(got - w)) <= TOL * np.max(np.abs(w))


def test_file_errors(emu, tmp_path):
    from kwhip import _abi
    cs = cases.case("p0_lossless", 2)
    _save(emu, cs, tmp_path / "in.h5")
    v = h5lite.read(tmp_path / "in.h5")
    variables = {k: (a, "uint64" if a.dtype == np.uint64 else "single") for k, a in v.items()}
    bad = dict(variables, elastic_flag=(np.array([1], dtype=np.uint64), "uint64"))
    h5lite.write_input(tmp_path / "elastic.h5", bad, {"file_type": "input"})
    with pytest.raises(_abi.KWError) as ei:
        _run_file(emu, tmp_path / "elastic.h5", tmp_path / "o.h5", ["p"])
    assert ei.value.code == _abi.KW_EUNSUP and "elastic_flag" in str(ei.value)
    missing = {k: x for k, x in variables.items() if k != "c0"}
    h5lite.write_input(tmp_path / "missing.h5", missing, {"file_type": "input"})
    with pytest.raises(_abi.KWError) as ei:
        _run_file(emu, tmp_path / "missing.h5", tmp_path / "o.h5", ["p"])
    assert ei.value.code == _abi.KW_EINVAL
    h5lite.write_input(tmp_path / "output.h5", variables, {"file_type": "output"})
    with pytest.raises(_abi.KWError):
        _run_file(emu, tmp_path / "output.h5", tmp_path / "o.h5", ["p"])
    (tmp_path / "junk.h5").write_bytes(b"not an hdf5 file at all" * 8)
    with pytest.raises(_abi.KWError):
        _run_file(emu, tmp_path / "junk.h5", tmp_path / "o.h5", ["p"])


def test_save_to_disk_option(tmp_path):
    import kwhip
    from kwhip import _abi
    cs = cases.case("psrc_add", 2)
    args = []
    for n in cs["shape"]:
        args += [n, cs["dx"]]
    kg = kwhip.kWaveGrid(*args)
    kg.setTime(cs["Nt"], cs["dt"])
    fn = tmp_path / "input"
    with _abi._backend_override("emu"):
        r = kwhip.kspaceFirstOrder2D(kg, cs["medium"], cs["source"], cs["sensor"],
                                     SaveToDisk=str(fn), **cs["opts"])
        assert r == []
        assert (tmp_path / "input.h5").exists()
        v = h5lite.read(tmp_path / "input.h5")
        assert int(v["Nx"].ravel()[0]) == cs["shape"][0]
        r2 = kwhip.kspaceFirstOrder2D(kg, cs["medium"], cs["source"], cs["sensor"],
                                      SaveToDisk=str(tmp_path / "again.h5"), SaveToDiskExit=False,
                                      HDFCompressionLevel=4, **cs["opts"])
        assert isinstance(r2, dict) and (tmp_path / "again.h5").exists()
        with pytest.raises(ValueError, match="not compatible with 1D"):
            kg1 = kwhip.kWaveGrid(64, 1e-4)
            kg1.setTime(10, 2e-8)
            kwhip.kspaceFirstOrder1D(kg1, {"sound_speed": 1500.0}, {"p0": np.ones(64)},
                                     {"mask": np.ones(64)}, SaveToDisk=str(tmp_path / "x.h5"))
        with pytest.raises(ValueError, match="HDFCompressionLevel"):
            kwhip.kspaceFirstOrder2D(kg, cs["medium"], cs["source"], cs["sensor"],
                                     SaveToDisk=str(tmp_path / "y.h5"), HDFCompressionLevel=12)
        with pytest.raises(ValueError, match="unknown file extension"):
            kwhip.kspaceFirstOrder2D(kg, cs["medium"], cs["source"], cs["sensor"],
                                     SaveToDisk=str(tmp_path / "z.txt"), **cs["opts"])


def test_command_line_arguments():
    """Argument handling of kspaceFirstOrder-HIP (no file is simulated, so
    no device is used)."""
    from kwhip import _abi
    exe = _abi.LIB_DIR / "kspaceFirstOrder-HIP"
    assert exe.exists(), f"{exe} not built"
    r = subprocess.run([str(exe), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "--p_raw" in r.stdout
    r = subprocess.run([str(exe), "-i", "in.h5"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "required" in r.stderr
    r = subprocess.run([str(exe), "-i", "a", "-o", "b", "--checkpoint_file", "c"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "unknown option" in r.stderr
    r = subprocess.run([str(exe), "-i", "/nonexistent/in.h5", "-o", "/tmp/o.h5"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1 and "cannot open" in r.stderr
