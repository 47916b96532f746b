"""kspaceFirstOrderAS on the GPU: every case of tests/as_cases.py (initial
pressure, pressure and velocity sources in every mode, time reversal;
lossless, Stokes, nonlinear; 'WSWA-FFT', and 'WSWS-FFT' for seven of them) on a 128 x 64 grid against the oracle (relative L-inf <= 1e-5
single, <= 1e-4 double), the spherical wave of a pulse centred on the axis against its
closed form (the scheme's discretisation error, as the oracle measures it),
the engine stepping axisymmetric problems beside others, and the file
contract:
'SaveToDisk' files run by kspaceFirstOrder-HIP against the oracle."""

import numpy as np
import pytest

import as_cases as ac

pytestmark = pytest.mark.gpu

TOL32, TOL64 = 1e-5, 1e-4


@pytest.mark.parametrize("name", ac.NAMES)
def test_gpu_as_matches_oracle(hip_lib, name):
    cs = ac.case(name, Nx=128, Ny=64, Nt=120)
    cs["pml"] = [12, 10]
    got = ac.run_kwhip(cs)
    e32 = ac.compare(got, ac.run_oracle(cs, "single"))
    e64 = ac.compare(got, ac.run_oracle(cs, "off"))
    print(f"[as] {name}: worst vs fp32 {max(e32.values()):.2e}, vs fp64 {max(e64.values()):.2e}",
          flush=True)
    for k in e32:
        assert e32[k] <= TOL32, (k, e32[k])
        assert e64[k] <= TOL64, (k, e64[k])


@pytest.mark.parametrize("name,Nx,Ny,Nt,pml", [(n, 128, 64, 120, [12, 10]) for n in ac.WS_NAMES] +
                         [("hetero", 64, 16, 200, [8, 4])])
def test_gpu_as_wsws_matches_oracle(hip_lib, name, Nx, Ny, Nt, pml):
    # the 64 x 16 run carries the wave through the radial PML to the last
    # radial column, whose u_y the reference updates outside its layout
    cs = ac.case(name, Nx=Nx, Ny=Ny, Nt=Nt, sym="WSWS-FFT")
    cs["pml"] = pml
    got = ac.run_kwhip(cs)
    e32 = ac.compare(got, ac.run_oracle(cs, "single"))
    e64 = ac.compare(got, ac.run_oracle(cs, "off"))
    print(f"[as-ws] {name} {Nx}x{Ny}: worst vs fp32 {max(e32.values()):.2e}, "
          f"vs fp64 {max(e64.values()):.2e}", flush=True)
    for k in e32:
        assert e32[k] <= TOL32, (k, e32[k])
        assert e64[k] <= TOL64, (k, e64[k])


def test_gpu_as_spherical_wave(hip_lib):
    import kwhip
    Nx, Ny, dx, c0, Nt = 128, 64, 1e-4, 1500.0, 60
    dt = 0.2 * dx / c0
    p0, exact, inside = ac.spherical_wave(Nx, Ny, dx, c0, Nt, dt, 6 * dx)
    kg = kwhip.kWaveGrid(Nx, dx, Ny, dx)
    kg.setTime(Nt, dt)
    out = kwhip.kspaceFirstOrderAS(kg, {"sound_speed": c0, "density": 1000.0}, {"p0": p0},
                                   {"mask": np.ones((Nx, Ny), bool), "record": ["p_final"]},
                                   PMLSize=16, Smooth=False)
    pf = np.asarray(out["p_final"]).reshape(Nx, Ny, order="F")
    err = float(np.max(np.abs(pf - exact)[inside]) / np.max(np.abs(exact)))
    print(f"[as] spherical wave, w = 6 points: rel L-inf {err:.2e}", flush=True)
    assert err < 6e-4


def test_gpu_engine_refuses_the_file_layout(hip_lib):
    # the radial grid of the file contract (axisymmetric = 2) is run by
    # kw_run_file, not by the engine
    import kwhip
    from kwhip.solver import prepare, _parse_options
    cs = ac.case("homogeneous")
    kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.setTime(4, cs["dt"])
    with kwhip.Engine() as eng:
        pp = prepare(kg, cs["medium"], {"p0": np.ones((cs["Nx"], cs["Ny"]))}, None, 2,
                     _parse_options((), {"Smooth": False, "PMLSize": 4}))
        pp.prob.axisymmetric = 2
        import ctypes as C
        from kwhip import _abi
        tid = C.c_int64()
        with pytest.raises(kwhip.KWError, match="kw_run_file"):
            _abi.call(eng.lib.kw_submit, eng.h, C.byref(pp.prob), C.byref(tid))


def test_gpu_engine_axisymmetric(hip_lib):
    # axisymmetric problems stepped by the engine beside a 2-D problem equal
    # kspaceFirstOrderAS run alone (full rounds and the last partial one)
    import test_engine
    test_engine._as_mix("hip", 8)


@pytest.mark.parametrize("name", ["hetero", "source_additive", "stokes_nonlinear",
                                  "velocity_additive"])
def test_gpu_as_command_line(hip_lib, tmp_path, name):
    """kspaceFirstOrderAS(..., 'SaveToDisk') writes the radial grid's input
    file (for one case re-written the way MATLAB writes it, chunked and
    deflated), and kspaceFirstOrder-HIP runs it with the flags kspaceFirstOrder3DC.m passes;
    the output file's records against the oracle (the kspaceFirstOrderASC
    route)."""
    import subprocess
    import sys
    from pathlib import Path
    import kwhip
    from kwhip import _abi
    sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
    import h5lite
    cs = ac.case(name, Nx=128, Ny=64, Nt=120)
    cs["pml"] = [12, 10]
    print(f"[as-cli] {name}", flush=True)
    kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.setTime(cs["Nt"], cs["dt"])
    assert kwhip.kspaceFirstOrderAS(kg, dict(cs["medium"]), dict(cs["source"]), dict(cs["sensor"]),
                                    PMLSize=cs["pml"], Smooth=False,
                                    SaveToDisk=str(tmp_path / "in.h5")) == []
    fn = tmp_path / "in.h5"
    if name == "hetero":
        # the test writer's chunk index holds one B-tree leaf: the grids of
        # this case, not the long source signals of the others
        v = h5lite.read(fn)
        variables = {k: (a, "uint64" if a.dtype == np.uint64 else "single") for k, a in v.items()}
        h5lite.write_input(tmp_path / "matlab.h5", variables, h5lite.read_attrs(fn), compression=4)
        fn = tmp_path / "matlab.h5"
    exe = _abi.LIB_DIR / "kspaceFirstOrder-HIP"
    r = subprocess.run([str(exe), "-i", str(fn), "-o", str(tmp_path / "out.h5"),
                        "--p_raw", "--p_max", "--p_rms", "--p_final", "--u_raw", "--u_final"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = ac.file_records(h5lite.read(tmp_path / "out.h5"), cs["Nx"], cs["Ny"])
    for cast, tol in (("single", TOL32), ("off", TOL64)):
        ref = ac.run_oracle(cs, cast)
        errs = ac.compare(got, {k: np.squeeze(np.asarray(ref[k], dtype=np.float64)) for k in got})
        print(f"[as-cli] {name} worst vs oracle ({cast}): {max(errs.values()):.2e}", flush=True)
        for k, e in errs.items():
            assert e <= tol, (k, cast, e)
