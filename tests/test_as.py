"""kspaceFirstOrderAS (the axisymmetric solver, 'WSWA-FFT') on the CPU lane
emulator against the oracle's restatement of kspaceFirstOrderAS.m, which
mirrors the fields onto the expanded radial grid every step (the library
instead steps the expanded grid itself and adds the u_r / r term): relative
L-inf <= 1e-5 against single precision and <= 1e-4 against double, for p,
its statistics, p_final, u and u_final; lossless, nonlinear, Stokes, Stokes
nonlinear, heterogeneous, smoothed p0.  The oracle is pinned by the
spherical-wave tests in tests/test_as_oracle.py.  Also the reference's
option checks and what the library refuses."""

import numpy as np
import pytest

import as_cases as ac

TOL32, TOL64 = 1e-5, 1e-4


@pytest.fixture(autouse=True)
def _emu(monkeypatch):
    monkeypatch.setenv("KW_MAX_PENCILS", "4")
    from kwhip import _abi
    with _abi._backend_override("emu"):
        yield


@pytest.mark.parametrize("name", ac.NAMES)
def test_as_matches_oracle(name):
    cs = ac.case(name)
    got = ac.run_kwhip(cs)
    e32 = ac.compare(got, ac.run_oracle(cs, "single"))
    e64 = ac.compare(got, ac.run_oracle(cs, "off"))
    for k in e32:
        assert e32[k] <= TOL32, (k, e32[k])
        assert e64[k] <= TOL64, (k, e64[k])


@pytest.mark.parametrize("name", ac.WS_NAMES)
def test_as_wsws_matches_oracle(name):
    cs = ac.case(name, sym="WSWS-FFT")
    got = ac.run_kwhip(cs)
    e32 = ac.compare(got, ac.run_oracle(cs, "single"))
    e64 = ac.compare(got, ac.run_oracle(cs, "off"))
    for k in e32:
        assert e32[k] <= TOL32, (k, e32[k])
        assert e64[k] <= TOL64, (k, e64[k])
    # 'WSWS' (the reference's discrete trigonometric transforms) gives the
    # field of 'WSWS-FFT'
    if name == "hetero":
        again = ac.run_kwhip(dict(cs, sym="WSWS"))
        assert np.array_equal(np.asarray(again["p"]), np.asarray(got["p"]))


def test_as_option_checks():
    import kwhip
    cs = ac.case("homogeneous")
    kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.setTime(4, cs["dt"])
    args = (kg, cs["medium"], dict(cs["source"]), cs["sensor"])
    with pytest.raises(ValueError, match="'RadialSymmetry' must be set to 'WSWA', 'WSWS'"):
        kwhip.kspaceFirstOrderAS(*args, RadialSymmetry="WSW")
    with pytest.raises(ValueError, match="'UseSG' is not supported for axisymmetric"):
        kwhip.kspaceFirstOrderAS(*args, UseSG=False)
    with pytest.raises(ValueError, match="medium.alpha_mode is not supported with the axisymmetric"):
        kwhip.kspaceFirstOrderAS(kg, dict(cs["medium"], alpha_coeff=0.5, alpha_mode="no_dispersion"),
                                 dict(cs["source"]), cs["sensor"])
    p0 = cs["source"]["p0"]
    with pytest.raises(ValueError, match="source.p0 and source.p can't be defined"):
        kwhip.kspaceFirstOrderAS(kg, cs["medium"], {"p0": p0, "p_mask": p0 > 0.5,
                                                    "p": np.ones((1, 4))}, cs["sensor"])
    edge = np.zeros_like(p0, dtype=bool)
    edge[4, -2] = True
    with pytest.raises(NotImplementedError, match="radial velocity sources"):
        kwhip.kspaceFirstOrderAS(kg, cs["medium"], {"u_mask": edge, "uy": np.ones((1, 4)),
                                                    "u_mode": "dirichlet"},
                                 cs["sensor"], RadialSymmetry="WSWS")
    with pytest.raises(NotImplementedError, match="non-staggered"):
        kwhip.kspaceFirstOrderAS(kg, cs["medium"], dict(cs["source"]),
                                 dict(cs["sensor"], record=["p", "I_avg"]))
    with pytest.raises(ValueError, match="source.uz is not defined"):
        kwhip.kspaceFirstOrderAS(kg, cs["medium"], {"u_mask": p0 > 0.5, "uz": np.ones((1, 4))},
                                 cs["sensor"])
    with pytest.raises(ValueError, match="number of time series must match"):
        kwhip.kspaceFirstOrderAS(kg, cs["medium"], {"p_mask": p0 > 0.5, "p": np.ones((2, 4))},
                                 cs["sensor"])
    with pytest.raises(ValueError, match="Unknown optional input 'RadialSymmetry'"):
        kwhip.kspaceFirstOrder2D(*args, RadialSymmetry="WSWA")
    # 'WSWA' (discrete trigonometric transforms in the reference) gives the
    # field of 'WSWA-FFT'
    a = kwhip.kspaceFirstOrderAS(*args, PMLSize=[6, 4], Smooth=False)
    b = kwhip.kspaceFirstOrderAS(*args, PMLSize=[6, 4], Smooth=False, RadialSymmetry="WSWA")
    assert np.array_equal(np.asarray(a["p"]), np.asarray(b["p"]))


@pytest.mark.parametrize("name", ["hetero", "stokes_nonlinear", "source_additive",
                                  "source_dirichlet", "velocity_additive"])
def test_as_file_contract(name, tmp_path):
    """'SaveToDisk' writes the radial grid (axisymmetric_flag 1, the
    variable set of a 2-D file), and kw_run_file runs it on the expanded
    grid: the records match the direct call and the oracle."""
    from kwhip import _abi
    cs = ac.case(name)
    out, v = ac.save_and_run_file(cs, tmp_path, _abi.load("emu"), matlab_style=name == "hetero")
    assert int(v["axisymmetric_flag"].ravel()[0]) == 1
    assert [int(v[k].ravel()[0]) for k in ("Nx", "Ny", "Nz")] == [cs["Nx"], cs["Ny"], 1]
    assert "rho0_sgz" not in v and v["c0"].size in (1, cs["Nx"] * cs["Ny"])
    assert [int(out[k].ravel()[0]) for k in ("Nx", "Ny")] == [cs["Nx"], cs["Ny"]]
    got = ac.file_records(out, cs["Nx"], cs["Ny"])
    direct = ac.run_kwhip(cs)
    ref32 = ac.run_oracle(cs, "single")
    for k in ("p", "p_max", "p_rms", "p_final", "ux", "uy", "ux_final", "uy_final"):
        assert got[k].shape == np.squeeze(np.asarray(direct[k])).shape, k
    # the file holds the inputs in single precision, so the operators built
    # from them differ from the direct call's by single-precision rounding
    e_direct = ac.compare(got, {k: np.squeeze(np.asarray(direct[k], dtype=np.float64))
                                for k in got})
    e32 = ac.compare(got, {k: np.squeeze(np.asarray(ref32[k], dtype=np.float64)) for k in got})
    for k in got:
        assert e_direct[k] <= TOL32, (k, e_direct[k])
        assert e32[k] <= TOL32, (k, e32[k])


def test_as_file_options(tmp_path, capsys):
    import kwhip
    from kwhip import _abi
    cs = ac.case("homogeneous")
    # 'WSWS' symmetries become 'WSWA' for the file (inputChecking.m:1402-1411)
    out, v = ac.save_and_run_file(cs, tmp_path, _abi.load("emu"), record=("p",),
                                  RadialSymmetry="WSWS")
    assert "changed to 'WSWA'" in capsys.readouterr().out
    assert int(v["axisymmetric_flag"].ravel()[0]) == 1
    ref = ac.run_oracle(dict(cs, sensor=dict(cs["sensor"], record=["p"])), "single")
    got = np.squeeze(out["p"]).astype(np.float64)
    assert np.max(np.abs(got - ref["p"])) <= TOL32 * np.max(np.abs(ref["p"]))
    # SaveToDiskExit false runs the simulation as well
    kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.setTime(cs["Nt"], cs["dt"])
    r = kwhip.kspaceFirstOrderAS(kg, cs["medium"], dict(cs["source"]), cs["sensor"],
                                 PMLSize=cs["pml"], Smooth=False, SaveToDisk=str(tmp_path / "b.h5"),
                                 SaveToDiskExit=False)
    assert isinstance(r, dict) and (tmp_path / "b.h5").exists()
    # the expanded layout is not a file layout, and the radial one is not run by kw_run
    from kwhip.solver import prepare, _parse_options
    pp = prepare(kg, cs["medium"], dict(cs["source"]), None, 2,
                 _parse_options((), {"Smooth": False, "PMLSize": [6, 4]}))
    pp.prob.axisymmetric = 1
    lib = _abi.load("emu")
    with pytest.raises(_abi.KWError, match="radial grid"):
        import ctypes as C
        _abi.call(lib.kw_save_problem, C.byref(pp.prob), str(tmp_path / "c.h5").encode())
    pp.prob.axisymmetric = 2
    out_arr = {}
    with pytest.raises(_abi.KWError, match="kw_run_file"):
        from kwhip.solver import run_prepared
        out_arr = run_prepared(pp)
    assert out_arr == {}


def test_as_pml_auto():
    """'PMLSize' 'auto' with the PML outside (getOptimalPMLSize.m with the
    radial symmetry): the radial PML is scored by the expanded length."""
    import kwhip
    from kwhip.solver import optimal_pml_size
    # 4 (16 + p) is a power of two at p = 16; 2 (16 + p) - 2 at p = 17
    assert optimal_pml_size((32, 16), (10, 40), "WSWA") == [16, 16]
    assert optimal_pml_size((32, 16), (10, 40), "WSWS") == [16, 17]
    cs = ac.case("hetero")
    kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.setTime(cs["Nt"], cs["dt"])
    a = kwhip.kspaceFirstOrderAS(kg, cs["medium"], dict(cs["source"]), cs["sensor"],
                                 PMLSize="auto", PMLInside=False, Smooth=False)
    b = kwhip.kspaceFirstOrderAS(kg, cs["medium"], dict(cs["source"]), cs["sensor"],
                                 PMLSize=[16, 16], PMLInside=False, Smooth=False)
    assert np.array_equal(np.asarray(a["p"]), np.asarray(b["p"]))
    with pytest.raises(ValueError, match="only supported with 'PMLInside' set to false"):
        kwhip.kspaceFirstOrderAS(kg, cs["medium"], dict(cs["source"]), cs["sensor"],
                                 PMLSize="auto")


def test_as_record_start_and_frequency_response():
    """sensor.record_start_index against the oracle, and
    sensor.frequency_response (kspaceFirstOrderAS.m:1401-1403: the recorded
    p filtered by gaussianFilter.m) against the oracle's record filtered by
    the oracle's own gaussian_filter."""
    from oracle import kwave_oracle as ko
    cs = ac.case("hetero")
    cs["sensor"] = dict(cs["sensor"], record=["p", "p_max", "u"], record_start_index=9)
    got = ac.run_kwhip(cs)
    ref = ac.run_oracle(cs, "single")
    e = ac.compare(got, ref)
    assert np.asarray(got["p"]).shape[1] == cs["Nt"] - 8
    for k in e:
        assert e[k] <= TOL32, (k, e[k])
    fr = [2e6, 80.0]
    cs2 = ac.case("hetero")
    cs2["sensor"] = dict(cs2["sensor"], record=["p"], frequency_response=fr)
    got2 = np.asarray(ac.run_kwhip(cs2)["p"], dtype=np.float64)
    ref2 = ac.run_oracle(dict(cs2, sensor=dict(cs2["sensor"], record=["p"])), "single")["p"]
    want = ko.gaussian_filter(ref2, 1.0 / cs2["dt"], fr[0], fr[1])
    assert np.max(np.abs(got2 - want)) <= TOL32 * np.max(np.abs(want))
