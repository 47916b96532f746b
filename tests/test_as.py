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


def test_as_option_checks():
    import kwhip
    cs = ac.case("homogeneous")
    kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
    kg.setTime(4, cs["dt"])
    args = (kg, cs["medium"], dict(cs["source"]), cs["sensor"])
    with pytest.raises(ValueError, match="'RadialSymmetry' must be set to 'WSWA', 'WSWS'"):
        kwhip.kspaceFirstOrderAS(*args, RadialSymmetry="WSW")
    with pytest.raises(NotImplementedError):
        kwhip.kspaceFirstOrderAS(*args, RadialSymmetry="WSWS-FFT")
    with pytest.raises(ValueError, match="'UseSG' is not supported for axisymmetric"):
        kwhip.kspaceFirstOrderAS(*args, UseSG=False)
    with pytest.raises(ValueError, match="medium.alpha_mode is not supported with the axisymmetric"):
        kwhip.kspaceFirstOrderAS(kg, dict(cs["medium"], alpha_coeff=0.5, alpha_mode="no_dispersion"),
                                 dict(cs["source"]), cs["sensor"])
    p0 = cs["source"]["p0"]
    with pytest.raises(ValueError, match="source.p0 and source.p can't be defined"):
        kwhip.kspaceFirstOrderAS(kg, cs["medium"], {"p0": p0, "p_mask": p0 > 0.5,
                                                    "p": np.ones((1, 4))}, cs["sensor"])
    with pytest.raises(NotImplementedError):
        kwhip.kspaceFirstOrderAS(kg, cs["medium"], {"u_mask": p0 > 0.5, "ux": np.ones((1, 4))},
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
