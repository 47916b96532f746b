"""Host-side checks of the Python mirror (no device work: every case raises
before the simulation starts, or runs on the CPU emulator build)."""

import numpy as np
import pytest

import cases


def _grid(n=24):
    import kwhip
    kg = kwhip.kWaveGrid(n, 1e-4)
    kg.setTime(5, 2e-8)
    return kg


def test_field_names_are_checked_as_the_reference_does():
    """kspaceFirstOrder_inputChecking.m / checkFieldNames.m: unknown fields
    of medium, source and sensor raise the reference's error; sensor
    directivity exists in 2-D only; sensor.record_mode is refused."""
    import kwhip
    med = {"sound_speed": 1500.0, "density": 1000.0}
    p0 = np.zeros(24)
    p0[10] = 1
    sens = {"mask": np.ones(24)}
    run = lambda m, s, se: kwhip.kspaceFirstOrder1D(_grid(), m, s, se, Smooth=False, Backend="emu")
    with pytest.raises(ValueError, match="speed is not a valid field for the structure medium"):
        run(dict(med, speed=1.0), {"p0": p0}, sens)
    with pytest.raises(ValueError, match="uy is not a valid field for the structure source"):
        run(med, {"p0": p0, "uy": np.zeros(5)}, sens)
    with pytest.raises(ValueError, match="directivity_angle is not a valid field for the structure sensor"):
        run(med, {"p0": p0}, dict(sens, directivity_angle=np.zeros(24)))
    with pytest.raises(ValueError, match="record_mode input has been deprecated"):
        run(med, {"p0": p0}, dict(sens, record_mode="statistics"))


def test_frequency_ref_and_alpha_filter_inputs_are_checked():
    import kwhip
    med = {"sound_speed": 1500.0, "density": 1000.0}
    m = np.zeros(24)
    m[8] = 1
    src = {"p": np.ones(5), "p_mask": m, "p_frequency_ref": 1e9}
    with pytest.raises(ValueError, match="maximum frequency supported"):
        kwhip.kspaceFirstOrder1D(_grid(), med, src, {"mask": np.ones(24)}, Backend="emu")
    stokes = {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 0.5, "alpha_mode": "stokes",
              "alpha_filter": np.ones(24)}
    with pytest.raises(ValueError, match="alpha_filter is not supported"):
        kwhip.kspaceFirstOrder1D(_grid(), stokes, {"p": np.ones(5), "p_mask": m},
                                 {"mask": np.ones(24)}, Backend="emu")


def test_cuboid_corner_errors():
    import kwhip
    kg = kwhip.kWaveGrid(16, 1e-4, 14, 1e-4)
    kg.setTime(5, 2e-8)
    med = {"sound_speed": 1500.0, "density": 1000.0}
    p0 = np.zeros((16, 14))
    p0[8, 7] = 1
    for corners, msg in (([[2], [3], [5.5], [6]], "integers"), ([[5], [3], [2], [6]], "x2 => x1"),
                         ([[2], [3], [17], [6]], "within the grid")):
        with pytest.raises(ValueError, match=msg):
            kwhip.kspaceFirstOrder2D(kg, med, {"p0": p0}, {"mask": np.array(corners, dtype=float)},
                                     Smooth=False, Backend="emu")


def test_options_of_other_solvers_are_checked():
    """'MultiAxialPMLRatio' is checked and has no effect on the fluid code;
    'UseFD' exists in 1-D only (finite differences are not built)."""
    import kwhip
    med = {"sound_speed": 1500.0, "density": 1000.0}
    p0 = np.zeros(24)
    p0[10] = 1
    sens = {"mask": np.ones(24)}
    a = kwhip.kspaceFirstOrder1D(_grid(), med, {"p0": p0}, sens, Smooth=False, Backend="emu",
                                 PMLSize=4)
    b = kwhip.kspaceFirstOrder1D(_grid(), med, {"p0": p0}, sens, Smooth=False, Backend="emu",
                                 PMLSize=4, MultiAxialPMLRatio=0.1, UseFD=False)
    assert np.array_equal(a, b)
    with pytest.raises(ValueError, match="single positive value"):
        kwhip.kspaceFirstOrder1D(_grid(), med, {"p0": p0}, sens, Backend="emu", MultiAxialPMLRatio=-1)
    with pytest.raises(NotImplementedError):
        kwhip.kspaceFirstOrder1D(_grid(), med, {"p0": p0}, sens, Backend="emu", UseFD=4)
    kg = kwhip.kWaveGrid(16, 1e-4, 16, 1e-4)
    kg.setTime(5, 2e-8)
    with pytest.raises(ValueError, match="only supported in 1D"):
        kwhip.kspaceFirstOrder2D(kg, med, {"p0": np.ones((16, 16))}, {"mask": np.ones((16, 16))},
                                 Backend="emu", UseFD=2)
