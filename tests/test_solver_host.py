"""Host-side checks of the Python mirror (no device work: every case raises
before the simulation starts, or runs on the CPU emulator build)."""

import numpy as np
import pytest

import cases


@pytest.fixture(autouse=True)
def _emulator():
    """These checks run the public API over the emulator build."""
    from kwhip import _abi
    with _abi._backend_override("emu"):
        yield


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
    run = lambda m, s, se: kwhip.kspaceFirstOrder1D(_grid(), m, s, se, Smooth=False)
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
        kwhip.kspaceFirstOrder1D(_grid(), med, src, {"mask": np.ones(24)})
    stokes = {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 0.5, "alpha_mode": "stokes",
              "alpha_filter": np.ones(24)}
    with pytest.raises(ValueError, match="alpha_filter is not supported"):
        kwhip.kspaceFirstOrder1D(_grid(), stokes, {"p": np.ones(5), "p_mask": m},
                                 {"mask": np.ones(24)})


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
                                     Smooth=False)


def test_options_of_other_solvers_are_checked():
    """'MultiAxialPMLRatio' is checked and has no effect on the fluid code;
    'UseFD' exists in 1-D only (finite differences are not built)."""
    import kwhip
    med = {"sound_speed": 1500.0, "density": 1000.0}
    p0 = np.zeros(24)
    p0[10] = 1
    sens = {"mask": np.ones(24)}
    a = kwhip.kspaceFirstOrder1D(_grid(), med, {"p0": p0}, sens, Smooth=False,
                                 PMLSize=4)
    b = kwhip.kspaceFirstOrder1D(_grid(), med, {"p0": p0}, sens, Smooth=False,
                                 PMLSize=4, MultiAxialPMLRatio=0.1, UseFD=False)
    assert np.array_equal(a, b)
    with pytest.raises(ValueError, match="single positive value"):
        kwhip.kspaceFirstOrder1D(_grid(), med, {"p0": p0}, sens, MultiAxialPMLRatio=-1)
    with pytest.raises(NotImplementedError):
        kwhip.kspaceFirstOrder1D(_grid(), med, {"p0": p0}, sens, UseFD=4)
    kg = kwhip.kWaveGrid(16, 1e-4, 16, 1e-4)
    kg.setTime(5, 2e-8)
    with pytest.raises(ValueError, match="only supported in 1D"):
        kwhip.kspaceFirstOrder2D(kg, med, {"p0": np.ones((16, 16))}, {"mask": np.ones((16, 16))},
                                 UseFD=2)


def test_stability_warning():
    """kspaceFirstOrder_inputChecking.m:1270-1281 with checkStability.m: a
    time step above the stability limit prints the reference's warning and
    runs; c_ref = max(c) in a lossless medium is unconditionally stable."""
    import math
    import warnings

    import kwhip
    from kwhip.solver import check_stability
    kg = kwhip.kWaveGrid(32, 1e-4, 32, 1e-4)
    kmax = math.pi * math.sqrt(2) / 1e-4
    c = np.array([1400.0, 1600.0])
    assert check_stability(kg, c, 1600.0, {}) == math.inf
    # closed form of the lossless case with c_ref < max(c)
    assert check_stability(kg, c, 1400.0, {}) == pytest.approx(
        2 / (1400.0 * kmax) * math.asin(1400.0 / 1600.0), rel=1e-12)
    # vanishing absorption: the fixed point of dt sinc(c k dt/2) = 2/(c k),
    # i.e. c k dt/2 = pi/2, which the reference's iteration (stopped at
    # steps of 1e-12 s) approaches from below within a percent
    weak = {"alpha_coeff": 1e-12, "alpha_power": 1.5}
    lim = check_stability(kg, np.array(1500.0), 1500.0, weak)
    assert 0.99 * math.pi / (1500.0 * kmax) < lim <= math.pi / (1500.0 * kmax)
    med = {"sound_speed": 1500.0, "density": 1000.0, "sound_speed_ref": 1000.0}
    p0 = np.zeros((32, 32))
    p0[16, 16] = 1
    limit = check_stability(kg, np.array(1500.0), 1000.0, {})
    for dt, warned in ((0.9 * limit, False), (1.1 * limit, True)):
        kg.setTime(3, dt)
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter("always")
            kwhip.kspaceFirstOrder2D(kg, med, {"p0": p0}, {"mask": np.ones((32, 32))}, Smooth=False,
                                     PMLSize=4)
        hit = [x for x in w if "time step may be too large" in str(x.message)]
        assert bool(hit) == warned


def test_alpha_sign_and_pml_range_rounding():
    """medium.alpha_sign passes MATLAB's sign, 0 included (a zero removes the
    term: the run equals 'no_dispersion'); 'PMLRange' is rounded half away
    from zero as MATLAB's round (12.5 -> 13)."""
    import kwhip
    from kwhip.solver import optimal_pml_size
    assert optimal_pml_size([101], (12.5, 20)) == [17]
    assert optimal_pml_size([101], (12.4, 20)) == [12]
    kg = kwhip.kWaveGrid(32, 1e-4, 30, 1e-4)
    kg.setTime(12, 1.5e-8)
    p0 = np.zeros((32, 30))
    p0[14:18, 13:17] = 1
    base = {"sound_speed": 1500.0, "density": 1000.0, "alpha_coeff": 2.0, "alpha_power": 1.5}
    sens = {"mask": np.ones((32, 30))}
    a = kwhip.kspaceFirstOrder2D(kg, dict(base, alpha_sign=[3.0, 0.0]), {"p0": p0}, sens,
                                 Smooth=False, PMLSize=4)
    b = kwhip.kspaceFirstOrder2D(kg, dict(base, alpha_mode="no_dispersion"), {"p0": p0}, sens,
                                 Smooth=False, PMLSize=4)
    c = kwhip.kspaceFirstOrder2D(kg, base, {"p0": p0}, sens, Smooth=False, PMLSize=4)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_stream_to_disk():
    """'StreamToDisk' (kspaceFirstOrder_checkOptionalInputs.m:353-374,
    kspaceFirstOrder_inputChecking.m:1369-1374): 3-D fluid forward
    simulations recording p only; the record is the one of a run without
    it (the reference reads its streamed file back whole)."""
    import kwhip
    cs = cases.case("p0_hetero_absorbing", 3)
    cs["sensor"] = {"mask": cs["sensor"]["mask"]}   # sensor.record absent: p
    args = []
    for n in cs["shape"]:
        args += [n, cs["dx"]]
    kg = kwhip.kWaveGrid(*args)
    kg.setTime(cs["Nt"], cs["dt"])
    run = lambda sen, **o: kwhip.kspaceFirstOrder3D(kg, cs["medium"], cs["source"], sen,  # noqa: E731
                                                    **dict(cs["opts"], **o))
    plain = run(cs["sensor"])
    for v in (True, 5, 200):
        assert np.array_equal(np.asarray(run(cs["sensor"], StreamToDisk=v)), np.asarray(plain))
    with pytest.raises(ValueError, match="single scalar or Boolean"):
        run(cs["sensor"], StreamToDisk=[1, 2])
    with pytest.raises(ValueError, match="sensor.record = {'p'}"):
        run(dict(cs["sensor"], record=["p", "p_max"]), StreamToDisk=True)
    with pytest.raises(ValueError, match="non-zero sensor mask"):
        run({"mask": np.zeros(cs["shape"])}, StreamToDisk=True)
    with pytest.raises(ValueError, match="only compatible with 3D fluid"):
        kwhip.kspaceFirstOrder1D(_grid(), {"sound_speed": 1500.0}, {"p0": np.ones(24)},
                                 {"mask": np.ones(24)}, StreamToDisk=True)
