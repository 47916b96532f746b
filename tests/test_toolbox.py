"""The new toolbox's acoustic solver API (kwhip.toolbox: Grid,
AcousticMedium, AcousticSource, AcousticSensor, AcousticSolver) on the CPU
lane emulator against the toolbox oracle (oracle/toolbox_oracle.py, pinned
by tests/test_toolbox_oracle.py), in single precision at relative L-inf
<= 1e-5 (double-precision oracle: <= 1e-4), for every output: the final
pressure, velocity and split density, the recorded sensor data and times.
Cases: repeated runs with a smaller and a larger time step, absorption
('on', 'noDispersion', 'noAbsorption'), an initial particle velocity,
'ongrid' velocity recorded every third step, a run(Nt=0) followed by more
runs (the reference then applies the initial condition again); and the
reference's input checks."""

import numpy as np
import pytest

import toolbox_cases as tc

TOL32, TOL64 = 1e-5, 1e-4


@pytest.fixture(autouse=True)
def _emu(monkeypatch):
    monkeypatch.setenv("KW_MAX_PENCILS", "4")
    from kwhip import _abi
    with _abi._backend_override("emu"):
        yield


@pytest.mark.parametrize("name", tc.NAMES)
@pytest.mark.parametrize("shape", [(24, 20), (10, 12, 8), (40,)])
def test_toolbox_matches_oracle(name, shape):
    cs = tc.case(name, shape, pad=4)
    got = tc.run_adapter(cs)
    e32 = tc.compare(got, tc.run_oracle(cs, "single"))
    e64 = tc.compare(got, tc.run_oracle(cs, "double"))
    for k in e32:
        assert e32[k] <= TOL32, (k, e32[k])
        assert e64[k] <= TOL64, (k, e64[k])


def test_input_checks():
    from kwhip import toolbox as tb
    kg = tb.Grid([16, 12], 1e-4, 2)
    other = tb.Grid([16, 12], 1e-4, 2)
    m = tb.AcousticMedium(kg)
    s = tb.AcousticSource(kg)
    with pytest.raises(ValueError, match="The property soundSpeed must be defined"):
        tb.AcousticSolver(kg, m, s)
    m.soundSpeed = 1500.0
    m.density = 1000.0
    with pytest.raises(ValueError, match="The property initialPressure must be defined"):
        tb.AcousticSolver(kg, m, s)
    s.initialPressure = np.ones((16, 12))
    with pytest.raises(ValueError, match="medium must be an object of Medium or AcousticMedium"):
        tb.AcousticSolver(kg, s, s)
    with pytest.raises(ValueError, match="different Grid object"):
        tb.AcousticSolver(other, m, s)
    with pytest.raises(ValueError, match="sensor must be an AcousticSensor"):
        tb.AcousticSolver(kg, m, s, sensor=m)
    with pytest.raises(ValueError, match="Grid:incorrectSize"):
        m.soundSpeed = np.ones((15, 12))
    with pytest.raises(ValueError, match="gridSpacing must be a scalar"):
        tb.Grid([16, 12], [1e-4, 1e-4, 1e-4])
    sol = tb.AcousticSolver(kg, m, s)
    with pytest.raises(ValueError, match="Nt and dt"):
        sol.run(Nt=5)
    assert sol.pressure is None


def test_auto_time_step_and_time_array():
    """run(CFL=, EndTime=) (autoComputeTimeStep.m:40-69) and run.m's timeArray."""
    from kwhip import toolbox as tb
    kg = tb.Grid([16, 12], 1e-4, 2)
    m = tb.AcousticMedium(kg)
    m.soundSpeed = 1500.0
    m.density = 1000.0
    s = tb.AcousticSource(kg)
    s.initialPressure = tc.gauss((16, 12), 1e-4)
    sol = tb.AcousticSolver(kg, m, s)
    sol.run(EndTime=2e-7)
    dt = 0.3 * 1e-4 / 1500.0
    Nt = int(np.ceil(2e-7 / dt))
    assert sol.timeStepsTaken == Nt
    assert sol.timeArray.shape == (Nt + 1,)
    assert np.isclose(sol.timeArray[-1], 2e-7, rtol=1e-6)
    sol.run(CFL=0.5)
    end = np.linalg.norm(np.array([16, 12]) * 1e-4) / 1500.0
    Nt2 = int(np.ceil(end / (0.5 * 1e-4 / 1500.0)))
    assert sol.timeStepsTaken == Nt + Nt2
    assert np.isclose(sol.timeArray[-1], 2e-7 + end, rtol=1e-5)
    sol.close()
