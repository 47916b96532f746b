"""Engine: simulations submitted together are stepped together (one launch
per pass for every simulation with the same dimension and equation of
state, whatever their sizes, media, sources, sensors and current step);
each result must equal the same problem run alone.  CPU tests use the
emulator build; the GPU test runs the product library."""

import numpy as np
import pytest

import cases

TOL = 1e-6   # same arithmetic per problem; only FFT plan variants may differ


def _alone(cs, backend):
    return cases.run_kwhip(cs, backend)


def _submit(eng, cs):
    import kwhip
    args = []
    for n in cs["shape"]:
        args += [n, cs["dx"]]
    kg = kwhip.kWaveGrid(*args)
    kg.setTime(cs["Nt"], cs["dt"])
    return eng.submit(kg, cs["medium"], cs["source"], cs["sensor"], cs["dim"], **cs["opts"])


def _same(a, b):
    if isinstance(a, list):
        # cuboid corners: one record per cuboid
        assert isinstance(b, list) and len(a) == len(b)
        for x, y in zip(a, b):
            _same(x, y)
        return
    if not isinstance(a, dict):
        a, b = {"p": a}, {"p": b}
    assert sorted(a) == sorted(b)
    for k in a:
        x, y = np.asarray(a[k], np.float64), np.asarray(b[k], np.float64)
        assert x.shape == y.shape, k
        scale = max(np.max(np.abs(y)), 1e-30)
        assert np.max(np.abs(x - y)) <= TOL * scale, (k, np.max(np.abs(x - y)) / scale)


MIX = [
    # (dim, case, shape): different sizes, media, sources and sensors in one batch
    (2, "p0_hetero_absorbing", (32, 30)),
    (2, "psrc_add", (24, 40)),
    (2, "p0_nonlinear_absorbing", (45, 27)),
    (2, "usrc_many", (30, 30)),
    (2, "time_reversal", (36, 32)),
    (2, "p0_lossless", (40, 24)),
    (2, "psrc_dir", (27, 36)),
    (2, "stokes", (30, 25)),
    (3, "p0_hetero_absorbing", (16, 18, 20)),
    (3, "psrc_label", (20, 16, 15)),
    (3, "time_reversal", (16, 16, 16)),
    (3, "cartesian_nearest", (18, 20, 16)),
    # sizes with fixed-length kernels: same-size groups on their own streams
    (2, "p0_hetero_absorbing", (64, 64)),
    (2, "psrc_add", (64, 64)),
    (2, "p0_lossless", (75, 75)),
    (2, "time_reversal", (75, 75)),
    # per-member passes inside a group: velocity on the non-staggered grid
    (2, "u_nonstaggered", (30, 27)),
    (3, "intensity", (16, 15, 18)),
    (2, "u_split", (27, 30)),
    (3, "u_split", (15, 18, 16)),
    # Cartesian points interpolated on simplices
    (2, "cartesian_linear", (30, 32)),
    # per-member directivity filters; host-side sensor forms
    (2, "directivity_gradient", (30, 32)),
    (3, "cuboid_corners", (16, 15, 18)),
    (2, "freq_response", (32, 30)),
    (2, "alpha_filter", (27, 30)),
    # 1-D problems (one-row grids) in the same engine
    (1, "psrc_add", (96,)),
    (1, "time_reversal", (75,)),
]


def _run_mix(backend, mix, steps_per_round=0, device_bytes=0):
    import kwhip
    css = [cases.case(name, dim, shape) for dim, name, shape in mix]
    from kwhip import _abi
    with _abi._backend_override(backend), \
            kwhip.Engine(device_bytes=device_bytes, steps_per_round=steps_per_round) as eng:
        tickets = [_submit(eng, cs) for cs in css]
        got = [eng.wait(t) for t in tickets]
        st = eng.stats()
    return css, got, st


@pytest.fixture
def few_pencils(monkeypatch):
    # fewer pencils per emulated block: same code paths, faster emulation
    monkeypatch.setenv("KW_MAX_PENCILS", "4")


def test_engine_mix_equals_alone_emu(few_pencils):
    css, got, st = _run_mix("emu", MIX, steps_per_round=5)
    assert st["completed"] == len(MIX) and st["failed"] == 0
    assert st["device_bytes"] == 0
    for cs, g in zip(css, got):
        _same(g, _alone(cs, "emu"))


def test_engine_small_budget_admits_in_turn_emu(few_pencils):
    """A budget that holds about one simulation at a time still completes
    every request, in order of submission."""
    mix = MIX[:4]
    css, got, st = _run_mix("emu", mix, device_bytes=1, steps_per_round=7)
    assert st["completed"] == len(mix)
    for cs, g in zip(css, got):
        _same(g, _alone(cs, "emu"))


def test_engine_poll_cancel_and_errors_emu(few_pencils):
    import kwhip
    cs = cases.case("p0_hetero_absorbing", 2, (32, 30))
    from kwhip import _abi
    with _abi._backend_override("emu"), kwhip.Engine(device_bytes=1) as eng:
        t1 = _submit(eng, cs)
        t2 = _submit(eng, cs)
        t2.cancel()
        assert eng.poll(t1)["status"] in ("queued", "running", "done")
        assert eng.poll(t2)["status"] == "cancelled"
        with pytest.raises(kwhip.KWError):
            eng.wait(t2)
        _same(eng.wait(t1), _alone(cs, "emu"))
        info = eng.stats()
        assert info["cancelled"] == 1 and info["completed"] == 1


def test_engine_full_tiling_emu():
    """Blocks of the GPU's size (up to 256 / 512 threads), so members with
    smaller tilings share blocks sized for the largest one."""
    mix = [MIX[0], MIX[2], MIX[8], MIX[9]]
    css, got, st = _run_mix("emu", mix, steps_per_round=4)
    for cs, g in zip(css, got):
        _same(g, _alone(cs, "emu"))


@pytest.mark.gpu
def test_engine_mix_equals_alone_hip(hip_lib):
    print("[engine] mixed batch of", len(MIX), "problems", flush=True)
    css, got, st = _run_mix("hip", MIX, steps_per_round=5)
    assert st["completed"] == len(MIX) and st["failed"] == 0
    for cs, g in zip(css, got):
        _same(g, _alone(cs, "hip"))


def _unstable_case(N=32, Nt=200):
    """A 2-D problem that becomes non-finite within Nt steps: half the grid at
    4500 m/s with the reference sound speed set to 1500 m/s and dt at the
    1500 m/s CFL limit of 1 (the k-space correction then under-corrects the
    fast half; the oracle in single precision goes to NaN within the same
    200 steps)."""
    dx = 1e-4
    c = np.full((N, N), 1500.0)
    c[N // 2:, :] = 4500.0
    p0 = np.zeros((N, N))
    p0[8:12, 8:12] = 1.0
    return {"dim": 2, "shape": (N, N), "dx": dx, "Nt": Nt, "dt": dx / 1500.0,
            "medium": {"sound_speed": c, "density": 1000.0, "sound_speed_ref": 1500.0},
            "source": {"p0": p0},
            "sensor": {"mask": np.ones((N, N)), "record": ["p_max", "p_final"]},
            "opts": {"PMLSize": 4, "Smooth": False}}


def _isolation(backend):
    """One request diverges in a group with two stable neighbours of the
    same size and equation of state: it alone fails, with KW_ENUMERIC, and
    the neighbours' results equal their runs alone."""
    import warnings

    import kwhip
    from kwhip import _abi
    good = [cases.case("p0_lossless", 2, (32, 32)), cases.case("p0_hetero_absorbing", 2, (32, 32))]
    for cs in good:
        cs["opts"] = dict(cs["opts"])
    bad = _unstable_case()
    bad["Nt"] = max(bad["Nt"], max(cs["Nt"] for cs in good))
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # the unstable request's stability warning
        with _abi._backend_override(backend), kwhip.Engine(steps_per_round=8) as eng:
            tg0 = _submit(eng, good[0])
            tb = _submit(eng, bad)
            tg1 = _submit(eng, good[1])
            with pytest.raises(kwhip.KWError) as ei:
                eng.wait(tb)
            r0, r1 = eng.wait(tg0), eng.wait(tg1)
            st = eng.stats()
    assert ei.value.code == _abi.KW_ENUMERIC, ei.value
    assert "non-finite" in str(ei.value)
    assert st["failed"] == 1 and st["completed"] == 2
    _same(r0, _alone(good[0], backend))
    _same(r1, _alone(good[1], backend))


def test_engine_isolates_diverging_request_emu(few_pencils):
    _isolation("emu")


@pytest.mark.gpu
def test_engine_isolates_diverging_request_hip(hip_lib):
    print("[engine] one diverging request between two stable ones", flush=True)
    _isolation("hip")


def _as_mix(backend, steps_per_round):
    """Axisymmetric problems ('WSWA-FFT' and 'WSWS-FFT', with p0, a source
    and time reversal) submitted beside a 2-D problem of the same expanded
    size: each result equals kspaceFirstOrderAS run alone."""
    import kwhip
    from kwhip import _abi
    import as_cases as ac
    css = [ac.case("hetero"), ac.case("source_additive", sym="WSWS-FFT"),
           ac.case("time_reversal"), ac.case("nonlinear", sym="WSWS-FFT"),
           ac.case("stokes_nonlinear")]
    plain = cases.case("p0_hetero_absorbing", 2, (32, 64))
    with _abi._backend_override(backend), kwhip.Engine(steps_per_round=steps_per_round) as eng:
        tickets = []
        for cs in css:
            kg = kwhip.kWaveGrid(cs["Nx"], cs["dx"], cs["Ny"], cs["dx"])
            kg.setTime(cs["Nt"], cs["dt"])
            tickets.append(eng.submit_as(kg, dict(cs["medium"]), dict(cs["source"]),
                                         dict(cs["sensor"]), PMLSize=cs["pml"],
                                         Smooth=cs["smooth"], RadialSymmetry=cs["sym"]))
        tp = _submit(eng, plain)
        got = [eng.wait(t) for t in tickets]
        got_plain = eng.wait(tp)
        st = eng.stats()
    assert st["completed"] == len(css) + 1 and st["failed"] == 0
    with _abi._backend_override(backend):
        for cs, g in zip(css, got):
            alone = ac.run_kwhip(cs)
            if "time_reversal_boundary_data" in cs["sensor"]:
                g = {"p_final": g}
            _same(g, alone)
    _same(got_plain, _alone(plain, backend))


def test_engine_axisymmetric_emu(few_pencils):
    _as_mix("emu", 5)
