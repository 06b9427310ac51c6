"""GPU parity of the per-step parts outside the fused stages, through the C ABI:
the one-launch CFL reduction (strip march with the reduction tail fused in),
the one-launch composed boundary, and the deferred copy of the new state's
ghost cells.

Tolerances: dt relative 1e-13 (FP64; the reduction order differs from the
reference's, the per-cell arithmetic does not); ghost cells bit-identical;
states after several manual boundary/calc_timestep/step calls within 1e-10
relative of the oracle (the fused-stage tolerance).
"""
import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu


def _mk(desc):
    from hydrodynamicsgpu_amd import capi

    return capi.Sim(desc)


def _rand_state(desc, seed=0, vscale=1.0):
    rng = np.random.default_rng(seed)
    x, _, _, _ = ics.grid(desc)
    shp = x.shape
    return ics.build_state_euler(desc, rng.uniform(0.5, 2.0, shp), vscale * rng.uniform(-0.8, 0.8, shp),
                                 vscale * rng.uniform(-0.8, 0.8, shp), np.zeros(shp), rng.uniform(0.5, 2.0, shp))


# sizes straddling the 60-column strips (64 = 60 interior columns exactly,
# 65/124/125 leave one column over), uneven row chunks, and 3D grids (the 3D
# march: planes x row chunks)
SIZES = [[64], [65], [124], [300], [64, 37], [125, 64], [124, 130], [61, 300], [24, 20, 18], [70, 33, 21],
         [130, 9, 40]]


@pytest.mark.parametrize("solver", ["Roe", "HLL", "HLLC"])
@pytest.mark.parametrize("size", SIZES)
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_cfl_matches_oracle(gpu_lib, solver, size, precision):
    d = config.make_desc(size, solverName="Euler" + solver, cfl=0.45, precision=precision)
    d64 = config.make_desc(size, solverName="Euler" + solver, cfl=0.45)
    q = _rand_state(d64, 11)
    g = _mk(d)
    g.set_state(q)
    g.boundary()
    o = oracle.OracleSim(d64)
    o.set_state(q if precision == "fp64" else g.get_state())
    o.boundary()
    o.init_step()  # HLL::initStep freezes the wave speeds the CFL reads
    # mixed: the CFL runs on the FP32 state in FP32 speeds, FP64 division
    rel = 1e-13 if precision == "fp64" else 5e-6
    assert g.calc_timestep() == pytest.approx(o.calc_timestep(), rel=rel)


@pytest.mark.parametrize("solver", ["Roe", "HLL", "HLLC"])
@pytest.mark.parametrize("size", [[300], [125, 64], [70, 33, 21]])
def test_cfl_skips_nan_speeds_like_the_reference(gpu_lib, solver, size):
    # cells with negative pressure have NaN sound speeds (per cell for
    # HLL/HLLC, where the Roe average is negative too for Roe); min/max with
    # NaN-skipping semantics absorb some (HLL's speed bounds), the reference's
    # max(0, .) / min(0, .) read the rest as 0, and its min() over the
    # per-cell dt counts every other cell (res/Roe.cl:51-57,
    # res/Common.cl:9-40).  The GPU minimum must be the oracle's.
    d = config.make_desc(size, solverName="Euler" + solver, cfl=0.5)
    q = _rand_state(d, 11)
    n = config.num_states(d)
    cells = config.volume(d)
    for c in (cells // 3, cells // 2 + 7, (2 * cells) // 3 + 1):
        q[n * c + n - 1] = -1.0  # negative total energy: negative pressure
    g = _mk(d)
    g.set_state(q)
    g.boundary()
    o = oracle.OracleSim(d)
    o.set_state(q)
    o.boundary()
    o.init_step()
    dt_o = o.calc_timestep()
    assert np.isfinite(dt_o)
    assert g.calc_timestep() == pytest.approx(dt_o, rel=1e-13)


def test_cfl_accumulators_reset_between_calls(gpu_lib):
    # the fused reduction leaves its accumulators reset: a smaller-speed state
    # after a larger-speed one gets its own (larger) dt, not the old minimum
    d = config.make_desc([200, 90], cfl=0.5)
    g = _mk(d)
    q1 = _rand_state(d, 1)
    q2 = _rand_state(d, 1, vscale=0.1)  # same density/pressure, slower flow
    dts = []
    for q in (q1, q2, q1):
        g.set_state(q)
        g.boundary()
        o = oracle.OracleSim(d)
        o.set_state(q)
        o.boundary()
        o.init_step()
        dg = g.calc_timestep()
        assert dg == pytest.approx(o.calc_timestep(), rel=1e-13)
        dts.append(dg)
    assert dts[0] == dts[2] and dts[1] > dts[0]


def test_calc_timestep_clips_against_the_clock(gpu_lib):
    # with HYDRO_FLAG_CLIP_TO_TEND, calc_timestep clips against the
    # simulation's own time (it leaves the clock as it is): the dt it returns
    # is the one the next update takes; at t_end it is 0 and updates stop
    # counting
    d = config.make_desc([256], xmin=[-1], xmax=[1], integratorName="RungeKutta2TVD", cfl=0.5, t_end=0.01,
                         flags=config.FLAG_REFERENCE | config.FLAG_CLIP_TO_TEND)
    g = _mk(d)
    g.set_state(ics.sod(d))
    while True:
        t, steps, _ = g.time()
        g.boundary()
        dt = g.calc_timestep()
        assert g.time()[:2] == (t, steps)  # the clock did not move
        assert t + dt <= 0.01 + 1e-15
        if dt == 0.0:
            break
        assert g.update() == dt
    assert g.time()[0] == 0.01
    n = g.time()[1]
    assert g.update() == 0.0 and g.time()[1] == n and g.health() == 0


@pytest.mark.parametrize("bm", [["PERIODIC", "MIRROR"], ["FREEFLOW", "PERIODIC"], ["MIRROR", "FREEFLOW"]])
def test_manual_step_sequence_matches_oracle(gpu_lib, bm):
    # boundary / calc_timestep / step called one by one, including two steps
    # without a boundary in between (the second step then sees the ghost
    # cells the first one carried over, as in the reference).  Every step
    # follows a calc_timestep: the frozen eigenbasis is the one of the state
    # the step starts from on both sides.
    d = config.make_desc([70, 46], integratorName="RungeKutta2", slopeLimiter="MonotizedCentral",
                         boundaryMethods=[{"min": bm[0], "max": bm[0]}, {"min": bm[1], "max": bm[1]}], cfl=0.4)
    q = _rand_state(d, 4)
    g = _mk(d)
    g.set_state(q)
    o = oracle.OracleSim(d)
    o.set_state(q)
    for obj in (g, o):
        obj.boundary()
        dt = obj.calc_timestep()  # Roe: init_step has nothing to freeze
        obj.step(dt)
        obj.calc_timestep()
        obj.step(0.5 * dt)
        obj.boundary()
        dt = obj.calc_timestep()
        obj.step(dt)
    a, b = g.get_state(), o.get_state()
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < 1e-10


def test_state_after_run_includes_carried_ghosts(gpu_lib):
    # after a run of updates the returned ghost cells are those the last
    # step started from (boundary of the previous state), bit-identical
    d = config.make_desc([90, 70], integratorName="RungeKutta4", slopeLimiter="Superbee",
                         boundaryMethods=[{"min": "MIRROR", "max": "FREEFLOW"}, {"min": "PERIODIC", "max": "PERIODIC"}])
    q = _rand_state(d, 6)
    g = _mk(d)
    g.set_state(q)
    o = oracle.OracleSim(d)
    o.set_state(q)
    g.run(3)
    for _ in range(3):
        o.update()
    a = g.get_state().reshape(70, 90, 4)
    b = o.get_state().reshape(70, 90, 4)
    ghost = np.ones((70, 90), bool)
    ghost[2:-2, 2:-2] = False
    assert np.max(np.abs(a[ghost] - b[ghost])) <= 1e-10 * np.max(np.abs(b))
    assert np.max(np.abs(a - b)) / np.max(np.abs(b)) < 1e-10


@pytest.mark.parametrize("integ,bm,precision", [("RungeKutta4", "PERIODIC", "fp32_mixed"),
                                                 ("RungeKutta3TVD", "MIRROR", "fp64"),
                                                 ("ForwardEuler", "FREEFLOW", "fp64")])
@pytest.mark.parametrize("graph_updates", ["2", "8"])
def test_graph_replays_match_direct_launches(gpu_lib, monkeypatch, integ, bm, precision, graph_updates):
    # run() replays captured runs of updates (graph_updates per graph, then
    # pairs), for either starting buffer parity; several replays, with eager
    # work and synchronisation in between, must equal direct launches bit
    # for bit
    d = config.make_desc([90, 66], integratorName=integ, slopeLimiter="MonotizedCentral",
                         boundaryMethods=[{"min": bm, "max": bm}] * 2, cfl=0.4, precision=precision)
    q = _rand_state(d, 8)
    monkeypatch.setenv("HYDRO_GRAPHS", "0")
    direct = _mk(d)
    monkeypatch.setenv("HYDRO_GRAPHS", "1")
    monkeypatch.setenv("HYDRO_GRAPH_UPDATES", graph_updates)
    graphed = _mk(d)
    direct.set_state(q)
    graphed.set_state(q)
    for n in (4, 3, 2, 5, 11, 17):
        direct.run(n)
        graphed.run(n)
        graphed.sync()
        a, b = graphed.get_state(), direct.get_state()
        np.testing.assert_array_equal(a, b)
        assert graphed.time() == direct.time()
        graphed.update()  # eager work between replays
        direct.update()


@pytest.mark.parametrize("family,precision", [("srhd", "fp64"), ("srhd", "fp32_mixed"),
                                              ("maxwell", "fp64"), ("maxwell", "fp32_mixed")])
@pytest.mark.parametrize("graph_updates", ["2", "8"])
def test_graph_replays_match_direct_launches_srhd_maxwell(gpu_lib, monkeypatch, family, precision, graph_updates):
    # SRHD runs its deferred ghost copy inside the captured con2prim launch
    # (no ghosts pending after a captured pair), Maxwell has its own mirror
    # sign masks: replays from either buffer parity, with eager updates and
    # state reads in between, equal direct launches bit for bit, and so do
    # the primitives
    from srhd_states import quadrants_for

    if family == "srhd":
        d = config.make_desc([70, 52], solverName="SRHDRoe", integratorName="RungeKutta2", xmin=[0, 0], xmax=[1, 1],
                             boundaryMethods=[{"min": "FREEFLOW", "max": "FREEFLOW"}] * 2, cfl=0.4,
                             defs={"idealGas_heatCapacityRatio": 5.0 / 3.0}, precision=precision)
        q = quadrants_for(d)
    else:
        d = config.make_desc([70, 52], solverName="MaxwellRoe", integratorName="RungeKutta4",
                             boundaryMethods=[{"min": "MIRROR", "max": "MIRROR"}, {"min": "PERIODIC", "max": "MIRROR"}],
                             cfl=0.4, precision=precision)
        q = ics.maxwell_1(d) + 0.1 * np.random.default_rng(3).standard_normal(6 * config.volume(d))
    monkeypatch.setenv("HYDRO_GRAPHS", "0")
    direct = _mk(d)
    monkeypatch.setenv("HYDRO_GRAPHS", "1")
    monkeypatch.setenv("HYDRO_GRAPH_UPDATES", graph_updates)
    graphed = _mk(d)
    direct.set_state(q)
    graphed.set_state(q)
    for n in (4, 3, 2, 5, 11, 17):
        direct.run(n)
        graphed.run(n)
        graphed.sync()
        np.testing.assert_array_equal(graphed.get_state(), direct.get_state())
        if family == "srhd":
            np.testing.assert_array_equal(graphed.get_primitives(), direct.get_primitives())
        assert graphed.time() == direct.time()
        graphed.update()
        direct.update()
    assert graphed.health() == direct.health() == 0


@pytest.mark.parametrize("solver,integ,phases", [
    ("EulerRoe", "RungeKutta4", {"head": 1, "dt_tail": 1, "stages": 4}),
    ("EulerHLLC", "RungeKutta3TVD", {"head": 1, "dt_tail": 1, "stages": 3}),
    ("EulerBurgers", "RungeKutta2", {"boundary": 4, "cfl": 1, "stages": 2, "pressure": 2}),
    ("SRHDRoe", "RungeKutta2", {"head": 1, "dt_tail": 1, "stages": 2, "con2prim": 1}),
    ("MaxwellRoe", "ForwardEuler", {"boundary": 1, "cfl": 1, "stages": 1}),
])
def test_phase_timing_counts_every_phase(gpu_lib, solver, integ, phases):
    """Per-phase device timing: each phase of an update is one interval per
    launch group, with a positive duration; untimed phases stay at zero."""
    from hydrodynamicsgpu_amd import capi

    size = [64, 48]
    kw = {"xmin": [0.0, 0.0], "xmax": [1.0, 1.0]} if solver == "SRHDRoe" else {}
    d = config.make_desc(size, solverName=solver, integratorName=integ, cfl=0.4,
                         boundaryMethods=[{"min": "FREEFLOW", "max": "FREEFLOW"}] * 2, **kw)
    if solver == "SRHDRoe":
        q = ics.srhd_quadrants(d)
    elif solver == "MaxwellRoe":
        q = ics.maxwell_1(d)
    else:
        q = _rand_state(d, 3)
    g = capi.Sim(d)
    g.set_state(q)
    g.set_profiling(True)
    n = 3
    for _ in range(n):
        g.update()
    g.totals()
    t = g.phase_timing()
    expect = dict(phases)
    expect["totals"] = 1 / n
    for name in capi.PHASES:
        ms, cnt = t[name]
        want = round(expect.get(name, 0) * n)
        if name == "pressure" and solver == "EulerBurgers":
            want = n * 2  # momentum pass and work pass
        if name == "boundary" and solver == "EulerBurgers":
            want = n * 4  # the update's own and one after each pass
        assert cnt == want, (name, cnt, want, t)
        assert (ms > 0) == (want > 0), (name, ms)
    ms, launches = g.stage_timing()
    assert launches == t["stages"][1] and ms > 0
    g.set_profiling(False)
    assert all(v == (0.0, 0) for v in g.phase_timing().values())
