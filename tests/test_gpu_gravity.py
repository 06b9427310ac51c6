"""GPU parity of self-gravity and the gravitational potential
(src/Solver/SelfGravitation.cpp, res/SelfGravitation.cl, the potential in
res/EulerRoe.cl, res/EulerHLL.cl and res/EulerBurgers.cl) against the oracle,
through the C ABI.

Cases: a fixed non-zero potential (the gravity pass with the potential's
source, the Roe / HLL / Burgers kernels reading it, HLLC not), and
useGravity (the potential relaxed from the density at reset and in every
stage of the gravity pass).

Tolerances (stated per test):
  * FP64: max |gpu - oracle| <= 1e-10 * max |oracle| after several updates,
    potential included (the GPU multiplies by 1/(2 dx) where the reference
    divides, and fuses a*b+c);
  * FP32-mixed against the FP64 GPU path: relative 5e-4;
  * graph replays (run) against direct updates: bit-identical.
"""
import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu


def _desc(size, solver, bm="FREEFLOW", **kw):
    kw.setdefault("xmin", [0.0, 0.0, 0.0])
    kw.setdefault("xmax", [1.0, 1.0, 1.0])
    return config.make_desc(size, solverName=solver, boundaryMethods=[{"min": bm, "max": bm}] * 3, **kw)


def _state(d, seed=0):
    rng = np.random.default_rng(seed)
    x, y, z, _ = ics.grid(d)
    k = 2 * np.pi
    rho = 1.0 + 0.3 * np.sin(k * x) * np.cos(k * y) + 0.05 * rng.random(x.shape)
    vx = 0.2 * np.cos(k * y)
    vy = -0.15 * np.sin(k * x) if d.dim > 1 else None
    vz = 0.1 * np.sin(k * (x + y)) if d.dim > 2 else None
    p = 1.0 + 0.1 * np.cos(k * x)
    return ics.build_state_euler(d, rho, vx, vy, vz, p)


def _potential(d):
    x, y, z, _ = ics.grid(d)
    return (0.3 * x - 0.2 * y + 0.1 * z + 0.05 * np.sin(3 * x)).ravel()


def _relerr(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _pair(d, q, pot):
    from hydrodynamicsgpu_amd import capi

    g = capi.Sim(d)
    g.set_state(q)
    g.set_potential(pot)
    o = oracle.OracleSim(d)
    o.set_state(q)
    o.set_potential(pot)
    return g, o


@pytest.mark.parametrize("solver,size", [
    ("EulerRoe", [40]), ("EulerRoe", [32, 28]), ("EulerRoe", [16, 14, 18]),
    ("EulerHLL", [40]), ("EulerHLL", [32, 28]),
    ("EulerHLLC", [32, 28]), ("EulerHLLC", [16, 14, 18]),
    ("EulerBurgers", [40]), ("EulerBurgers", [32, 28]), ("EulerBurgers", [16, 14, 18]),
])
@pytest.mark.parametrize("integ", ["ForwardEuler", "RungeKutta2TVD", "RungeKutta4"])
def test_fixed_potential_matches_oracle(gpu_lib, solver, size, integ):
    d = _desc(size, solver, integratorName=integ, cfl=0.4)
    g, o = _pair(d, _state(d, 1), _potential(d))
    assert _relerr(g.get_state(), o.get_state()) < 1e-13  # resetState's E += phi
    for _ in range(4):
        dg = g.update()
        do = o.update()
        assert dg == pytest.approx(do, rel=1e-12)
    assert _relerr(g.get_state(), o.get_state()) < 1e-10
    assert g.health() == 0


@pytest.mark.parametrize("solver,size,bm", [
    ("EulerRoe", [36], "FREEFLOW"), ("EulerRoe", [30, 26], "PERIODIC"), ("EulerRoe", [30, 26], "MIRROR"),
    ("EulerRoe", [14, 16, 12], "FREEFLOW"), ("EulerHLL", [30, 26], "FREEFLOW"), ("EulerHLLC", [30, 26], "MIRROR"),
    ("EulerBurgers", [30, 26], "FREEFLOW"), ("EulerBurgers", [36], "PERIODIC"),
])
@pytest.mark.parametrize("integ", ["ForwardEuler", "RungeKutta3TVD"])
def test_self_gravity_matches_oracle(gpu_lib, solver, size, bm, integ):
    d = _desc(size, solver, bm, integratorName=integ, cfl=0.3, useGravity=True, gaussSeidelMaxIter=7)
    g, o = _pair(d, _state(d, 2), None)
    np.testing.assert_allclose(g.get_potential(), o.get_potential(), rtol=0,
                               atol=1e-11 * np.max(np.abs(o.get_potential())))
    for _ in range(3):
        dg = g.update()
        do = o.update()
        assert dg == pytest.approx(do, rel=1e-11)
    assert _relerr(g.get_state(), o.get_state()) < 1e-10
    np.testing.assert_allclose(g.get_potential(), o.get_potential(), rtol=0,
                               atol=1e-10 * np.max(np.abs(o.get_potential())))


def test_zero_potential_keeps_the_update_unchanged(gpu_lib):
    """A zero potential (and no useGravity) adds no gravity pass: the update
    is bit-identical to one without set_potential."""
    from hydrodynamicsgpu_amd import capi

    d = _desc([40, 30], "EulerRoe", integratorName="RungeKutta4", cfl=0.4)
    q = _state(d, 3)
    a = capi.Sim(d)
    a.set_state(q)
    b = capi.Sim(d)
    b.set_state(q)
    b.set_potential(np.zeros(config.volume(d)))
    for _ in range(3):
        a.update()
        b.update()
    np.testing.assert_array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("solver", ["EulerRoe", "EulerBurgers"])
def test_graph_replays_match_direct_updates(gpu_lib, solver):
    from hydrodynamicsgpu_amd import capi

    d = _desc([48, 40], solver, "PERIODIC", integratorName="RungeKutta2", cfl=0.4, useGravity=True,
              gaussSeidelMaxIter=3)
    q = _state(d, 4)
    a = capi.Sim(d)
    a.set_state(q)
    a.set_potential()
    b = capi.Sim(d)
    b.set_state(q)
    b.set_potential()
    for n in (4, 3, 5):
        for _ in range(n):
            a.update()
        b.run(n)
        b.sync()
        np.testing.assert_array_equal(a.get_state(), b.get_state())
        np.testing.assert_array_equal(a.get_potential(), b.get_potential())


@pytest.mark.parametrize("kahan", [False, True])
def test_fp32_mixed_tracks_fp64(gpu_lib, kahan):
    from hydrodynamicsgpu_amd import capi

    flags = config.FLAG_REFERENCE | (config.FLAG_KAHAN if kahan else 0)
    d64 = _desc([40, 36], "EulerRoe", integratorName="RungeKutta2TVD", cfl=0.4, useGravity=True,
                gaussSeidelMaxIter=5)
    d32 = _desc([40, 36], "EulerRoe", integratorName="RungeKutta2TVD", cfl=0.4, useGravity=True,
                gaussSeidelMaxIter=5, precision="fp32_mixed", flags=flags)
    q = _state(d64, 5)
    a = capi.Sim(d64)
    a.set_state(q)
    a.set_potential()
    b = capi.Sim(d32)
    b.set_state(q)
    b.set_potential()
    for _ in range(4):
        a.update()
        b.update()
    assert _relerr(b.get_state(), a.get_state()) < 5e-4
    assert _relerr(b.get_potential(), a.get_potential()) < 5e-4


def test_phase_timing_counts_the_gravity_pass(gpu_lib):
    from hydrodynamicsgpu_amd import capi

    d = _desc([48, 40], "EulerRoe", integratorName="RungeKutta2", cfl=0.4, useGravity=True, gaussSeidelMaxIter=2)
    g = capi.Sim(d)
    g.set_state(_state(d, 6))
    g.set_potential()
    g.set_profiling(True)
    for _ in range(3):
        g.update()
    t = g.phase_timing()
    assert t["gravity"][1] == 3 and t["gravity"][0] > 0
    assert t["cfl"][1] == 3 and t["head"][1] == 0  # the potential's speeds: the generic CFL kernel


def test_refusals(gpu_lib):
    from hydrodynamicsgpu_amd import capi

    d = _desc([16, 14, 12], "EulerHLL", useGravity=True)
    with pytest.raises(capi.HydroError, match="3D HLL"):
        capi.Sim(d)
    d = _desc([16, 14, 12], "EulerHLL")
    g = capi.Sim(d)
    g.set_state(_state(d, 7))
    with pytest.raises(capi.HydroError, match="3D HLL"):
        g.set_potential(_potential(d))
    d = _desc([32, 30], "SRHDRoe")
    g = capi.Sim(d)
    with pytest.raises(capi.HydroError, match="Euler"):
        g.set_potential(np.zeros(config.volume(d)))
    with pytest.raises(capi.HydroError, match="self-gravity"):
        capi.Batch([_desc([64], "EulerRoe", useGravity=True)] * 2)
