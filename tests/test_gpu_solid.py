"""GPU parity of solid cells (the SOLID builds of EulerRoe and EulerBurgers:
res/EulerRoe.cl:48-65, res/Roe.cl:38-44, 110-124, 235-290,
res/CalcFluxDeriv.cl:28-30, res/EulerBurgers.cl:42-47, 119-127, 209-236,
307-309, 372-400, 425-447) against the oracle, through the C ABI.

Masks: random cells (ghost cells included) and a disk, so walls of every
orientation, solid cells two over, solid pairs and solid ghost cells occur.

Tolerances (stated per test):
  * FP64: max |gpu - oracle| <= 1e-10 * max |oracle| after several updates
    (the GPU forms the Roe averages from (rho, v, E) per face and fuses
    a*b+c), dt relative 1e-12;
  * FP32-mixed against the FP64 GPU path: relative 5e-4;
  * graph replays against direct updates, HLL/HLLC with and without a mask,
    a cleared mask against none: bit-identical.
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
    vx = 0.25 * np.cos(k * y) + 0.1
    vy = -0.15 * np.sin(k * x) if d.dim > 1 else None
    vz = 0.1 * np.sin(k * (x + y)) if d.dim > 2 else None
    p = 1.0 + 0.1 * np.cos(k * x)
    return ics.build_state_euler(d, rho, vx, vy, vz, p)


def _mask(d, seed, kind="random"):
    x, y, z, _ = ics.grid(d)
    if kind == "disk":
        r2 = (x - 0.45) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2 * (d.dim > 2)
        return (r2 < 0.04).astype(np.uint8).ravel()
    rng = np.random.default_rng(seed)
    return (rng.random(config.volume(d)) < 0.12).astype(np.uint8)


def _relerr(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _pair(d, q, m, pot=None):
    from hydrodynamicsgpu_amd import capi

    g = capi.Sim(d)
    g.set_state(q)
    o = oracle.OracleSim(d)
    o.set_state(q)
    if pot is not None:
        g.set_potential(pot)
        o.set_potential(pot)
    g.set_solid(m)
    o.set_solid(m)
    return g, o


def _run(g, o, n, rel=1e-10):
    for _ in range(n):
        dg = g.update()
        do = o.update()
        assert dg == pytest.approx(do, rel=1e-12)
    assert np.all(np.isfinite(o.get_state()))
    assert _relerr(g.get_state(), o.get_state()) < rel


@pytest.mark.parametrize("size", [[40], [32, 28], [16, 14, 18]])
@pytest.mark.parametrize("integ,lim", [("ForwardEuler", "Superbee"), ("RungeKutta2TVD", "MonotizedCentral"),
                                       ("RungeKutta4", "MinMod"), ("RungeKutta3TVD", "Sweby")])
@pytest.mark.parametrize("kind", ["random", "disk"])
def test_roe_solid_matches_oracle(gpu_lib, size, integ, lim, kind):
    d = _desc(size, "EulerRoe", integratorName=integ, slopeLimiter=lim, cfl=0.4)
    g, o = _pair(d, _state(d, 1), _mask(d, 2, kind))
    _run(g, o, 4)


@pytest.mark.parametrize("size", [[40], [32, 28], [16, 14, 18]])
@pytest.mark.parametrize("integ,lim", [("ForwardEuler", "Superbee"), ("RungeKutta2TVD", "MonotizedCentral"),
                                       ("RungeKutta4", "MinMod")])
@pytest.mark.parametrize("kind", ["random", "disk"])
def test_burgers_solid_matches_oracle(gpu_lib, size, integ, lim, kind):
    d = _desc(size, "EulerBurgers", integratorName=integ, slopeLimiter=lim, cfl=0.4)
    g, o = _pair(d, _state(d, 3), _mask(d, 4, kind))
    _run(g, o, 4)


@pytest.mark.parametrize("solver,size,bm", [("EulerRoe", [36], "MIRROR"), ("EulerRoe", [30, 26], "PERIODIC"),
                                            ("EulerRoe", [14, 16, 12], "FREEFLOW"),
                                            ("EulerBurgers", [30, 26], "MIRROR"),
                                            ("EulerBurgers", [14, 12, 16], "PERIODIC")])
def test_solid_with_potential_and_self_gravity(gpu_lib, solver, size, bm):
    d = _desc(size, solver, bm, integratorName="RungeKutta2", cfl=0.3)
    x, y, z, _ = ics.grid(d)
    pot = (0.3 * x - 0.2 * y + 0.1 * z).ravel()
    g, o = _pair(d, _state(d, 5), _mask(d, 6), pot)
    _run(g, o, 3)
    d = _desc(size, solver, bm, integratorName="ForwardEuler", cfl=0.3, useGravity=True, gaussSeidelMaxIter=4)
    q = _state(d, 7).reshape(-1, d.dim + 2)
    q[:, -1] += 1.25 * q[:, 0]  # keeps the pressure positive with the -rho guess in the energy
    g, o = _pair(d, q.ravel(), _mask(d, 8), np.zeros(config.volume(d)))
    _run(g, o, 3)
    np.testing.assert_allclose(g.get_potential(), o.get_potential(), rtol=0,
                               atol=1e-10 * np.max(np.abs(o.get_potential())))


@pytest.mark.parametrize("solver", ["EulerHLL", "EulerHLLC"])
def test_hll_family_ignores_the_mask(gpu_lib, solver):
    from hydrodynamicsgpu_amd import capi

    d = _desc([40, 30], solver, integratorName="RungeKutta2", cfl=0.4)
    q = _state(d, 9)
    a = capi.Sim(d)
    a.set_state(q)
    b = capi.Sim(d)
    b.set_state(q)
    b.set_solid(_mask(d, 10))
    for _ in range(3):
        a.update()
        b.update()
    np.testing.assert_array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("solver", ["EulerRoe", "EulerBurgers"])
def test_cleared_mask_restores_the_plain_update(gpu_lib, solver):
    from hydrodynamicsgpu_amd import capi

    d = _desc([48, 40], solver, integratorName="RungeKutta4", cfl=0.4)
    q = _state(d, 11)
    a = capi.Sim(d)
    a.set_state(q)
    b = capi.Sim(d)
    b.set_state(q)
    b.set_solid(_mask(d, 12))
    b.set_solid(None)
    for _ in range(3):
        a.update()
        b.update()
    np.testing.assert_array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("solver,size", [("EulerRoe", [48, 40]), ("EulerRoe", [20, 18, 22]),
                                         ("EulerBurgers", [48, 40]), ("EulerRoe", [200])])
def test_graph_replays_match_direct_updates(gpu_lib, solver, size):
    from hydrodynamicsgpu_amd import capi

    d = _desc(size, solver, "PERIODIC", integratorName="RungeKutta3TVD", cfl=0.4)
    q = _state(d, 13)
    m = _mask(d, 14, "disk")
    a = capi.Sim(d)
    a.set_state(q)
    a.set_solid(m)
    b = capi.Sim(d)
    b.set_state(q)
    b.set_solid(m)
    for n in (4, 3, 5):
        for _ in range(n):
            a.update()
        b.run(n)
        b.sync()
        np.testing.assert_array_equal(a.get_state(), b.get_state())


@pytest.mark.parametrize("solver,size", [("EulerRoe", [40, 36]), ("EulerRoe", [18, 16, 20]),
                                         ("EulerBurgers", [40, 36])])
@pytest.mark.parametrize("kahan", [False, True])
def test_fp32_mixed_tracks_fp64(gpu_lib, solver, size, kahan):
    from hydrodynamicsgpu_amd import capi

    flags = config.FLAG_REFERENCE | (config.FLAG_KAHAN if kahan else 0)
    d64 = _desc(size, solver, integratorName="RungeKutta2TVD", cfl=0.4)
    d32 = _desc(size, solver, integratorName="RungeKutta2TVD", cfl=0.4, precision="fp32_mixed", flags=flags)
    q = _state(d64, 15)
    m = _mask(d64, 16, "disk")
    a = capi.Sim(d64)
    a.set_state(q)
    a.set_solid(m)
    b = capi.Sim(d32)
    b.set_state(q)
    b.set_solid(m)
    for _ in range(4):
        a.update()
        b.update()
    assert _relerr(b.get_state(), a.get_state()) < 5e-4


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_solid_mask_reads_back(gpu_lib, precision):
    from hydrodynamicsgpu_amd import capi

    d = _desc([40, 36], "EulerRoe", precision=precision)
    g = capi.Sim(d)
    assert not g.get_solid().any()
    m = _mask(d, 17)
    g.set_solid(m * 3)  # any nonzero value is solid
    np.testing.assert_array_equal(g.get_solid(), (m != 0).astype(np.uint8))
    g.set_solid(None)
    assert not g.get_solid().any()


def test_refusals(gpu_lib):
    from hydrodynamicsgpu_amd import capi

    d = _desc([32, 30], "SRHDRoe")
    g = capi.Sim(d)
    with pytest.raises(capi.HydroError, match="Euler"):
        g.set_solid(np.zeros(config.volume(d), np.uint8))
    d = _desc([32, 30], "EulerRoe")
    g = capi.Sim(d)
    with pytest.raises(capi.HydroError, match="volume"):
        g.set_solid(np.zeros(10, np.uint8))
