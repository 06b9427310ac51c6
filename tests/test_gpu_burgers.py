"""GPU parity of the EulerBurgers path (src/Solver/EulerBurgers.cpp,
res/EulerBurgers.cl) against the oracle, through the C ABI.

The GPU step is the reference's three-pass split: the advection pass in the
fused stage kernels (BurgersPolicy), the boundary, the momentum pass
(pressure of each stage state, central pressure gradient), the boundary,
the work pass with the pressure the last momentum stage left, the boundary.

Tolerances (stated per test):
  * FP64: max |gpu - oracle| <= 1e-11 * max |oracle| after several updates
    (the GPU multiplies by 1/dx where the reference divides, fuses a*b+c,
    and applies the constant work derivative once as c dt W instead of
    through every Runge-Kutta row: roundings differ at the 1e-16 level);
  * CFL dt: relative 1e-13;
  * FP32-mixed against the FP64 GPU path: relative 3e-4 on smooth flow
    (plain and compensated storage);
  * graph replays (run) against direct updates: bit-identical.
"""
import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu


def _desc(size, **kw):
    kw.setdefault("boundaryMethods", [{"min": "PERIODIC", "max": "PERIODIC"}] * 3)
    return config.make_desc(size, solverName="EulerBurgers", **kw)


def _smooth(d, seed=0, amp=0.3):
    """A smooth periodic flow with every velocity component non-zero."""
    rng = np.random.default_rng(seed)
    x, y, z, _ = ics.grid(d)
    ph = rng.uniform(0, 2 * np.pi, 6)
    k = 2 * np.pi
    rho = 1.0 + amp * np.sin(k * x + ph[0]) * np.cos(k * y + ph[1]) + 0.1 * np.sin(k * z + ph[2])
    vx = 0.4 * np.sin(k * y + ph[3]) + 0.2
    vy = -0.3 * np.cos(k * x + ph[4]) if d.dim > 1 else np.zeros_like(x)
    vz = 0.25 * np.sin(k * (x + y) + ph[5]) if d.dim > 2 else np.zeros_like(x)
    p = 1.0 + 0.2 * np.cos(k * x - ph[1]) * np.sin(k * z + ph[0])
    return ics.build_state_euler(d, rho, vx, vy, vz, p)


def _rand(d, seed=0):
    rng = np.random.default_rng(seed)
    x, y, z, _ = ics.grid(d)
    shp = x.shape
    return ics.build_state_euler(d, rng.uniform(0.5, 2.0, shp), rng.uniform(-0.6, 0.6, shp),
                                 rng.uniform(-0.6, 0.6, shp), rng.uniform(-0.6, 0.6, shp),
                                 rng.uniform(0.5, 2.0, shp))


def _relerr(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def _pair(d, q):
    from hydrodynamicsgpu_amd import capi

    g = capi.Sim(d)
    g.set_state(q)
    o = oracle.OracleSim(d)
    o.set_state(q)
    return g, o


def _updates(g, o, n, rtol_dt=1e-13):
    for _ in range(n):
        dg = g.update()
        do = o.update()
        assert abs(dg - do) <= rtol_dt * abs(do), (dg, do)


@pytest.mark.parametrize("size", [[40], [36, 28], [20, 18, 16]])
@pytest.mark.parametrize("integ", ["ForwardEuler", "RungeKutta2TVD", "RungeKutta4", "RungeKutta3TVD"])
def test_fp64_updates_match_oracle(gpu_lib, size, integ):
    d = _desc(size, integratorName=integ, slopeLimiter="Superbee", cfl=0.5)
    g, o = _pair(d, _smooth(d, 1))
    _updates(g, o, 5)
    assert _relerr(g.get_state(), o.get_state()) < 1e-11
    assert g.health() == 0


@pytest.mark.parametrize("lim", ["MonotizedCentral", "MinMod", "VanLeer", "DonorCell", "LaxWendroff", "Koren"])
@pytest.mark.parametrize("size", [[48], [34, 30], [18, 16, 20]])
def test_fp64_limiters_match_oracle(gpu_lib, size, lim):
    d = _desc(size, integratorName="RungeKutta2", slopeLimiter=lim, cfl=0.4)
    g, o = _pair(d, _smooth(d, 2))
    _updates(g, o, 4)
    assert _relerr(g.get_state(), o.get_state()) < 1e-11


@pytest.mark.parametrize("integ", ["RungeKutta2Heun", "RungeKutta2Ralston", "RungeKutta3", "RungeKutta4_3_8thsRule",
                                   "RungeKutta2NonTVD", "RungeKutta4TVD", "RungeKutta4NonTVD", "RungeKutta2"])
def test_fp64_every_tableau_matches_oracle(gpu_lib, integ):
    d = _desc([30, 26], integratorName=integ, slopeLimiter="Superbee", cfl=0.3)
    g, o = _pair(d, _smooth(d, 3, amp=0.1))
    _updates(g, o, 3)
    assert _relerr(g.get_state(), o.get_state()) < 1e-11


@pytest.mark.parametrize("bm", ["MIRROR", "FREEFLOW", "NONE"])
@pytest.mark.parametrize("size", [[40], [36, 28], [18, 16, 16]])
def test_fp64_boundaries_match_oracle(gpu_lib, size, bm):
    d = _desc(size, integratorName="RungeKutta2TVD", slopeLimiter="Superbee", cfl=0.4,
              boundaryMethods=[{"min": bm, "max": bm}] * 3)
    g, o = _pair(d, _rand(d, 4) if bm != "NONE" else _smooth(d, 4, amp=0.1))
    _updates(g, o, 4)
    assert _relerr(g.get_state(), o.get_state()) < 1e-11


def test_fp64_mixed_boundaries_match_oracle(gpu_lib):
    d = _desc([40, 30], integratorName="RungeKutta4", slopeLimiter="MonotizedCentral", cfl=0.4,
              boundaryMethods=[{"min": "MIRROR", "max": "FREEFLOW"}, {"min": "PERIODIC", "max": "PERIODIC"}])
    g, o = _pair(d, _rand(d, 5))
    _updates(g, o, 4)
    assert _relerr(g.get_state(), o.get_state()) < 1e-11


def test_sod_tube_matches_oracle_long_run(gpu_lib):
    """1D Sod shock tube, Superbee, TVD-RK2, free-flow: 200 updates (shock,
    contact and rarefaction all on the grid) against the oracle."""
    d = _desc([256], xmin=[-1.0], xmax=[1.0], integratorName="RungeKutta2TVD", slopeLimiter="Superbee",
              cfl=0.5, boundaryMethods=[{"min": "FREEFLOW", "max": "FREEFLOW"}])
    x, y, z, _ = ics.grid(d)
    left = x <= 0
    q = ics.build_state_euler(d, np.where(left, 1.0, 0.125), np.zeros_like(x), None, None,
                              np.where(left, 1.0, 0.1))
    g, o = _pair(d, q)
    _updates(g, o, 200, rtol_dt=1e-12)
    assert _relerr(g.get_state(), o.get_state()) < 1e-10


def test_calc_timestep_and_fixed_dt_step(gpu_lib):
    d = _desc([30, 22], integratorName="RungeKutta4", cfl=0.45)
    q = _smooth(d, 6)
    g, o = _pair(d, q)
    o.init_step()
    dt_o = o.calc_timestep()
    dt_g = g.calc_timestep()
    assert dt_g == pytest.approx(dt_o, rel=1e-13)
    g.step(0.9 * dt_o)
    o.step(0.9 * dt_o)
    assert _relerr(g.get_state(), o.get_state()) < 1e-11


def test_run_graphs_bit_identical_to_updates(gpu_lib):
    from hydrodynamicsgpu_amd import capi

    d = _desc([66, 40], integratorName="RungeKutta4", slopeLimiter="Superbee", cfl=0.5)
    q = _smooth(d, 7)
    a = capi.Sim(d)
    a.set_state(q)
    b = capi.Sim(d)
    b.set_state(q)
    for n in (4, 3, 5):
        for _ in range(n):
            a.update()
        b.run(n)
        b.sync()
        np.testing.assert_array_equal(a.get_state(), b.get_state())
        assert a.time() == b.time()


@pytest.mark.parametrize("size", [[40], [36, 28], [18, 16, 20]])
@pytest.mark.parametrize("kahan", [False, True])
def test_fp32_mixed_tracks_fp64(gpu_lib, size, kahan):
    from hydrodynamicsgpu_amd import capi

    flags = config.FLAG_REFERENCE | (config.FLAG_KAHAN if kahan else 0)
    d64 = _desc(size, integratorName="RungeKutta2TVD", cfl=0.4)
    d32 = _desc(size, integratorName="RungeKutta2TVD", cfl=0.4, precision="fp32_mixed", flags=flags)
    q = _smooth(d64, 8, amp=0.2)
    a = capi.Sim(d64)
    a.set_state(q)
    b = capi.Sim(d32)
    b.set_state(q)
    for _ in range(6):
        da = a.update()
        db = b.update()
        assert db == pytest.approx(da, rel=1e-5)
    assert _relerr(b.get_state(), a.get_state()) < 3e-4
    assert b.health() == 0


def test_generic_entry_points_refuse_burgers(gpu_lib):
    from hydrodynamicsgpu_amd import capi

    d = _desc([30])
    g = capi.Sim(d)
    g.set_state(_smooth(d, 9))
    with pytest.raises(capi.HydroError, match="EulerBurgers"):
        g.init_step()


def test_batch_refuses_burgers(gpu_lib):
    from hydrodynamicsgpu_amd import capi

    d = _desc([64])
    with pytest.raises(capi.HydroError):
        capi.Batch([d, d])
