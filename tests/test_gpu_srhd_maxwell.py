"""GPU parity of the SRHD Roe and Maxwell Roe paths against the oracle.

Tolerances: FP64 max |gpu - oracle| <= 1e-9 * max |oracle| after several
steps for SRHD (the con2prim Newton stops at a relative change of 1e-7, so
round-off differences can move the stopping iteration; the converged
pressure then differs at the 1e-12 level, not at the 1e-7 level, because the
iteration is quadratically convergent); 1e-10 for Maxwell.  Primitives are
compared the same way.
"""
import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics
from srhd_states import quadrants_for

pytestmark = pytest.mark.gpu


def _mk(desc):
    from hydrodynamicsgpu_amd import capi

    return capi.Sim(desc)


def _relerr(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


SRHD_CASES = [
    ([128], "RungeKutta2", "Superbee", "FREEFLOW"),
    ([70, 40], "RungeKutta2", "Superbee", "FREEFLOW"),
    ([66, 50], "RungeKutta4", "MonotizedCentral", "PERIODIC"),
    ([40, 36], "ForwardEuler", "MinMod", "MIRROR"),
    ([24, 20, 22], "RungeKutta2", "Superbee", "FREEFLOW"),
]


@pytest.mark.parametrize("size,integ,lim,bm", SRHD_CASES)
def test_srhd_fp64_parity(gpu_lib, size, integ, lim, bm):
    kw = dict(xmin=[0.0] * 3, xmax=[1.0] * 3, solverName="SRHDRoe", integratorName=integ, slopeLimiter=lim,
              boundaryMethods=[{"min": bm, "max": bm}] * 3, cfl=0.4,
              defs={"idealGas_heatCapacityRatio": 5.0 / 3.0})
    d = config.make_desc(size, **kw)
    if len(size) == 1:
        q = ics.riemann(d, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    else:
        q = quadrants_for(d) if len(size) == 2 else ics.riemann(d, (1.0, 0.3, 1.0), (0.2, -0.2, 0.05))
    g = _mk(d)
    g.set_state(q)
    o = oracle.OracleSim(d)
    o.set_state(q)
    np.testing.assert_allclose(g.get_primitives(), o.get_primitives(), rtol=1e-14, atol=1e-14)
    for _ in range(5):
        assert g.update() == pytest.approx(o.update(), rel=1e-9)
    assert _relerr(g.get_state(), o.get_state()) < 1e-9
    assert _relerr(g.get_primitives(), o.get_primitives()) < 1e-9
    assert g.health() == 0


def test_srhd_mixed_quadrants(gpu_lib):
    kw = dict(xmin=[0.0] * 2, xmax=[1.0] * 2, solverName="SRHDRoe", integratorName="RungeKutta2",
              slopeLimiter="Superbee", boundaryMethods=[{"min": "FREEFLOW", "max": "FREEFLOW"}] * 2, cfl=0.4,
              defs={"idealGas_heatCapacityRatio": 5.0 / 3.0})
    d64 = config.make_desc([96, 96], **kw)
    d32 = config.make_desc([96, 96], precision="fp32_mixed", flags=config.FLAG_REFERENCE | config.FLAG_KAHAN, **kw)
    q = ics.srhd_quadrants(d64)
    o = oracle.OracleSim(d64)
    o.set_state(q)
    g = _mk(d32)
    g.set_state(q)
    for _ in range(4):
        o.update()
        g.update()
    assert _relerr(g.get_state(), o.get_state()) < 1e-3
    assert g.health() == 0


def _max_lorentz(prims, dim):
    P = np.asarray(prims).reshape(-1, dim + 2)
    v2 = (P[:, 1:1 + dim] ** 2).sum(axis=1)
    return float(1.0 / np.sqrt(1.0 - v2.max()))


def test_srhd_full_size_quadrants(gpu_lib):
    # the SRHD workload at its full size (1024^2 quadrants, RK2, gamma 5/3,
    # fastest quadrant at W = 9.14): three FP64 updates against the oracle
    # (state and primitives, 1e-9 as above), and the FP32-mixed path against
    # the FP64 one (1e-3); no con2prim failure (health bit 2) in either
    kw = dict(xmin=[0.0] * 2, xmax=[1.0] * 2, solverName="SRHDRoe", integratorName="RungeKutta2",
              slopeLimiter="Superbee", boundaryMethods=[{"min": "FREEFLOW", "max": "FREEFLOW"}] * 2,
              defs={"idealGas_heatCapacityRatio": 5.0 / 3.0})
    d64 = config.make_desc([1024, 1024], **kw)
    d32 = config.make_desc([1024, 1024], precision="fp32_mixed", **kw)
    q = ics.srhd_quadrants(d64)
    g, m = _mk(d64), _mk(d32)
    g.set_state(q)
    m.set_state(q)
    assert _max_lorentz(g.get_primitives(), 2) > 9.0
    o = oracle.OracleSim(d64)
    o.set_state(q)
    for _ in range(3):
        assert g.update() == pytest.approx(o.update(), rel=1e-9)
        m.update()
    assert _relerr(g.get_state(), o.get_state()) < 1e-9
    assert _relerr(g.get_primitives(), o.get_primitives()) < 1e-9
    assert _relerr(m.get_state(), g.get_state()) < 1e-3
    assert _max_lorentz(m.get_primitives(), 2) > 9.0
    assert g.health() == 0 and m.health() == 0


@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_srhd_lorentz_guard_keeps_primitives(gpu_lib, precision):
    # Forced from the host: a uniform conserved state D = 1, S = (4, 0),
    # tau = 1 has |S| = 2 (tau + D), beyond any physical state.  With the
    # velocity margin srhd_solvePrimVelEpsilon = 0 the Newton bracket starts
    # at P = |S| - tau - D = 2, where |v| = |S| / (tau + D + P) = 1 exactly
    # and the Lorentz factor would overflow (res/SRHDCommon.cl:163-165).
    # The recovery stops there, keeps every cell's previous primitives and
    # raises health bit 2; the uniform state has zero flux divergence, so
    # it is unchanged.
    d = config.make_desc([16, 16], xmin=[0.0] * 2, xmax=[1.0] * 2, solverName="SRHDRoe",
                         integratorName="RungeKutta2", precision=precision,
                         defs={"idealGas_heatCapacityRatio": 5.0 / 3.0, "srhd_solvePrimVelEpsilon": 0.0})
    V = config.volume(d)
    cons = np.tile([1.0, 4.0, 0.0, 1.0], V)
    prims = np.tile([0.5, 0.3, 0.0, 1.0], V)
    g = _mk(d)
    g.restore(cons, 0.0, 0, prims=prims)
    dt = g.update()
    assert np.isfinite(dt) and dt > 0
    assert g.health() & 4
    # the FP32-mixed form stores the primitives in FP32
    kept = prims.astype(np.float32).astype(np.float64) if precision == "fp32_mixed" else prims
    np.testing.assert_array_equal(g.get_primitives(), kept)
    np.testing.assert_array_equal(g.get_state(), cons)


MAXWELL_CASES = [
    ([128], "RungeKutta2", "Superbee", "FREEFLOW"),
    ([70, 40], "RungeKutta4", "MonotizedCentral", "PERIODIC"),
    ([40, 36], "ForwardEuler", "DonorCell", "MIRROR"),
    ([24, 26, 20], "RungeKutta3TVD", "MinMod", "PERIODIC"),
]


@pytest.mark.parametrize("size,integ,lim,bm", MAXWELL_CASES)
def test_maxwell_fp64_parity(gpu_lib, size, integ, lim, bm):
    d = config.make_desc(size, solverName="MaxwellRoe", integratorName=integ, slopeLimiter=lim,
                         boundaryMethods=[{"min": bm, "max": bm}] * 3, cfl=0.4)
    rng = np.random.default_rng(5)
    q = ics.maxwell_1(d) + 0.1 * rng.standard_normal(6 * config.volume(d))
    g = _mk(d)
    g.set_state(q)
    o = oracle.OracleSim(d)
    o.set_state(q)
    for _ in range(5):
        assert g.update() == pytest.approx(o.update(), rel=1e-12)
    assert _relerr(g.get_state(), o.get_state()) < 1e-10
