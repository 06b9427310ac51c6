"""Seeded random configurations against the oracle.

A fixed-seed draw of grid dimension and size (odd and strip-straddling
sizes, the smallest grids included), equation and solver, limiter,
integrator, boundary method per side and axis, and precision.  Each
configuration runs three FP64 updates against the oracle (max |gpu -
oracle| <= 1e-10 * max |oracle|, SRHD 1e-9), run(3) bit-identical to three
update() calls, and the FP32-mixed form (plain or compensated) within 1e-4
of the FP64 GPU state.  The draw is part of the test (same seed, same
cases).  Two more draws cover the adjacent solvers (EulerBurgers, a
potential, useGravity) and solid masks.
"""
import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu

NAN_UNSAFE = ("BeamWarming", "Fromm", "CHARM", "Ospre", "VanAlbada1", "VanAlbada2", "VanLeer", "BarthJespersen")
SIZES_1D = [8, 9, 13, 61, 64, 65, 124, 125, 183, 300]
SIZES_2D = [8, 9, 13, 24, 61, 64, 65, 70]
SIZES_3D = [8, 9, 12, 17, 21, 26]
SOLVERS = ["EulerRoe", "EulerHLL", "EulerHLLC", "SRHDRoe", "MaxwellRoe"]
BOUNDS = ["PERIODIC", "MIRROR", "FREEFLOW", "NONE"]


def _draw(n=160, seed=20261204):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dim = int(rng.integers(1, 4))
        pool = {1: SIZES_1D, 2: SIZES_2D, 3: SIZES_3D}[dim]
        size = [int(rng.choice(pool)) for _ in range(dim)]
        solver = str(rng.choice(SOLVERS))
        lim = str(rng.choice([l for l in config.LIMITERS if l not in NAN_UNSAFE]))
        integ = str(rng.choice(config.INTEGRATORS))
        bm = [(str(rng.choice(BOUNDS)), str(rng.choice(BOUNDS))) for _ in range(3)]
        kahan = bool(rng.integers(0, 2))
        out.append((i, size, solver, lim, integ, bm, kahan))
    return out


CASES = _draw()


def _state(d, solver, seed):
    rng = np.random.default_rng(seed)
    if solver == "MaxwellRoe":
        return ics.maxwell_1(d) + 0.1 * rng.standard_normal(6 * config.volume(d))
    shp = ics.grid(d)[0].shape
    z = np.zeros(shp)
    vmax = 0.3 if solver == "SRHDRoe" else 0.5
    v = [rng.uniform(-vmax, vmax, shp) if a < d.dim else z for a in range(3)]
    return ics.build_state_euler(d, rng.uniform(0.6, 1.6, shp), v[0], v[1], v[2], rng.uniform(0.6, 1.6, shp))


def _relerr(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("case", CASES, ids=["c%02d-%s-%s-%s" % (c[0], "x".join(map(str, c[1])), c[2], c[4])
                                              for c in CASES])
def test_random_configuration(gpu_lib, case):
    from hydrodynamicsgpu_amd import capi

    i, size, solver, lim, integ, bm, kahan = case
    cfl = 0.02 if integ == "RungeKutta2NonTVD" else 0.4
    kw = dict(xmin=[0.0] * 3, xmax=[1.0] * 3, solverName=solver, integratorName=integ, slopeLimiter=lim,
              boundaryMethods=[{"min": a, "max": b} for a, b in bm], cfl=cfl)
    if solver == "SRHDRoe":
        kw["defs"] = {"idealGas_heatCapacityRatio": 5.0 / 3.0}
    d = config.make_desc(size, **kw)
    q = _state(d, solver, i)
    g, r = capi.Sim(d), capi.Sim(d)
    g.set_state(q)
    r.set_state(q)
    o = oracle.OracleSim(d)
    o.set_state(q)
    for _ in range(3):
        assert g.update() == pytest.approx(o.update(), rel=1e-11)
    r.run(3)
    ref = o.get_state()
    tol = 1e-9 if solver == "SRHDRoe" else 1e-10
    assert _relerr(g.get_state(), ref) < tol
    np.testing.assert_array_equal(r.get_state(), g.get_state())
    assert r.time() == g.time()
    assert g.health() == 0
    flags = config.FLAG_REFERENCE | (config.FLAG_KAHAN if kahan else 0)
    m = capi.Sim(config.make_desc(size, precision="fp32_mixed", flags=flags, **kw))
    m.set_state(q)
    for _ in range(3):
        m.update()
    assert _relerr(m.get_state(), g.get_state()) < 1e-4
    assert m.health() == 0


# A second fixed-seed draw over the adjacent solvers: EulerBurgers, and the
# Euler solvers with a non-zero potential or useGravity (3D HLL takes no
# potential).  Same checks: three FP64 updates against the oracle, run(3)
# bit-identical to update() x 3, FP32-mixed within 1e-4 of FP64.
ADJ_SOLVERS = ["EulerBurgers", "EulerRoe", "EulerHLL", "EulerHLLC"]


def _draw_adjacent(n=64, seed=20261205):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dim = int(rng.integers(1, 4))
        pool = {1: SIZES_1D, 2: SIZES_2D, 3: SIZES_3D}[dim]
        size = [int(rng.choice(pool)) for _ in range(dim)]
        solver = str(rng.choice(ADJ_SOLVERS))
        gravity = str(rng.choice(["none", "potential", "relax"])) if not (solver == "EulerHLL" and dim == 3) else "none"
        if solver == "EulerBurgers" and gravity == "none":
            gravity = str(rng.choice(["none", "potential", "relax"]))
        lim = str(rng.choice([l for l in config.LIMITERS if l not in NAN_UNSAFE]))
        integ = str(rng.choice(config.INTEGRATORS))
        bm = [(str(rng.choice(BOUNDS[:3])), str(rng.choice(BOUNDS[:3]))) for _ in range(3)]
        kahan = bool(rng.integers(0, 2))
        sweeps = int(rng.integers(0, 6))
        out.append((i, size, solver, gravity, lim, integ, bm, kahan, sweeps))
    return out


CASES_ADJ = _draw_adjacent()


@pytest.mark.parametrize("case", CASES_ADJ, ids=["a%02d-%s-%s-%s-%s" % (c[0], "x".join(map(str, c[1])), c[2], c[3], c[5])
                                                  for c in CASES_ADJ])
def test_random_adjacent_configuration(gpu_lib, case):
    from hydrodynamicsgpu_amd import capi

    i, size, solver, gravity, lim, integ, bm, kahan, sweeps = case
    cfl = 0.02 if integ == "RungeKutta2NonTVD" else 0.3
    kw = dict(xmin=[0.0] * 3, xmax=[1.0] * 3, solverName=solver, integratorName=integ, slopeLimiter=lim,
              boundaryMethods=[{"min": a, "max": b} for a, b in bm], cfl=cfl,
              useGravity=gravity == "relax", gaussSeidelMaxIter=sweeps)
    d = config.make_desc(size, **kw)
    q = _state(d, solver, 1000 + i)
    if gravity == "relax":
        # resetState adds the guess -rho into the energy, which lowers the
        # pressure by (gamma - 1) rho: start with that much more, so the
        # state stays physical
        qa = q.reshape(-1, d.dim + 2)
        qa[:, -1] += 1.25 * qa[:, 0]
    x, y, z, _ = ics.grid(d)
    pot = (0.2 * x - 0.1 * y + 0.05 * z).ravel() if gravity == "potential" else None
    g, r = capi.Sim(d), capi.Sim(d)
    o = oracle.OracleSim(d)
    for s in (g, r, o):
        s.set_state(q)
        if gravity != "none":
            s.set_potential(pot)
    for _ in range(3):
        assert g.update() == pytest.approx(o.update(), rel=1e-11)
    r.run(3)
    assert _relerr(g.get_state(), o.get_state()) < 1e-10
    np.testing.assert_array_equal(r.get_state(), g.get_state())
    assert r.time() == g.time()
    assert g.health() == 0
    flags = config.FLAG_REFERENCE | (config.FLAG_KAHAN if kahan else 0)
    m = capi.Sim(config.make_desc(size, precision="fp32_mixed", flags=flags, **kw))
    m.set_state(q)
    if gravity != "none":
        m.set_potential(pot)
    for _ in range(3):
        m.update()
    assert _relerr(m.get_state(), g.get_state()) < 1e-4
    assert m.health() == 0


# A third draw over solid cells: random masks (ghost cells included) in
# EulerRoe and EulerBurgers, which read them, and in EulerHLL/HLLC, which
# ignore them, with or without a potential.  Same checks; FP32-mixed within
# 2e-4 of FP64 (walls keep sharp jumps in the state).
SOLID_SOLVERS = ["EulerRoe", "EulerBurgers", "EulerRoe", "EulerBurgers", "EulerHLLC", "EulerHLL"]


def _draw_solid(n=48, seed=20261206):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        dim = int(rng.integers(1, 4))
        pool = {1: SIZES_1D, 2: SIZES_2D, 3: SIZES_3D}[dim]
        size = [int(rng.choice(pool)) for _ in range(dim)]
        solver = str(rng.choice(SOLID_SOLVERS))
        pot = bool(rng.integers(0, 2)) and not (solver == "EulerHLL" and dim == 3)
        lim = str(rng.choice([l for l in config.LIMITERS if l not in NAN_UNSAFE]))
        integ = str(rng.choice(config.INTEGRATORS))
        bm = [(str(rng.choice(BOUNDS[:3])), str(rng.choice(BOUNDS[:3]))) for _ in range(3)]
        frac = float(rng.choice([0.05, 0.15, 0.3]))
        kahan = bool(rng.integers(0, 2))
        out.append((i, size, solver, pot, lim, integ, bm, frac, kahan))
    return out


CASES_SOLID = _draw_solid()


@pytest.mark.parametrize("case", CASES_SOLID, ids=["s%02d-%s-%s-%s-%s" % (c[0], "x".join(map(str, c[1])), c[2],
                                                                          "pot" if c[3] else "nopot", c[5])
                                                    for c in CASES_SOLID])
def test_random_solid_configuration(gpu_lib, case):
    from hydrodynamicsgpu_amd import capi

    i, size, solver, with_pot, lim, integ, bm, frac, kahan = case
    cfl = 0.02 if integ == "RungeKutta2NonTVD" else 0.3
    kw = dict(xmin=[0.0] * 3, xmax=[1.0] * 3, solverName=solver, integratorName=integ, slopeLimiter=lim,
              boundaryMethods=[{"min": a, "max": b} for a, b in bm], cfl=cfl)
    d = config.make_desc(size, **kw)
    q = _state(d, solver, 3000 + i)
    x, y, z, _ = ics.grid(d)
    pot = (0.2 * x - 0.1 * y + 0.05 * z).ravel() if with_pot else None
    mask = (np.random.default_rng(4000 + i).random(config.volume(d)) < frac).astype(np.uint8)
    g, r = capi.Sim(d), capi.Sim(d)
    o = oracle.OracleSim(d)
    for s in (g, r, o):
        s.set_state(q)
        if with_pot:
            s.set_potential(pot)
        s.set_solid(mask)
    for _ in range(3):
        assert g.update() == pytest.approx(o.update(), rel=1e-11)
    r.run(3)
    assert _relerr(g.get_state(), o.get_state()) < 1e-10
    np.testing.assert_array_equal(r.get_state(), g.get_state())
    assert r.time() == g.time()
    assert g.health() == 0
    flags = config.FLAG_REFERENCE | (config.FLAG_KAHAN if kahan else 0)
    m = capi.Sim(config.make_desc(size, precision="fp32_mixed", flags=flags, **kw))
    m.set_state(q)
    if with_pot:
        m.set_potential(pot)
    m.set_solid(mask)
    for _ in range(3):
        m.update()
    assert _relerr(m.get_state(), g.get_state()) < 2e-4
    assert m.health() == 0
