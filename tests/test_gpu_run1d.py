"""The single-workgroup run of small 1D Euler grids (run1d.hip).

hydro_sim_run on a small 1D Euler grid (up to 24 strips of 60 cells in FP64,
48 in FP32) runs every update in one workgroup with the same device code as
the launch-per-kernel path, so the two must agree bit for bit: state (ghost
cells included), clock, step count and health, for every solver, precision,
integrator and boundary.  The
launch-per-kernel path is selected with HYDRO_RUN1D=0 at creation, or by
calling update() step by step.  Clipped runs past t_end are checked against
the oracle (FP64, rtol 1e-10).
"""
import os

import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu


def _sim(desc, fused):
    from hydrodynamicsgpu_amd import capi

    old = os.environ.get("HYDRO_RUN1D")
    os.environ["HYDRO_RUN1D"] = "1" if fused else "0"
    try:
        return capi.Sim(desc)
    finally:
        if old is None:
            del os.environ["HYDRO_RUN1D"]
        else:
            os.environ["HYDRO_RUN1D"] = old


def _smooth(desc, seed):
    rng = np.random.default_rng(seed)
    x = ics.cell_centers(desc)[0]
    rho = 1.0 + 0.2 * np.sin(2 * np.pi * x) + 0.01 * rng.uniform(-1, 1, x.shape)
    vx = 0.3 * np.cos(2 * np.pi * x)
    p = 1.0 + 0.1 * np.sin(4 * np.pi * x)
    z = np.zeros_like(x)
    return ics.build_state_euler(desc, rho, vx, z, z, p)


def _same(a, b):
    np.testing.assert_array_equal(a.get_state(), b.get_state())
    assert a.time() == b.time()
    assert a.health() == b.health()


CASES = [
    # size, solver, limiter, integrator, precision, boundary, flags
    (1024, "EulerRoe", "Superbee", "RungeKutta2TVD", "fp64", "FREEFLOW", 0),
    (1024, "EulerRoe", "MonotizedCentral", "RungeKutta4", "fp32_mixed", "PERIODIC", 0),
    (700, "EulerRoe", "Koren", "RungeKutta3", "fp32_mixed", "MIRROR", config.FLAG_KAHAN),
    (1444, "EulerRoe", "MinMod", "RungeKutta2Heun", "fp64", "PERIODIC", 0),
    (2884, "EulerRoe", "Superbee", "RungeKutta4NonTVD", "fp32_mixed", "PERIODIC", config.FLAG_KAHAN),
    (64, "EulerHLL", "Superbee", "RungeKutta3TVD", "fp64", "MIRROR", 0),
    (1000, "EulerHLLC", "Superbee", "ForwardEuler", "fp32_mixed", "FREEFLOW", 0),
    (513, "EulerRoe", "DonorCell", "ForwardEuler", "fp64", "PERIODIC", 0),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}-{c[2]}-{c[3]}-{c[4]}-{c[5]}" for c in CASES])
def test_fused_run_matches_launch_per_kernel(gpu_lib, case):
    n, solver, lim, integ, prec, bnd, flags = case
    d = config.make_desc([n], xmin=[0], xmax=[1], solverName=solver, slopeLimiter=lim, integratorName=integ,
                         boundaryMethods=[{"min": bnd, "max": bnd}], precision=prec,
                         flags=config.FLAG_REFERENCE | flags)
    q = _smooth(d, seed=n)
    a, b = _sim(d, True), _sim(d, False)
    a.set_state(q)
    b.set_state(q)
    for k in (1, 2, 37, 300):  # odd counts flip the buffer parity; 300 spans two launches
        a.run(k)
        b.run(k)
        _same(a, b)
    assert a.health() == 0


def test_fused_run_matches_update_steps(gpu_lib):
    d = config.make_desc([1024], xmin=[-1], xmax=[1], integratorName="RungeKutta2TVD", slopeLimiter="Superbee",
                         cfl=0.5)
    q = ics.sod(d)
    a, b = _sim(d, True), _sim(d, True)
    a.set_state(q)
    b.set_state(q)
    a.run(25)
    for _ in range(25):
        b.update()
    _same(a, b)


def test_fused_run_fixed_dt(gpu_lib):
    d = config.make_desc([256], xmin=[0], xmax=[1], integratorName="RungeKutta4", slopeLimiter="MonotizedCentral",
                         boundaryMethods=[{"min": "PERIODIC", "max": "PERIODIC"}], useFixedDT=True, fixedDT=1e-3)
    q = _smooth(d, seed=3)
    a, b = _sim(d, True), _sim(d, False)
    a.set_state(q)
    b.set_state(q)
    a.run(41)
    b.run(41)
    _same(a, b)
    assert a.time()[0] == pytest.approx(41e-3, rel=1e-12)


@pytest.mark.parametrize("prec", ["fp64", "fp32_mixed"])
def test_clipped_run_stops_at_t_end(gpu_lib, prec):
    # runs past t_end are no-ops: the clock stops at t_end, the step count at
    # the last real update, the health stays clean; both paths agree
    n = 1024
    kw = dict(xmin=[-1], xmax=[1], integratorName="RungeKutta2TVD", slopeLimiter="Superbee", cfl=0.5, t_end=0.05)
    flags = config.FLAG_REFERENCE | config.FLAG_CLIP_TO_TEND
    d = config.make_desc([n], precision=prec, flags=flags, **kw)
    q = ics.sod(d)
    a, b = _sim(d, True), _sim(d, False)
    a.set_state(q)
    b.set_state(q)
    for k in (64, 64, 63):
        a.run(k)
        b.run(k)
    np.testing.assert_array_equal(a.get_state(), b.get_state())
    ta, sa, _ = a.time()
    tb, sb, _ = b.time()
    assert ta == tb == 0.05 and sa == sb
    assert a.health() == 0 and b.health() == 0
    if prec == "fp64":
        o = oracle.OracleSim(config.make_desc([n], flags=flags, **kw))
        o.set_state(q)
        steps = 0
        while o.update() > 0:
            steps += 1
        assert o.time == 0.05 and steps == sa
        assert o.update() == 0.0 and o.time == 0.05
        ref = o.get_state()
        assert np.max(np.abs(a.get_state() - ref)) <= 1e-10 * np.max(np.abs(ref))


def test_fused_run_flags_nan(gpu_lib):
    d = config.make_desc([128])
    q = ics.sod(d)
    q[3 * 40] = -1.0  # negative density: NaN sound speed in the CFL
    a, b = _sim(d, True), _sim(d, False)
    a.set_state(q)
    b.set_state(q)
    a.run(3)
    b.run(3)
    assert a.health() & 1 and a.health() == b.health()
    assert a.time()[1] == b.time()[1]
