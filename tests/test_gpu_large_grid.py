"""A large grid on the GPU: 32768 x 8192 cells (ghosts included, 268M cells;
Euler Roe, SRHD Roe, Maxwell Roe and EulerBurgers),
whose state planes lie beyond 4 GB of their buffer's base and whose CFL has
more row chunks than the strip march takes.

The state is uniform in y (v_y = 0) with periodic y boundaries.  Away from the
ghost rows every y-face has equal states on both sides, so the y-divergence is
exactly zero; near them it is not, as the ghosts are refreshed once per update
and not between the stages.  Two RK2 updates reach at most 2 x 2 x 2 rows from
the ghosts, so the large grid's first and last 24 rows go through the same
arithmetic as the rows of a 32768 x 48 grid with the same x-profile and the
same cell size (which the oracle runs in a second), and every row between them as that grid's middle row.
Checks:
  * the rows between the first and last 24 bit-identical;
  * the dt of every update equal to the oracle's (relative 1e-12; FP32-mixed
    1e-6, its CFL reads the FP32 state);
  * the first 24 rows, the last 24 rows and a middle row against the oracle's
    rows: FP64 relative 1e-11, FP32-mixed relative 2e-4 (the tolerances of the
    full-size parity tests).
"""
import functools

import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu

SX, SY = 32768, 8192
NY, K = 48, 24  # rows of the oracle's grid; rows of each edge region


def _desc(sy, precision, family="euler", sx=SX):
    # dy = 1/8192 on both grids: the y-extent grows with the row count
    h = sy / 16384.0
    per = {"min": "PERIODIC", "max": "PERIODIC"}
    if family == "srhd":
        return config.make_desc([sx, sy], xmin=[-0.5, -h], xmax=[0.5, h], solverName="SRHDRoe",
                                integratorName="RungeKutta2", slopeLimiter="Superbee",
                                boundaryMethods=[{"min": "FREEFLOW", "max": "FREEFLOW"}, per], cfl=0.5,
                                precision=precision, defs={"idealGas_heatCapacityRatio": 5.0 / 3.0})
    if family == "burgers":
        return config.make_desc([sx, sy], xmin=[-0.5, -h], xmax=[0.5, h], solverName="EulerBurgers",
                                integratorName="RungeKutta2", slopeLimiter="MonotizedCentral",
                                boundaryMethods=[per] * 2, cfl=0.5, precision=precision)
    if family == "maxwell":
        return config.make_desc([sx, sy], xmin=[-0.5, -h], xmax=[0.5, h], solverName="MaxwellRoe",
                                integratorName="RungeKutta2", slopeLimiter="Superbee",
                                boundaryMethods=[per] * 2, cfl=0.5, precision=precision)
    return config.make_desc([sx, sy], xmin=[-0.5, -h], xmax=[0.5, h], integratorName="RungeKutta2",
                            slopeLimiter="MonotizedCentral", boundaryMethods=[per] * 2, cfl=0.5,
                            precision=precision)


def _dt_tol(precision):
    return 1e-12 if precision == "fp64" else 1e-6


def _narrow_state(d, vmax=0.2):
    x = ics.grid(d)[0]
    if d.equation == 2:  # Maxwell: E and B fields varying in x
        eps = d.maxwell_permittivity
        comps = [0.1 * np.sin(2 * np.pi * x), np.cos(2 * np.pi * x) + (x > 0.1), np.full(x.shape, eps),
                 0.2 * np.cos(4 * np.pi * x), np.sin(2 * np.pi * x) - (x < -0.2), np.where(x <= 0, 1.0, -1.0)]
        return np.stack([c.reshape(-1) for c in comps], axis=1).reshape(-1).copy()
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x) + 0.5 * (x > 0.1)
    vx = vmax * np.cos(2 * np.pi * x)
    p = 1.0 + 0.2 * np.sin(4 * np.pi * x) + 0.4 * (x < -0.2)
    z = np.zeros(x.shape)
    return ics.build_state_euler(d, rho, vx, z, z, p)


# SRHD: velocities up to 0.8 (W up to 1.67), FREEFLOW in x; tolerances of the
# full-size SRHD parity test (FP64 1e-9; FP32-mixed 1e-3).  Maxwell: E and B
# varying in x (FP64 1e-10, as the Maxwell parity tests; FP32-mixed 5e-4).
# EulerBurgers: its pressure passes see zero y-gradients away from the edges
@pytest.mark.parametrize("family,precision,tol", [("euler", "fp64", 1e-11), ("euler", "fp32_mixed", 2e-4),
                                                  ("srhd", "fp64", 1e-9), ("srhd", "fp32_mixed", 1e-3),
                                                  ("maxwell", "fp64", 1e-10), ("maxwell", "fp32_mixed", 5e-4),
                                                  ("burgers", "fp64", 1e-11), ("burgers", "fp32_mixed", 2e-4)])
def test_large_grid_rows_match_oracle(gpu_lib, family, precision, tol):
    from hydrodynamicsgpu_amd import capi

    vmax = 0.8 if family == "srhd" else 0.2
    dn = _desc(NY, "fp64", family)
    qn = _narrow_state(dn, vmax)
    n = qn.size // (SX * NY)  # states per cell
    row = qn[: n * SX]
    assert np.array_equal(qn.reshape(NY, n * SX), np.broadcast_to(row, (NY, n * SX)))  # y-uniform
    o = oracle.OracleSim(dn)
    o.set_state(qn)

    d = _desc(SY, precision, family)
    print(f"large grid {SX}x{SY} {family} {precision}", flush=True)
    g = capi.Sim(d)
    big = np.empty(SY * row.size)
    big.reshape(SY, -1)[:] = row
    g.set_state(big)
    del big
    for _ in range(2):
        assert g.update() == pytest.approx(o.update(), rel=_dt_tol(precision))
    assert g.health() == 0
    q = g.get_state().reshape(SY, n * SX)
    del g
    assert (q[K:SY - K] == q[K]).all(), "middle rows of a y-uniform state differ"
    ref = o.get_state().reshape(NY, n * SX)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(q[:K] - ref[:K])) / scale < tol
    assert np.max(np.abs(q[SY - K:] - ref[K:])) / scale < tol
    assert np.max(np.abs(q[SY // 2] - ref[K])) / scale < tol


# 3D: 4096 x 256 x 256 (268M cells), 3D HLLC with TVD-RK3 as in the 3D blast
# workload, a state uniform in y and z (periodic there, mirror in x); the
# oracle runs 4096 x 48 x 48 with the same cell size.  Cell (x, y, z) of the
# large grid goes through the arithmetic of cell (x, m(y), m(z)) of the
# oracle's grid, m(i) = i in the first K rows, i - S + 2K in the last K, K
# between them (first-order fluxes over 2 updates x 3 stages reach 6 rows).
SX3, S3 = 4096, 256


def _desc3(s, precision, sx=SX3):
    h = s / 512.0  # dy = dz = 1/256
    return config.make_desc([sx, s, s], xmin=[-0.5, -h, -h], xmax=[0.5, h, h], solverName="EulerHLLC",
                            integratorName="RungeKutta3TVD", slopeLimiter="MinMod",
                            boundaryMethods=[{"min": "MIRROR", "max": "MIRROR"},
                                             {"min": "PERIODIC", "max": "PERIODIC"},
                                             {"min": "PERIODIC", "max": "PERIODIC"}],
                            cfl=0.5, precision=precision)


def _edge_map(s):
    i = np.arange(s)
    return np.where(i < K, i, np.where(i >= s - K, i - s + 2 * K, K))


@functools.lru_cache(maxsize=1)
def _oracle3():
    # the oracle's grid, its two dts and its state (shared by both precisions)
    dn = _desc3(NY, "fp64")
    qn = _narrow_state(dn)
    o = oracle.OracleSim(dn)
    o.set_state(qn)
    dts = [o.update() for _ in range(2)]
    return qn, dts, o.get_state()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32_mixed", 2e-4)])
def test_large_grid_3d_hllc_matches_oracle(gpu_lib, precision, tol):
    from hydrodynamicsgpu_amd import capi

    qn, dts, qo = _oracle3()
    n = qn.size // (SX3 * NY * NY)
    row = qn[: n * SX3]

    print(f"large grid {SX3}x{S3}x{S3} HLLC {precision}", flush=True)
    g = capi.Sim(_desc3(S3, precision))
    big = np.empty(S3 * S3 * row.size)
    big.reshape(S3 * S3, -1)[:] = row
    g.set_state(big)
    del big
    for dt in dts:
        assert g.update() == pytest.approx(dt, rel=_dt_tol(precision))
    assert g.health() == 0
    q = g.get_state().reshape(S3, S3, n * SX3)
    del g
    assert (q[K:S3 - K, K:S3 - K] == q[K, K]).all(), "middle rows of a y/z-uniform state differ"
    ref = qo.reshape(NY, NY, n * SX3)
    m = _edge_map(S3)
    scale = np.max(np.abs(ref))
    for z in range(S3):  # row blocks of one plane at a time
        assert np.max(np.abs(q[z] - ref[m[z]][m])) / scale < tol, f"plane {z}"
