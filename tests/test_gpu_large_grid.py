"""A large grid on the GPU: 32768 x 8192 cells (ghosts included, 268M cells),
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
  * the dt of every update equal to the oracle's (relative 1e-12);
  * the first 24 rows, the last 24 rows and a middle row against the oracle's
    rows: FP64 relative 1e-11, FP32-mixed relative 2e-4 (the tolerances of the
    full-size parity tests).
"""
import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu

SX, SY = 32768, 8192
NY, K = 48, 24  # rows of the oracle's grid; rows of each edge region


def _desc(sy, precision):
    # dy = 1/8192 on both grids: the y-extent grows with the row count
    h = sy / 16384.0
    return config.make_desc([SX, sy], xmin=[-0.5, -h], xmax=[0.5, h], integratorName="RungeKutta2",
                            slopeLimiter="MonotizedCentral",
                            boundaryMethods=[{"min": "PERIODIC", "max": "PERIODIC"}] * 2, cfl=0.5,
                            precision=precision)


def _narrow_state(d):
    x = ics.grid(d)[0]
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * x) + 0.5 * (x > 0.1)
    vx = 0.2 * np.cos(2 * np.pi * x)
    p = 1.0 + 0.2 * np.sin(4 * np.pi * x) + 0.4 * (x < -0.2)
    z = np.zeros(x.shape)
    return ics.build_state_euler(d, rho, vx, z, z, p)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("fp32_mixed", 2e-4)])
def test_large_grid_rows_match_oracle(gpu_lib, precision, tol):
    from hydrodynamicsgpu_amd import capi

    dn = _desc(NY, precision)
    qn = _narrow_state(dn)
    n = qn.size // (SX * NY)  # states per cell
    row = qn[: n * SX]
    assert np.array_equal(qn.reshape(NY, n * SX), np.broadcast_to(row, (NY, n * SX)))  # y-uniform
    o = oracle.OracleSim(_desc(NY, "fp64"))
    o.set_state(qn)

    d = _desc(SY, precision)
    print(f"large grid {SX}x{SY} {precision}", flush=True)
    g = capi.Sim(d)
    big = np.empty(SY * row.size)
    big.reshape(SY, -1)[:] = row
    g.set_state(big)
    del big
    for _ in range(2):
        assert g.update() == pytest.approx(o.update(), rel=1e-12)
    assert g.health() == 0
    q = g.get_state().reshape(SY, n * SX)
    del g
    assert (q[K:SY - K] == q[K]).all(), "middle rows of a y-uniform state differ"
    ref = o.get_state().reshape(NY, n * SX)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(q[:K] - ref[:K])) / scale < tol
    assert np.max(np.abs(q[SY - K:] - ref[K:])) / scale < tol
    assert np.max(np.abs(q[SY // 2] - ref[K])) / scale < tol
