"""Initial-condition generators (CPU): the counter hash, the reference's
cell-centre and buildStateEuler conventions, and the uneven batch mix."""
import numpy as np

from hydrodynamicsgpu_amd import config, ics
from hydrodynamicsgpu_amd.riemann import star_state


def test_splitmix64_published_sequence():
    # SplitMix64 seeded with 0 emits mix(k * golden) for k = 1, 2, 3, ...
    golden = 0x9E3779B97F4A7C15
    x = np.array([0, golden, (2 * golden) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    got = [int(v) for v in ics.splitmix64(x)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform01_range_and_streams():
    u = ics.uniform01(7, np.arange(100000, dtype=np.uint64))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    v = ics.uniform01(7, np.arange(100000, dtype=np.uint64), stream=1)
    assert not np.array_equal(u, v)


def test_cell_centres_include_ghosts():
    # Solver::resetState: pos = (xmax - xmin) * (i + .5) / size + xmin over all cells
    d = config.make_desc([8], xmin=[-1.0], xmax=[1.0])
    cx, _, _ = ics.cell_centers(d)
    np.testing.assert_allclose(cx, -1.0 + 0.25 * (np.arange(8) + 0.5))


def test_build_state_euler_closed_form():
    d = config.make_desc([4, 4])
    shape = (1, 4, 4)
    rho = np.full(shape, 2.0)
    vx = np.full(shape, 0.5)
    vy = np.full(shape, -1.0)
    p = np.full(shape, 3.0)
    q = ics.build_state_euler(d, rho, vx, vy, None, p).reshape(-1, 4)
    g = d.gamma
    E = 2.0 * (0.5 * (0.25 + 1.0) + 3.0 / ((g - 1.0) * 2.0))
    np.testing.assert_allclose(q[0], [2.0, 1.0, -2.0, E])
    r2, v2, p2 = ics.primitives_euler(d, q.reshape(-1))
    np.testing.assert_allclose(p2, 3.0)


def test_sod_split_at_cell_511():
    d = config.make_desc([1024], xmin=[-1.0], xmax=[1.0])
    rho, _, p = ics.primitives_euler(d, ics.sod(d))
    assert (rho[:512] == 1.0).all() and (rho[512:] == 0.125).all()
    assert (p[:512] == 1.0).all() and np.allclose(p[512:], 0.1)


def test_riemann_batch_mix_is_deterministic_and_filtered():
    a = ics.riemann_batch(48, seed=11)
    b = ics.riemann_batch(48, seed=11)
    c = ics.riemann_batch(48, seed=12)
    assert all(np.array_equal(qa, qb) for (_, qa), (_, qb) in zip(a, b))
    assert any(not np.array_equal(qa, qc) for (_, qa), (_, qc) in zip(a, c) if qa.shape == qc.shape) or \
        [d.size for d, _ in a] != [d.size for d, _ in c]
    dims = set()
    for d, q in a:
        dims.add(d.dim)
        n = d.size[0]
        if d.dim == 1:
            assert 256 <= n <= 65536
        else:
            assert 64 <= n <= 512 and d.size[1] == n
        rho, v, p = ics.primitives_euler(d, q)
        # the two constant sides of the Riemann problem, split at the x-centre
        left = (rho[0], v[0, 0], p[0])
        right = (rho[-1], v[-1, 0], p[-1])
        for s in (left, right):
            assert 0.1 <= s[0] <= 10.0 and 0.1 <= s[2] * (1 + 1e-12) and s[2] <= 10.0 * (1 + 1e-12)
            assert abs(s[1]) < 1.0
        ps, _ = star_state(left, right, d.gamma)
        pmin = min(left[2], right[2])
        assert 0.05 * pmin * (1 - 1e-9) <= ps <= 3.0 * pmin * (1 + 1e-9)
    assert dims == {1, 2}


def test_srhd_quadrants_reach_lorentz_factor_nine():
    # rho in [.1, 1], p in [.01, 1], W in [3, 10] with the fastest quadrant
    # in [9, 10]; one shared direction within +-.05 rad
    rho, p, W, vx, vy = ics.srhd_quadrant_states()
    assert ((0.1 <= rho) & (rho <= 1.0)).all() and ((0.01 <= p) & (p <= 1.0)).all()
    assert W.max() >= 9.0 and ((3.0 <= W) & (W <= 10.0)).all()
    v = np.hypot(vx, vy)
    np.testing.assert_allclose(1.0 / np.sqrt(1.0 - v * v), W, rtol=1e-10)
    ang = np.arctan2(vy, vx)
    assert np.ptp(np.unwrap(ang)) <= 0.1 + 1e-12
    # relative Lorentz factor of every pair W1 W2 (1 - v1.v2) stays below 1.5
    for i in range(4):
        for j in range(i + 1, 4):
            wrel = W[i] * W[j] * (1.0 - vx[i] * vx[j] - vy[i] * vy[j])
            assert 1.0 <= wrel < 1.5


def test_srhd_quadrants_stay_physical_in_the_oracle():
    # 64^2, RK2, Superbee, CFL .5: every primitive stays finite and above the
    # floors for 30 steps, and the fastest flow keeps W > 9
    import oracle

    d = config.make_desc([64, 64], xmin=[0, 0], xmax=[1, 1], solverName="SRHDRoe", integratorName="RungeKutta2",
                         cfl=0.5, defs={"idealGas_heatCapacityRatio": 5.0 / 3.0})
    o = oracle.OracleSim(d)
    o.set_state(ics.srhd_quadrants(d))
    for _ in range(30):
        assert np.isfinite(o.update())
    P = o.get_primitives().reshape(64, 64, 4)
    assert np.isfinite(P).all() and P[..., 0].min() > 0.05 and P[..., 3].min() > 0.05
    v2 = P[..., 1] ** 2 + P[..., 2] ** 2
    assert 1.0 / np.sqrt(1.0 - v2.max()) > 9.0
