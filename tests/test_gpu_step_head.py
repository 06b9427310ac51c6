"""The update head (boundary and CFL strip march in one launch,
step_head_kernel, then the dt tail) against the same parts as separate
launches (HYDRO_STEP_HEAD=0), through the C ABI.

The fused form reads ghost cells through the boundary's composed map while
other workgroups of the same launch write them.  Every update's dt, the
state, the clock, the health bits and (SRHD) the primitives must agree bit
for bit with the separate launches, for 1D, 2D and 3D grids, every
strip-march solver, both precisions (and compensated storage), per-side
boundary mixes with NONE, single updates and captured runs; and a NaN in the
state must raise the same health bit.
"""
import numpy as np
import pytest

from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu


def _pair(monkeypatch, d):
    from hydrodynamicsgpu_amd import capi

    monkeypatch.setenv("HYDRO_STEP_HEAD", "0")
    split = capi.Sim(d)
    monkeypatch.setenv("HYDRO_STEP_HEAD", "1")
    fused = capi.Sim(d)
    return split, fused


def _state(d, seed):
    rng = np.random.default_rng(seed)
    x, _, _, _ = ics.grid(d)
    shp = x.shape
    return ics.build_state_euler(d, rng.uniform(0.5, 2.0, shp), rng.uniform(-0.8, 0.8, shp),
                                 rng.uniform(-0.8, 0.8, shp), rng.uniform(-0.8, 0.8, shp), rng.uniform(0.5, 2.0, shp))


def _check(split, fused, srhd=False):
    np.testing.assert_array_equal(fused.get_state(), split.get_state())
    if srhd:
        np.testing.assert_array_equal(fused.get_primitives(), split.get_primitives())
    assert fused.time() == split.time()
    assert fused.health() == split.health()


def _drive(split, fused, srhd=False):
    for _ in range(3):
        assert fused.update() == split.update()
    _check(split, fused, srhd)
    for n in (2, 5, 17):
        split.run(n)
        fused.run(n)
        _check(split, fused, srhd)
        assert fused.update() == split.update()


BMIX = [
    [("PERIODIC", "PERIODIC")] * 3,
    [("MIRROR", "MIRROR")] * 3,
    [("FREEFLOW", "FREEFLOW")] * 3,
    [("MIRROR", "NONE"), ("NONE", "FREEFLOW"), ("PERIODIC", "PERIODIC")],
    [("FREEFLOW", "MIRROR"), ("MIRROR", "PERIODIC"), ("NONE", "MIRROR")],
]


@pytest.mark.parametrize("solver", ["Roe", "HLL", "HLLC"])
@pytest.mark.parametrize("size", [[300], [125, 64], [61, 130], [24, 20, 18], [70, 33, 21]])
@pytest.mark.parametrize("bm", range(len(BMIX)))
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed", "fp32_kahan"])
def test_step_head_matches_separate_launches(gpu_lib, monkeypatch, solver, size, bm, precision):
    flags = config.FLAG_REFERENCE | (config.FLAG_KAHAN if precision == "fp32_kahan" else 0)
    d = config.make_desc(size, solverName="Euler" + solver, integratorName="RungeKutta2TVD",
                         slopeLimiter="MonotizedCentral", cfl=0.4,
                         boundaryMethods=[{"min": a, "max": b} for a, b in BMIX[bm][:len(size)]],
                         precision="fp64" if precision == "fp64" else "fp32_mixed", flags=flags)
    split, fused = _pair(monkeypatch, d)
    q = _state(d, 5 + bm)
    split.set_state(q)
    fused.set_state(q)
    _drive(split, fused)


@pytest.mark.parametrize("size", [[260], [70, 52], [125, 64]])
@pytest.mark.parametrize("bm", [0, 2, 3])
@pytest.mark.parametrize("precision", ["fp64", "fp32_mixed"])
def test_step_head_matches_separate_launches_srhd(gpu_lib, monkeypatch, size, bm, precision):
    from srhd_states import quadrants_for

    d = config.make_desc(size, solverName="SRHDRoe", integratorName="RungeKutta2", xmin=[0] * len(size),
                         xmax=[1] * len(size), cfl=0.4, defs={"idealGas_heatCapacityRatio": 5.0 / 3.0},
                         boundaryMethods=[{"min": a, "max": b} for a, b in BMIX[bm][:len(size)]],
                         precision=precision)
    q = quadrants_for(d) if len(size) == 2 else ics.riemann(d, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1))
    split, fused = _pair(monkeypatch, d)
    split.set_state(q)
    fused.set_state(q)
    _drive(split, fused, srhd=True)


@pytest.mark.parametrize("size", [[125, 64], [24, 20, 18]])
def test_step_head_nan_raises_the_same_health_bit(gpu_lib, monkeypatch, size):
    d = config.make_desc(size, solverName="EulerRoe", integratorName="RungeKutta2", cfl=0.4)
    q = _state(d, 3)
    n = d.size[0]
    cell = 5 + n * (4 if len(size) > 1 else 0) + (n * d.size[1] * 3 if len(size) > 2 else 0)
    q[cell * config.num_states(d)] = np.nan
    split, fused = _pair(monkeypatch, d)
    split.set_state(q)
    fused.set_state(q)
    da, db = split.update(), fused.update()
    assert da == db or (np.isnan(da) and np.isnan(db))
    assert fused.health() == split.health() and fused.health() & 1
    np.testing.assert_array_equal(np.isnan(fused.get_state()), np.isnan(split.get_state()))
