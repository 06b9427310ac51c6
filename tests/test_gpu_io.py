"""Checkpoint resume and FITS output of a running grid, through the C ABI
(hydro_sim_restore).  A resumed run repeats the uninterrupted one bit for
bit: the state, its clock and, for SRHD, the primitives it continues from."""
import numpy as np
import pytest

from hydrodynamicsgpu_amd import config, ics, io
from srhd_states import quadrants_for

pytestmark = pytest.mark.gpu

CASES = [
    ("fp64", 0, "EulerRoe", [64, 48], "RungeKutta4"),
    ("fp32_mixed", 0, "EulerRoe", [64, 48], "RungeKutta4"),
    ("fp32_mixed", config.FLAG_KAHAN, "EulerHLLC", [40, 24, 20], "RungeKutta3TVD"),
    ("fp64", 0, "SRHDRoe", [48, 40], "RungeKutta2"),
]


@pytest.mark.parametrize("precision,flags,solver,size,integ", CASES)
def test_checkpoint_resume_is_exact(gpu_lib, tmp_path, precision, flags, solver, size, integ):
    from hydrodynamicsgpu_amd import capi

    kw = dict(solverName=solver, integratorName=integ, precision=precision, flags=config.FLAG_REFERENCE | flags,
              boundaryMethods=[{"min": "FREEFLOW", "max": "FREEFLOW"}] * 3, cfl=0.4)
    if solver == "SRHDRoe":
        d = config.make_desc(size, xmin=[0, 0], xmax=[1, 1], defs={"idealGas_heatCapacityRatio": 5 / 3}, **kw)
        q = quadrants_for(d)
    elif len(size) == 3:
        d = config.make_desc(size, xmin=[-1] * 3, xmax=[1] * 3, **kw)
        q = ics.blast(d)
    else:
        d = config.make_desc(size, **kw)
        q = ics.riemann(d, (1.0, 0.2, 1.0), (0.4, -0.1, 0.5))
    s = capi.Sim(d)
    s.set_state(q)
    s.run(6)
    path = str(tmp_path / "run.ckpt")
    io.save_checkpoint(s, path)
    s.run(9)
    a, ta = s.get_state(), s.time()
    s.close()
    r = io.load_checkpoint(path)
    assert r.time()[1] == 6
    r.run(9)
    np.testing.assert_array_equal(r.get_state(), a)
    assert r.time()[:2] == ta[:2]
    # FITS of the running grid: the state channels as 32-bit floats
    paths = io.save_fits(d, r.get_state(), str(tmp_path))
    N = config.num_states(d)
    for c, p in enumerate(paths):
        _, img = io.read_fits(p)
        np.testing.assert_array_equal(img.reshape(-1), a.reshape(-1, N)[:, c].astype(np.float32).astype(np.float64))
    r.close()


def test_restore_rejects_mismatched_primitives(gpu_lib):
    from hydrodynamicsgpu_amd import capi

    d = config.make_desc([32, 32], solverName="SRHDRoe")
    s = capi.Sim(d)
    q = np.ones(4 * 32 * 32)
    with pytest.raises(capi.HydroError, match="primitives"):
        s.restore(q, 0.0, 0)
    e = capi.Sim(config.make_desc([32, 32]))
    with pytest.raises(capi.HydroError, match="SRHD only"):
        e.restore(q, 0.0, 0, prims=q)
    with pytest.raises(capi.HydroError, match=">= 0"):
        e.restore(q, -1.0, 0)
