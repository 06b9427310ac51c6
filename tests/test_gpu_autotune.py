"""The single-grid autotuner (hydro_sim_autotune) through the C ABI.

Every decomposition of the stage kernels gives the same bits, so a tuned
simulation must continue exactly as an untuned one from the same state; the
tuner restores the state it trials on; its pick is one of its candidates; it
is kept per grid shape (this process, and the HYDRO_TUNE_FILE file read by a
new process)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from hydrodynamicsgpu_amd import config, ics

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _mk(d):
    from hydrodynamicsgpu_amd import capi

    return capi.Sim(d)


CASES = {
    "roe2d_f32": lambda: config.make_desc([260, 200], integratorName="RungeKutta4", slopeLimiter="MonotizedCentral",
                                          boundaryMethods=[{"min": "PERIODIC", "max": "PERIODIC"}] * 2,
                                          precision="fp32_mixed"),
    "roe2d_f64": lambda: config.make_desc([130, 98], integratorName="RungeKutta2", slopeLimiter="Superbee"),
    "hllc3d_f32k": lambda: config.make_desc([70, 45, 38], solverName="EulerHLLC", integratorName="RungeKutta3TVD",
                                            boundaryMethods=[{"min": "MIRROR", "max": "MIRROR"}] * 3,
                                            precision="fp32_mixed",
                                            flags=config.FLAG_REFERENCE | config.FLAG_KAHAN),
    "roe3d_f64": lambda: config.make_desc([36, 30, 28], integratorName="RungeKutta2", slopeLimiter="MinMod"),
}


def _ic(d):
    if d.dim == 2:
        return ics.shear_kh(d)
    return ics.blast(d)


@pytest.mark.parametrize("name", sorted(CASES))
def test_autotune_restores_state_and_keeps_the_bits(gpu_lib, name):
    d = CASES[name]()
    q = _ic(d)
    ref = _mk(d)
    ref.set_state(q)
    ref.run(3)
    tuned = _mk(d)
    tuned.set_state(q)
    tuned.run(3)
    before = tuned.get_state()
    t_before = tuned.time()
    params, ms = tuned.autotune(updates=2)
    assert ms > 0
    assert tuned.decomposition() == params
    np.testing.assert_array_equal(tuned.get_state(), before)
    assert tuned.time() == t_before
    for s in (ref, tuned):
        s.run(5)
        s.update()
    np.testing.assert_array_equal(tuned.get_state(), ref.get_state())
    assert tuned.time() == ref.time() and tuned.health() == 0
    # a new simulation of the same shape starts from the winner
    again = _mk(d)
    assert again.decomposition() == params


def test_tune_file_is_read_by_a_new_process(gpu_lib, tmp_path):
    d = config.make_desc([190, 150], integratorName="RungeKutta2", slopeLimiter="Superbee")
    key = "d2_190x150x1_e0_s0_l%d_i%d_p0_k0" % (d.limiter, d.integrator)
    path = tmp_path / "tune.txt"
    path.write_text("%s 13 2 0 4 0.5\n" % key)
    code = ("from hydrodynamicsgpu_amd import capi, config\n"
            "d = config.make_desc([190, 150], integratorName='RungeKutta2', slopeLimiter='Superbee')\n"
            "print(capi.Sim(d).decomposition())\n")
    env = dict(os.environ, HYDRO_TUNE_FILE=str(path))
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().endswith("(13, 2, 0, 0)")
