"""The C++ host layer (HydroGPU::HydroGPUApp / Solver / Integrator / Equation
over the C ABI), exercised by tests/cpp/host_api_test.cpp the way the
reference's HydroGPUApp drives its solvers, and compared with the oracle.

CPU: configuration errors throw HydroGPU::Exception before any device work
(unknown solver / integrator / limiter names, self-gravity requested).
GPU: Sod (1024 cells, TVD-RK2, Superbee, to t = 0.2) within 1e-10 relative of
the oracle, run(n) continuing it, and short runs of EulerRoe / EulerHLL /
EulerHLLC / SRHDRoe / MaxwellRoe within 1e-10 (SRHD 1e-9: the con2prim
Newton stops at a relative change of 1e-7, so rounding differences can move
an iteration count).
"""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "host_api_test")


def _binary():
    if not os.path.exists(BIN):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp")])
    return BIN


def test_host_layer_config_errors_cpu():
    r = subprocess.run([_binary(), "cpu"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "passed cpu checks" in r.stdout


@pytest.mark.gpu
def test_host_layer_solvers_match_oracle(gpu_lib, tmp_path):
    # cwd: Solver::save and the checkpoint check write their files there
    r = subprocess.run([_binary(), "gpu"], capture_output=True, text=True, timeout=300, cwd=str(tmp_path))
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "passed gpu checks" in r.stdout
