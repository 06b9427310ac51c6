"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code.

The oracle (oracle/hydro_oracle.c) and the C++ host layer
(hydrodynamicsgpu_amd/host/src/*.cpp, with the test program
tests/cpp/host_api_test.cpp) are rebuilt with gcc/g++
-fsanitize=address,undefined into a temporary directory, and then:
  * the oracle pin suites run against the sanitized oracle (loaded through
    ORACLE_LIB into a Python that preloads the ASan runtime);
  * the host layer runs its CPU modes: configuration errors, the FITS
    writer, and a checkpoint written and read back.
Any invalid access or undefined operation aborts the run (halt on error),
which fails the test.  Host code only: the HIP library is linked as built.
"""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "hydrodynamicsgpu_amd")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


def _runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def san_build(tmp_path_factory):
    asan = _runtime("libasan.so")
    if not asan or not _runtime("libubsan.so"):
        pytest.skip("gcc sanitizer runtimes not installed")
    out = tmp_path_factory.mktemp("san")
    lib = str(out / "liboracle_san.so")
    subprocess.check_call(["gcc"] + SAN + ["-fPIC", "-shared", "-std=c11", "-fopenmp", "-ffp-contract=off", "-o", lib,
                                           os.path.join(ROOT, "oracle", "hydro_oracle.c"), "-lm"])
    exe = None
    if os.path.exists(os.path.join(PKG, "libhydro_hip.so")):
        exe = str(out / "host_api_test_san")
        srcs = sorted(glob.glob(os.path.join(PKG, "host", "src", "*.cpp")))
        subprocess.check_call(["g++"] + SAN + ["-std=c++17", "-I", os.path.join(PKG, "host", "include"),
                                               "-I", os.path.join(ROOT, "include"), "-o", exe] + srcs +
                              [os.path.join(ROOT, "tests", "cpp", "host_api_test.cpp"), "-L", PKG, "-lhydro_hip",
                               "-L", str(out), "-loracle_san", "-Wl,-rpath," + PKG, "-Wl,-rpath," + str(out)])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    return {"asan": asan, "lib": lib, "exe": exe, "dir": str(out), "env": env}


def test_sanitized_oracle_passes_its_pins(san_build):
    env = dict(san_build["env"], LD_PRELOAD=san_build["asan"], ORACLE_LIB=san_build["lib"])
    probe = "import oracle; print(oracle._LIB_PATH)"
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == san_build["lib"], r.stderr
    tests = ["tests/test_oracle_pins.py", "tests/test_oracle_kernels.py", "tests/test_oracle_solver.py",
             "tests/test_oracle_maxwell_srhd.py", "tests/test_ics.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider"] + tests, cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "passed" in r.stdout and "failed" not in r.stdout


def test_sanitized_host_layer_cpu_modes(san_build):
    if not san_build["exe"]:
        pytest.skip("libhydro_hip.so not built")
    exe, env, d = san_build["exe"], san_build["env"], san_build["dir"]

    def run(*args):
        r = subprocess.run([exe] + list(args), cwd=d, env=env, capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, (args, r.stdout[-2000:], r.stderr[-3000:])
        return r.stdout

    assert "passed cpu checks" in run("cpu")
    run("fits", os.path.join(d, "x.fits"), "5", "4", "3")
    assert os.path.getsize(os.path.join(d, "x.fits")) % 2880 == 0
    run("ckpt-write", os.path.join(d, "c.ckpt"))
    assert "steps 7" in run("ckpt-read", os.path.join(d, "c.ckpt"))
