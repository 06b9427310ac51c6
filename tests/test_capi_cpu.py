"""CPU-side checks of the C ABI library: it loads, exports every entry point
include/hydro_capi.h declares, and the descriptor mirror matches the header."""
import ctypes
import os
import re

import pytest

from hydrodynamicsgpu_amd import capi, config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "hydro_capi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(hydro_[a-z_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(capi.LIB_PATH):
        pytest.fail("libhydro_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(capi.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) >= 25
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared) == set(capi.EXPORTED_SYMBOLS)


def test_desc_default_matches_reference_config():
    lib = capi.lib()
    d = config.HydroSimDesc()
    lib.hydro_sim_desc_default(ctypes.byref(d))
    assert d.limiter == config.LIMITERS.index("Superbee")
    assert d.integrator == config.INTEGRATORS.index("ForwardEuler")
    assert d.gamma == 1.4
    assert d.srhd_solve_prim_max_iter == 1000
    assert [d.boundary[a][m] for a in range(3) for m in range(2)] == [config.BOUNDARY["FREEFLOW"]] * 6
    assert lib.hydro_version().decode().startswith("hydro-mi355x")


def test_struct_size_matches_header():
    # one int dim, 3 ints size, 6 doubles xmin/xmax, 5 ints, 6 ints boundary,
    # doubles and ints with natural alignment: compare with a C compile
    import subprocess
    import tempfile

    src = '#include "hydro_capi.h"\n#include <stdio.h>\n#include <stddef.h>\nint main(){printf("%zu %zu %zu\\n", sizeof(hydro_sim_desc), offsetof(hydro_sim_desc, cfl), offsetof(hydro_sim_desc, flags));return 0;}\n'
    with tempfile.TemporaryDirectory() as t:
        c = os.path.join(t, "s.c")
        open(c, "w").write(src)
        exe = os.path.join(t, "s")
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = subprocess.check_output([exe]).decode().split()
    assert int(out[0]) == ctypes.sizeof(config.HydroSimDesc)
    assert int(out[1]) == config.HydroSimDesc.cfl.offset
    assert int(out[2]) == config.HydroSimDesc.flags.offset


def test_bad_descriptor_rejected_without_gpu():
    # validation happens before any device call
    lib = capi.lib()
    d = config.make_desc([16, 16])
    d.dim = 5
    h = ctypes.c_void_p()
    rc = lib.hydro_sim_create(ctypes.c_void_p(1), ctypes.byref(d), ctypes.byref(h))
    assert rc == -1
    assert b"dim" in lib.hydro_last_error()
