"""The allocation-free CFL entry point (hydro_calc_timestep_noalloc) against
hydro_sim_calc_timestep, through the C ABI.

Each case builds a simulation, applies its boundary and takes the CFL
timestep it computes; the same stored state (and SRHD primitives) is then
handed to the allocation-free entry point in caller-owned device memory with
a padded plane stride.  The two must agree bit for bit: the same kernels run
on the same values.  Every case calls the entry point twice in a row, the
second time on a smaller grid, since both calls share the scratch array.
"""
import ctypes

import numpy as np
import pytest

from hydrodynamicsgpu_amd import capi, config, ics

pytestmark = pytest.mark.gpu


def _entry():
    f = capi.lib().hydro_calc_timestep_noalloc
    vp = ctypes.c_void_p
    f.argtypes = [vp, vp, ctypes.c_size_t, vp, ctypes.c_double, vp, vp, vp]
    f.restype = ctypes.c_int
    return f


class _Hip:
    """Device memory for the caller's side of the test, straight from the
    HIP runtime the library uses."""

    H2D, D2H = 1, 2

    def __init__(self):
        import os

        path = "/opt/rocm/lib/libamdhip64.so"
        L = ctypes.CDLL(path if os.path.exists(path) else "libamdhip64.so")
        L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        L.hipFree.argtypes = [ctypes.c_void_p]
        L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L = L
        self.live = []

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), arr.nbytes) == 0
        self.live.append(p)
        assert self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, self.H2D) == 0
        return p

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, self.D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def _planes(hip, aos, N, vol, plane, dtype):
    a = np.asarray(aos).reshape(vol, N).T
    out = np.zeros((N, plane), dtype=dtype)
    out[:, :vol] = a
    return hip.upload(out)


def _ic(desc):
    if desc.equation == config.EQUATIONS["SRHD"]:
        return ics.srhd_quadrants(desc)
    if desc.equation == config.EQUATIONS["Maxwell"]:
        rng = np.random.default_rng(5)
        return ics.maxwell_1(desc) + 0.1 * rng.standard_normal(6 * config.volume(desc))
    if desc.dim == 2:
        return ics.shear_kh(desc)
    if desc.dim == 3:
        return ics.blast(desc)
    return ics.sod(desc)


def _one(desc, t=0.0, poison=False):
    """(dt, flags) of the simulation and of the allocation-free entry."""
    N = config.num_states(desc)
    vol = config.volume(desc)
    q = _ic(desc)
    if poison:
        # an interior cell (ghost cells are rewritten by the boundary)
        q = q.copy()
        sx, sy = desc.size[0], desc.size[1]
        q[N * ((sy // 2) * sx + sx // 2)] = np.nan
    sim = capi.Sim(desc)
    sim.set_state(q)
    sim.boundary()
    dt_ref = sim.calc_timestep()
    health_ref = sim.health()
    hi, _ = sim.get_storage()
    mixed = desc.precision == config.PRECISION["fp32_mixed"]
    plane = vol + 37
    hip = _Hip()
    try:
        state = _planes(hip, hi, N, vol, plane, np.float32 if mixed else np.float64)
        prims = None
        if desc.equation == config.EQUATIONS["SRHD"]:
            prims = _planes(hip, sim.get_primitives(), N, vol, plane, np.float64)
        sim.close()
        dt_out = hip.upload(np.zeros(1))
        flags = hip.upload(np.zeros(1, dtype=np.uint32))
        rc = _entry()(ctypes.byref(desc), state, plane, prims, float(t), dt_out, flags, None)
        if rc != 0:
            raise capi.HydroError(capi.lib().hydro_last_error().decode())
        # the null stream: the copies back wait for the entry's kernels
        dt = float(hip.download(dt_out, np.zeros(1))[0])
        fl = int(hip.download(flags, np.zeros(1, dtype=np.uint32))[0])
    finally:
        hip.free()
    return (dt_ref, health_ref & 1), (dt, fl & 1)


def _kw(name):
    periodic = [{"min": "PERIODIC", "max": "PERIODIC"}] * 3
    mirror = [{"min": "MIRROR", "max": "MIRROR"}] * 3
    return {
        "roe2d_f64": dict(integratorName="RungeKutta4", slopeLimiter="MonotizedCentral", boundaryMethods=periodic),
        "roe2d_f32k": dict(integratorName="RungeKutta4", slopeLimiter="MonotizedCentral", boundaryMethods=periodic,
                           precision="fp32_mixed", flags=config.FLAG_REFERENCE | config.FLAG_KAHAN),
        "hllc1d_f64": dict(solverName="EulerHLLC", integratorName="RungeKutta2", xmin=[-1], xmax=[1]),
        "hllc3d_f32": dict(solverName="EulerHLLC", integratorName="RungeKutta3TVD", boundaryMethods=mirror,
                           precision="fp32_mixed", cfl=0.4),
        "roe3d_f64": dict(integratorName="RungeKutta2", boundaryMethods=mirror),
        "srhd2d_f64": dict(solverName="SRHDRoe", integratorName="RungeKutta2", xmin=[0, 0], xmax=[1, 1],
                           defs={"idealGas_heatCapacityRatio": 5.0 / 3.0}),
        "srhd2d_f32": dict(solverName="SRHDRoe", integratorName="RungeKutta2", xmin=[0, 0], xmax=[1, 1],
                           defs={"idealGas_heatCapacityRatio": 5.0 / 3.0}, precision="fp32_mixed"),
        "maxwell2d_f64": dict(solverName="MaxwellRoe", integratorName="RungeKutta2", boundaryMethods=periodic),
    }[name]


CASES = [
    ("roe2d_f64", [130, 70], [64, 40]),
    ("roe2d_f32k", [2048, 2048], [128, 128]),
    ("hllc1d_f64", [1024], [100]),
    ("hllc3d_f32", [64, 64, 64], [24, 20, 22]),
    ("roe3d_f64", [40, 36, 30], [16, 16, 16]),
    ("srhd2d_f64", [96, 96], [40, 36]),
    ("srhd2d_f32", [1024, 1024], [64, 64]),
    ("maxwell2d_f64", [70, 40], [30, 30]),
]


@pytest.mark.parametrize("name,big,small", CASES, ids=[c[0] for c in CASES])
def test_noalloc_timestep_matches_sim(gpu_lib, name, big, small):
    for size in (big, small):
        d = config.make_desc(size, **_kw(name))
        ref, got = _one(d)
        assert got[0] == ref[0], (name, size, got, ref)
        assert got[1] == ref[1] == 0


def test_noalloc_fixed_dt_clip_and_nan_flag(gpu_lib):
    # fixed dt, clipped to t_end - t
    d = config.make_desc([66, 50], useFixedDT=True, fixedDT=0.01, t_end=0.105,
                         flags=config.FLAG_REFERENCE | config.FLAG_CLIP_TO_TEND)
    ref, got = _one(d)
    assert got == ref and ref[0] == 0.01
    _, got = _one(d, t=0.1)
    assert got[0] == pytest.approx(0.005, rel=1e-12)
    # a NaN cell: skipped by the minimum, health bit 0 raised in both
    d = config.make_desc([130, 70], integratorName="RungeKutta4")
    ref, got = _one(d, poison=True)
    assert got == ref and ref[1] == 1
    # a second, smaller call after the poisoned one starts from clean scratch
    d = config.make_desc([40, 40], integratorName="RungeKutta4")
    ref, got = _one(d)
    assert got == ref and ref[1] == 0


def test_noalloc_rejects_bad_arguments(gpu_lib):
    d = config.make_desc([64, 40])
    f = _entry()
    dummy = ctypes.c_void_p(1)
    # plane stride below the cell count: rejected before any GPU work
    assert f(ctypes.byref(d), dummy, 64 * 40 - 1, None, 0.0, dummy, None, None) == -1
    assert b"plane stride" in capi.lib().hydro_last_error()
    # SRHD without primitives
    ds = config.make_desc([64, 40], solverName="SRHDRoe")
    assert f(ctypes.byref(ds), dummy, 64 * 40, None, 0.0, dummy, None, None) == -1
    # null output
    assert f(ctypes.byref(d), dummy, 64 * 40, None, 0.0, None, None, None) == -1
