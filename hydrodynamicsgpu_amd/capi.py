"""ctypes binding of libhydro_hip.so (include/hydro_capi.h).

The library is built in-tree by hydrodynamicsgpu_amd/build.py.  Loading fails
loudly when it is missing: there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

from . import config

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhydro_hip.so")

# include/hydro_capi.h HYDRO_PHASE_*, in order
PHASES = ("boundary", "cfl", "head", "dt_tail", "stages", "con2prim", "pressure", "ghosts", "totals", "gravity")
_lib = None


class HydroError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HydroError("libhydro_hip.so is not built: run hydrodynamicsgpu_amd/build.py "
                             "(or __graft_entry__.build()) first")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip, sz = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_size_t
        L.hydro_last_error.restype = ctypes.c_char_p
        L.hydro_version.restype = ctypes.c_char_p
        L.hydro_sim_desc_default.argtypes = [vp]
        L.hydro_ctx_create.argtypes = [ip, ctypes.POINTER(vp)]
        L.hydro_ctx_destroy.argtypes = [vp]
        L.hydro_sim_create.argtypes = [vp, vp, ctypes.POINTER(vp)]
        L.hydro_sim_destroy.argtypes = [vp]
        L.hydro_sim_num_states.argtypes = [vp]
        L.hydro_sim_volume.argtypes = [vp]
        L.hydro_sim_volume.restype = sz
        for n in ("hydro_sim_set_state", "hydro_sim_get_state", "hydro_sim_get_primitives"):
            getattr(L, n).argtypes = [vp, dp, sz]
        fp = ctypes.POINTER(ctypes.c_float)
        L.hydro_sim_restore.argtypes = [vp, dp, fp, dp, sz, ctypes.c_double, ctypes.c_longlong]
        L.hydro_sim_get_storage.argtypes = [vp, dp, fp, sz]
        L.hydro_sim_boundary.argtypes = [vp]
        L.hydro_sim_calc_timestep.argtypes = [vp, dp]
        L.hydro_sim_step.argtypes = [vp, ctypes.c_double]
        L.hydro_sim_update.argtypes = [vp, dp]
        L.hydro_sim_run.argtypes = [vp, ip]
        L.hydro_sim_sync.argtypes = [vp]
        L.hydro_sim_time.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_longlong), dp]
        L.hydro_sim_totals.argtypes = [vp, dp]
        L.hydro_sim_health.argtypes = [vp, ctypes.POINTER(ctypes.c_uint)]
        L.hydro_sim_set_profiling.argtypes = [vp, ip]
        L.hydro_sim_stage_timing.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_longlong)]
        L.hydro_sim_phase_timing.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_longlong)]
        L.hydro_sim_set_potential.argtypes = [vp, dp, sz]
        L.hydro_sim_get_potential.argtypes = [vp, dp, sz]
        L.hydro_sim_set_solid.argtypes = [vp, ctypes.c_void_p, sz]
        L.hydro_sim_get_solid.argtypes = [vp, ctypes.c_void_p, sz]
        L.hydro_sim_stream.argtypes = [vp]
        L.hydro_sim_stream.restype = vp
        L.hydro_sim_element_size.argtypes = [vp]
        L.hydro_sim_element_size.restype = sz
        L.hydro_integrator_tableau.argtypes = [ip, ctypes.POINTER(ip), dp, dp]
        L.hydro_sim_buffer_create.argtypes = [vp, ctypes.POINTER(vp)]
        L.hydro_sim_buffer_destroy.argtypes = [vp, vp]
        L.hydro_sim_state_buffer.argtypes = [vp, ctypes.POINTER(vp)]
        L.hydro_sim_init_step.argtypes = [vp]
        L.hydro_sim_calc_deriv.argtypes = [vp, ctypes.c_double, vp]
        L.hydro_sim_buffer_copy.argtypes = [vp, vp, vp]
        L.hydro_sim_buffer_zero.argtypes = [vp, vp]
        L.hydro_sim_mult_add.argtypes = [vp, vp, vp, vp, ctypes.c_double]
        L.hydro_sim_finish_step.argtypes = [vp, ctypes.c_double]
        L.hydro_batch_create.argtypes = [vp, vp, ip, ctypes.POINTER(vp)]
        L.hydro_batch_destroy.argtypes = [vp]
        L.hydro_batch_set_state.argtypes = [vp, ip, dp, sz]
        L.hydro_batch_get_state.argtypes = [vp, ip, dp, sz]
        L.hydro_batch_run.argtypes = [vp, ip, ip]
        L.hydro_batch_sync.argtypes = [vp]
        L.hydro_batch_member_dt.argtypes = [vp, ip, dp]
        L.hydro_batch_interior_cells.argtypes = [vp]
        L.hydro_batch_interior_cells.restype = ctypes.c_longlong
        L.hydro_batch_member_time.argtypes = [vp, ip, dp, ctypes.POINTER(ctypes.c_longlong), dp]
        L.hydro_batch_member_health.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_uint)]
        L.hydro_batch_set_decomposition.argtypes = [vp, ip, ip]
        L.hydro_batch_autotune.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip), dp]
        L.hydro_calc_timestep_noalloc.argtypes = [vp, vp, sz, vp, ctypes.c_double, vp, vp, vp]
        L.hydro_sim_decomposition.argtypes = [vp, ctypes.POINTER(ip)]
        L.hydro_sim_autotune.argtypes = [vp, ip, ctypes.POINTER(ip), dp]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise HydroError("hydro error %d: %s" % (rc, lib().hydro_last_error().decode()))
    return rc


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


_ctx = {}


def context(device=0):
    """Process-wide context per device."""
    if device not in _ctx:
        h = ctypes.c_void_p()
        check(lib().hydro_ctx_create(device, ctypes.byref(h)))
        _ctx[device] = h
    return _ctx[device]


class Sim:
    """One grid on the GPU through the C ABI."""

    def __init__(self, desc, device=0):
        self.desc = desc
        h = ctypes.c_void_p()
        check(lib().hydro_sim_create(context(device), ctypes.byref(desc), ctypes.byref(h)))
        self._h = h
        self.N = lib().hydro_sim_num_states(h)
        self.volume = lib().hydro_sim_volume(h)

    def close(self):
        if getattr(self, "_h", None):
            lib().hydro_sim_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, aos):
        a = np.ascontiguousarray(aos, dtype=np.float64)
        check(lib().hydro_sim_set_state(self._h, _dp(a), a.size))

    def get_state(self):
        out = np.empty(self.N * self.volume)
        check(lib().hydro_sim_get_state(self._h, _dp(out), out.size))
        return out

    def set_potential(self, pot=None):
        """SelfGravitation::resetState after set_state: the initial potential
        (None: zero; with use_gravity the guess -rho, then relaxed), added
        into every cell's total energy."""
        if pot is None:
            check(lib().hydro_sim_set_potential(self._h, None, 0))
        else:
            a = np.ascontiguousarray(pot, dtype=np.float64).ravel()
            check(lib().hydro_sim_set_potential(self._h, _dp(a), a.size))

    def reset_state(self, state, potential=None, solid=None):
        """Solver::resetState's upload (src/Solver/Solver.cpp:351-392 and, for
        the Euler solvers, SelfGravitation::resetState): the state, then the
        potential channel (added into the energy; relaxed from -rho with
        useGravity) and the solid channel, when the entry gives them."""
        self.set_state(state)
        if self.desc.equation != config.EQUATIONS["Euler"]:
            return
        pot = None if potential is None else np.asarray(potential, dtype=np.float64).ravel()
        if (pot is not None and np.any(pot != 0)) or self.desc.use_gravity:
            self.set_potential(pot)
        if solid is not None and np.any(np.asarray(solid) != 0):
            self.set_solid(solid)

    def set_solid(self, solid=None):
        """The solid flag per cell (nonzero = solid; None clears them), read
        by EulerRoe and EulerBurgers."""
        if solid is None:
            check(lib().hydro_sim_set_solid(self._h, None, 0))
        else:
            a = np.ascontiguousarray(np.asarray(solid).ravel() != 0, dtype=np.uint8)
            check(lib().hydro_sim_set_solid(self._h, a.ctypes.data, a.size))

    def get_solid(self):
        """The solid flag of every cell (uint8, 0 or 1)."""
        out = np.zeros(self.volume, dtype=np.uint8)
        check(lib().hydro_sim_get_solid(self._h, out.ctypes.data, out.size))
        return out

    def get_potential(self):
        out = np.empty(self.volume)
        check(lib().hydro_sim_get_potential(self._h, _dp(out), out.size))
        return out

    def get_primitives(self):
        out = np.empty(self.N * self.volume)
        check(lib().hydro_sim_get_primitives(self._h, _dp(out), out.size))
        return out

    def get_storage(self):
        """(high words as float64, low words as float32) of compensated FP32
        storage; (state, zeros) otherwise."""
        hi = np.empty(self.N * self.volume)
        lo = np.empty(self.N * self.volume, dtype=np.float32)
        check(lib().hydro_sim_get_storage(self._h, _dp(hi), lo.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                          hi.size))
        return hi, lo

    def restore(self, aos, t, steps, prims=None, low=None):
        """Install a state read back by get_state / get_storage (with its low
        words, and for SRHD its primitives) unchanged, with the clock at
        (t, steps)."""
        a = np.ascontiguousarray(aos, dtype=np.float64)
        p = None if prims is None else np.ascontiguousarray(prims, dtype=np.float64)
        lo = None if low is None else np.ascontiguousarray(low, dtype=np.float32)
        for x in (p, lo):
            if x is not None and x.size != a.size:
                raise HydroError("primitives and low words must have the state's size")
        fp = None if lo is None else lo.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        check(lib().hydro_sim_restore(self._h, _dp(a), fp, None if p is None else _dp(p), a.size, float(t),
                                      int(steps)))

    def boundary(self):
        check(lib().hydro_sim_boundary(self._h))

    def calc_timestep(self):
        dt = ctypes.c_double()
        check(lib().hydro_sim_calc_timestep(self._h, ctypes.byref(dt)))
        return dt.value

    def step(self, dt):
        check(lib().hydro_sim_step(self._h, float(dt)))

    def update(self):
        dt = ctypes.c_double()
        check(lib().hydro_sim_update(self._h, ctypes.byref(dt)))
        return dt.value

    def run(self, nsteps):
        check(lib().hydro_sim_run(self._h, int(nsteps)))

    def sync(self):
        check(lib().hydro_sim_sync(self._h))

    def time(self):
        t = ctypes.c_double()
        n = ctypes.c_longlong()
        dt = ctypes.c_double()
        check(lib().hydro_sim_time(self._h, ctypes.byref(t), ctypes.byref(n), ctypes.byref(dt)))
        return t.value, n.value, dt.value

    def totals(self):
        out = np.empty(self.N)
        check(lib().hydro_sim_totals(self._h, _dp(out)))
        return out

    def health(self):
        f = ctypes.c_uint()
        check(lib().hydro_sim_health(self._h, ctypes.byref(f)))
        return f.value

    # generic stage evaluation (device buffers are opaque handles here)
    def element_size(self):
        return lib().hydro_sim_element_size(self._h)

    def buffer_create(self):
        b = ctypes.c_void_p()
        check(lib().hydro_sim_buffer_create(self._h, ctypes.byref(b)))
        return b

    def buffer_destroy(self, buf):
        check(lib().hydro_sim_buffer_destroy(self._h, buf))

    def state_buffer(self):
        b = ctypes.c_void_p()
        check(lib().hydro_sim_state_buffer(self._h, ctypes.byref(b)))
        return b

    def init_step(self):
        check(lib().hydro_sim_init_step(self._h))

    def calc_deriv(self, dt, deriv):
        check(lib().hydro_sim_calc_deriv(self._h, float(dt), deriv))

    def buffer_copy(self, dst, src):
        check(lib().hydro_sim_buffer_copy(self._h, dst, src))

    def buffer_zero(self, buf):
        check(lib().hydro_sim_buffer_zero(self._h, buf))

    def mult_add(self, a, b, c, k):
        check(lib().hydro_sim_mult_add(self._h, a, b, c, float(k)))

    def finish_step(self, dt):
        check(lib().hydro_sim_finish_step(self._h, float(dt)))

    def set_profiling(self, enable):
        check(lib().hydro_sim_set_profiling(self._h, 1 if enable else 0))

    def stage_timing(self):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        check(lib().hydro_sim_stage_timing(self._h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def phase_timing(self):
        """{phase: (total ms, intervals)} of the device time spent per phase
        of the updates since profiling was enabled (PHASES names them)."""
        ms = (ctypes.c_double * len(PHASES))()
        n = (ctypes.c_longlong * len(PHASES))()
        check(lib().hydro_sim_phase_timing(self._h, ms, n))
        return {name: (ms[i], n[i]) for i, name in enumerate(PHASES)}

    def decomposition(self):
        """(rows, waves per workgroup, 3D form, its waves per workgroup) of
        the stage kernels."""
        out = (ctypes.c_int * 4)()
        check(lib().hydro_sim_decomposition(self._h, out))
        return tuple(out)

    def autotune(self, updates=4):
        """Time the candidate decompositions, keep the fastest (the state is
        restored); returns (decomposition, ms per update)."""
        out = (ctypes.c_int * 4)()
        ms = ctypes.c_double()
        check(lib().hydro_sim_autotune(self._h, int(updates), out, ctypes.byref(ms)))
        return tuple(out), ms.value

    @property
    def stream(self):
        return lib().hydro_sim_stream(self._h)


EXPORTED_SYMBOLS = [
    "hydro_sim_desc_default", "hydro_last_error", "hydro_version", "hydro_ctx_create",
    "hydro_ctx_destroy", "hydro_sim_create", "hydro_sim_destroy", "hydro_sim_num_states",
    "hydro_sim_volume", "hydro_sim_set_state", "hydro_sim_get_state", "hydro_sim_get_primitives",
    "hydro_sim_restore", "hydro_sim_get_storage",
    "hydro_sim_boundary", "hydro_sim_calc_timestep", "hydro_sim_step", "hydro_sim_update",
    "hydro_sim_run", "hydro_sim_sync", "hydro_sim_time", "hydro_sim_totals", "hydro_sim_health",
    "hydro_sim_stream", "hydro_sim_set_profiling", "hydro_sim_stage_timing", "hydro_sim_phase_timing",
    "hydro_sim_set_potential", "hydro_sim_get_potential", "hydro_sim_set_solid", "hydro_sim_get_solid",
    "hydro_sim_element_size", "hydro_integrator_tableau", "hydro_sim_buffer_create", "hydro_sim_buffer_destroy",
    "hydro_sim_state_buffer", "hydro_sim_init_step", "hydro_sim_calc_deriv", "hydro_sim_buffer_copy",
    "hydro_sim_buffer_zero", "hydro_sim_mult_add", "hydro_sim_finish_step",
    "hydro_batch_create", "hydro_batch_destroy", "hydro_batch_set_state",
    "hydro_batch_get_state", "hydro_batch_run", "hydro_batch_sync", "hydro_batch_member_dt",
    "hydro_batch_interior_cells", "hydro_batch_member_time", "hydro_batch_member_health",
    "hydro_batch_set_decomposition", "hydro_batch_autotune", "hydro_calc_timestep_noalloc",
    "hydro_sim_decomposition", "hydro_sim_autotune",
]

def tableau(integrator):
    """(alphas, betas) of an integrator id, each (order, order)
    (include/HydroGPU/Integrator/RungeKutta.h:129-294)."""
    order = ctypes.c_int()
    A = np.zeros(16)
    B = np.zeros(16)
    check(lib().hydro_integrator_tableau(int(integrator), ctypes.byref(order), _dp(A), _dp(B)))
    n = order.value
    return A.reshape(4, 4)[:n, :n].copy(), B.reshape(4, 4)[:n, :n].copy()


def integrate(sim, dt, callback=None):
    """One step of the simulation's integrator with the combination on the
    host side, as the reference's Integrator::integrate(dt, callback) does
    (include/HydroGPU/Integrator/RungeKutta.h:59-125,
    src/Integrator/ForwardEuler.cpp:8-32): u(0) = state; for each stage the
    callback fills a zeroed derivative buffer for the state buffer's current
    contents; the state is rebuilt as sum_k alphas(i-1,k) u(k) +
    dt betas(i-1,k) L(u(k)) in that order, skipping zero coefficients; stage
    states and derivatives are kept only when a later row uses them.  The
    default callback is the solver's derivative (Roe::step's lambda,
    src/Solver/Roe.cpp:94-99).  Frames the step with init_step /
    finish_step."""
    if callback is None:
        def callback(deriv):
            sim.calc_deriv(dt, deriv)
    A, B = tableau(sim.desc.integrator)
    n = A.shape[0]
    state = sim.state_buffer()
    sim.init_step()
    bufs = []

    def new():
        b = sim.buffer_create()
        bufs.append(b)
        return b

    try:
        u = [new() if any(A[m, i] != 0 for m in range(i, n)) else None for i in range(n)]
        d = [new() if any(B[m, i] != 0 for m in range(i, n)) else None for i in range(n)]
        if u[0] is not None:
            sim.buffer_copy(u[0], state)
        if d[0] is not None:
            sim.buffer_zero(d[0])
            callback(d[0])
        for i in range(1, n + 1):
            sim.buffer_zero(state)
            for k in range(i):
                if A[i - 1, k]:
                    sim.mult_add(state, state, u[k], A[i - 1, k])
                if B[i - 1, k]:
                    sim.mult_add(state, state, d[k], B[i - 1, k] * dt)
            if i < n:
                if u[i] is not None:
                    sim.buffer_copy(u[i], state)
                if d[i] is not None:
                    sim.buffer_zero(d[i])
                    callback(d[i])
        sim.finish_step(dt)
    finally:
        sim.sync()
        for b in bufs:
            sim.buffer_destroy(b)


__all__ = ["Sim", "Batch", "HydroError", "lib", "check", "context", "config", "EXPORTED_SYMBOLS", "LIB_PATH",
           "tableau", "integrate"]


class Batch:
    """Many independent Euler grids advanced together (hydro_batch_*)."""

    def __init__(self, descs, device=0):
        self.descs = list(descs)
        arr = (config.HydroSimDesc * len(self.descs))(*self.descs)
        self._h = ctypes.c_void_p()
        check(lib().hydro_batch_create(context(device), ctypes.byref(arr), len(self.descs), ctypes.byref(self._h)))
        self.sizes = [int(np.prod(list(d.size)[:d.dim])) * (d.dim + 2) for d in self.descs]

    def close(self):
        if self._h:
            lib().hydro_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, member, aos):
        aos = np.ascontiguousarray(aos, dtype=np.float64)
        check(lib().hydro_batch_set_state(self._h, member, _dp(aos), aos.size))

    def get_state(self, member):
        out = np.empty(self.sizes[member])
        check(lib().hydro_batch_get_state(self._h, member, _dp(out), out.size))
        return out

    def run(self, nsteps, graph=True):
        check(lib().hydro_batch_run(self._h, nsteps, 1 if graph else 0))

    def sync(self):
        check(lib().hydro_batch_sync(self._h))

    def time(self, member):
        t, dt = ctypes.c_double(), ctypes.c_double()
        n = ctypes.c_longlong()
        check(lib().hydro_batch_member_time(self._h, member, ctypes.byref(t), ctypes.byref(n), ctypes.byref(dt)))
        return t.value, n.value, dt.value

    def health(self, member):
        h = ctypes.c_uint()
        check(lib().hydro_batch_member_health(self._h, member, ctypes.byref(h)))
        return h.value

    def interior_cells(self):
        return lib().hydro_batch_interior_cells(self._h)

    def set_decomposition(self, rows, strips):
        check(lib().hydro_batch_set_decomposition(self._h, rows, strips))

    def autotune(self):
        r, k = ctypes.c_int(), ctypes.c_int()
        ms = ctypes.c_double()
        check(lib().hydro_batch_autotune(self._h, ctypes.byref(r), ctypes.byref(k), ctypes.byref(ms)))
        return r.value, k.value, ms.value
