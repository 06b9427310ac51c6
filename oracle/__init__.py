"""TEST INFRASTRUCTURE ONLY: ctypes binding of the oracle restatement.

The oracle (hydro_oracle.c) is the FP64 CPU checker of the HIP update path.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB names another build of the same sources (the sanitizer build
# of tests/test_sanitizers.py)
_LIB_PATH = os.environ.get("ORACLE_LIB") or os.path.join(_HERE, "liboracle_hydro.so")
_lib = None

ORACLE_OPT_FRESH_BASIS_FIXED_DT = 0x1


def build(force=False):
    """Compile liboracle_hydro.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", _HERE])
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, dp, ip, sz = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_size_t
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [vp, ctypes.c_uint]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_num_states.argtypes = [vp]
        L.oracle_volume.argtypes = [vp]
        L.oracle_volume.restype = sz
        for n in ("oracle_set_state", "oracle_get_state", "oracle_get_primitives", "oracle_set_potential",
                  "oracle_get_potential"):
            getattr(L, n).argtypes = [vp, dp, sz]
            getattr(L, n).restype = ip
        L.oracle_boundary.argtypes = [vp]
        L.oracle_init_step.argtypes = [vp]
        L.oracle_calc_timestep.argtypes = [vp]
        L.oracle_calc_timestep.restype = ctypes.c_double
        L.oracle_step.argtypes = [vp, ctypes.c_double]
        L.oracle_update.argtypes = [vp]
        L.oracle_update.restype = ctypes.c_double
        L.oracle_time.argtypes = [vp]
        L.oracle_time.restype = ctypes.c_double
        L.oracle_set_threads.argtypes = [ip]
        L.oracle_slope_limiter.argtypes = [ip, ctypes.c_double]
        L.oracle_slope_limiter.restype = ctypes.c_double
        L.oracle_kernel_dx.argtypes = [ctypes.c_double, ip]
        L.oracle_kernel_dx.restype = ctypes.c_double
        L.oracle_rk_tableau.argtypes = [ip, ctypes.POINTER(ip), dp, dp]
        L.oracle_euler_eigenbasis.argtypes = [ip, ip, ctypes.c_double, dp, dp, ctypes.c_double,
                                              ctypes.c_double, dp, dp, dp]
        L.oracle_euler_eigenbasis_solid.argtypes = [ip, ip, ctypes.c_double, dp, dp, ctypes.c_double,
                                                    ctypes.c_double, ip, ip, dp, dp, dp]
        L.oracle_set_solid.argtypes = [vp, ctypes.c_void_p, sz]
        L.oracle_srhd_eigenbasis.argtypes = [ip, ip, ctypes.c_double, dp, dp, dp, dp, dp]
        L.oracle_euler_hll_speeds.argtypes = [ip, ip, ip, ctypes.c_double, dp, dp, dp]
        L.oracle_euler_hll_flux.argtypes = [ip, ip, ip, ctypes.c_double, dp, dp, dp, dp]
        L.oracle_euler_hll_flux.restype = ip
        L.oracle_euler_hll_speeds_pot.argtypes = [ip, ip, ip, ctypes.c_double, dp, dp, ctypes.c_double,
                                                  ctypes.c_double, dp]
        L.oracle_euler_hll_flux_pot.argtypes = [ip, ip, ip, ctypes.c_double, dp, dp, ctypes.c_double,
                                                ctypes.c_double, dp, dp]
        L.oracle_euler_hll_flux_pot.restype = ip
        L.oracle_srhd_con2prim.argtypes = [vp, ip, dp, dp]
        L.oracle_srhd_prim2cons.argtypes = [ip, ctypes.c_double, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleSim:
    """One grid advanced by the oracle (reference dispatch order)."""

    def __init__(self, desc, opts=0):
        self.desc = desc
        self._h = lib().oracle_create(ctypes.byref(desc), opts)
        if not self._h:
            raise ValueError("oracle_create rejected the descriptor")
        self.N = lib().oracle_num_states(self._h)
        self.volume = lib().oracle_volume(self._h)

    def close(self):
        if self._h:
            lib().oracle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, aos):
        a = f64(aos)
        if lib().oracle_set_state(self._h, _dp(a), a.size) != 0:
            raise ValueError("bad state size")

    def set_potential(self, pot=None):
        """SelfGravitation::resetState after set_state: the initial potential
        (None: zero; ignored with useGravity, whose guess is -rho and which
        relaxes it), then that initial potential added into the energy."""
        if pot is None:
            rc = lib().oracle_set_potential(self._h, None, 0)
        else:
            a = f64(pot)
            rc = lib().oracle_set_potential(self._h, _dp(a), a.size)
        if rc != 0:
            raise ValueError("bad potential")

    def set_solid(self, solid=None):
        """solidBuffer: one flag per cell (nonzero = solid), None clears it."""
        if solid is None:
            rc = lib().oracle_set_solid(self._h, None, 0)
        else:
            a = np.ascontiguousarray(np.asarray(solid).ravel() != 0, dtype=np.uint8)
            rc = lib().oracle_set_solid(self._h, a.ctypes.data, a.size)
        if rc != 0:
            raise ValueError("bad solid mask")

    def get_potential(self):
        out = np.empty(self.volume)
        lib().oracle_get_potential(self._h, _dp(out), out.size)
        return out

    def get_state(self):
        out = np.empty(self.N * self.volume)
        lib().oracle_get_state(self._h, _dp(out), out.size)
        return out

    def get_primitives(self):
        out = np.empty(self.N * self.volume)
        if lib().oracle_get_primitives(self._h, _dp(out), out.size) != 0:
            raise ValueError("no primitives for this equation")
        return out

    def boundary(self):
        lib().oracle_boundary(self._h)

    def init_step(self):
        lib().oracle_init_step(self._h)

    def calc_timestep(self):
        return lib().oracle_calc_timestep(self._h)

    def step(self, dt):
        lib().oracle_step(self._h, dt)

    def update(self):
        return lib().oracle_update(self._h)

    def run(self, nsteps):
        dts = [self.update() for _ in range(nsteps)]
        return dts

    @property
    def time(self):
        return lib().oracle_time(self._h)


def slope_limiter(limiter, r):
    return lib().oracle_slope_limiter(int(limiter), float(r))


def kernel_dx(dx, six_digits=True):
    return lib().oracle_kernel_dx(float(dx), 1 if six_digits else 0)


def rk_tableau(integrator):
    order = ctypes.c_int()
    A = np.zeros(16)
    B = np.zeros(16)
    if lib().oracle_rk_tableau(int(integrator), ctypes.byref(order), _dp(A), _dp(B)) != 0:
        raise ValueError("unknown integrator")
    n = order.value
    return A.reshape(4, 4)[:n, :n].copy(), B.reshape(4, 4)[:n, :n].copy()


def euler_eigenbasis(dim, side, gamma, qL, qR, potL=0.0, potR=0.0):
    N = dim + 2
    ev = np.zeros(N)
    L = np.zeros(N * N)
    R = np.zeros(N * N)
    lib().oracle_euler_eigenbasis(dim, side, gamma, _dp(f64(qL)), _dp(f64(qR)), potL, potR,
                                  _dp(ev), _dp(L), _dp(R))
    # column-major m[i + N*j] -> numpy [i, j]
    return ev, L.reshape(N, N).T.copy(), R.reshape(N, N).T.copy()


def euler_eigenbasis_solid(dim, side, gamma, qL, qR, potL=0.0, potR=0.0, solidL=0, solidR=0):
    N = dim + 2
    ev = np.zeros(N)
    L = np.zeros(N * N)
    R = np.zeros(N * N)
    lib().oracle_euler_eigenbasis_solid(dim, side, gamma, _dp(f64(qL)), _dp(f64(qR)), potL, potR, int(solidL),
                                        int(solidR), _dp(ev), _dp(L), _dp(R))
    return ev, L.reshape(N, N).T.copy(), R.reshape(N, N).T.copy()


def srhd_eigenbasis(dim, side, gamma, primL, primR):
    N = dim + 2
    ev = np.zeros(N)
    L = np.zeros(N * N)
    R = np.zeros(N * N)
    lib().oracle_srhd_eigenbasis(dim, side, gamma, _dp(f64(primL)), _dp(f64(primR)),
                                 _dp(ev), _dp(L), _dp(R))
    return ev, L.reshape(N, N).T.copy(), R.reshape(N, N).T.copy()


def hll_speeds(solver, dim, side, gamma, qL, qR, potL=0.0, potR=0.0):
    ev = np.zeros(dim + 2)
    lib().oracle_euler_hll_speeds_pot(solver, dim, side, gamma, _dp(f64(qL)), _dp(f64(qR)), potL, potR, _dp(ev))
    return ev


def hll_flux(solver, dim, side, gamma, qL, qR, ev, potL=0.0, potR=0.0):
    f = np.full(dim + 2, np.nan)
    ok = lib().oracle_euler_hll_flux_pot(solver, dim, side, gamma, _dp(f64(qL)), _dp(f64(qR)), potL, potR,
                                         _dp(f64(ev)), _dp(f))
    return f if ok else None


def srhd_con2prim(desc, cons, prim_guess):
    prim = f64(prim_guess).copy()
    it = lib().oracle_srhd_con2prim(ctypes.byref(desc), desc.dim, _dp(f64(cons)), _dp(prim))
    return it, prim


def srhd_prim2cons(dim, gamma, newtonian):
    N = dim + 2
    cons = np.zeros(N)
    prim = np.zeros(N)
    lib().oracle_srhd_prim2cons(dim, gamma, _dp(f64(newtonian)), _dp(cons), _dp(prim))
    return cons, prim
