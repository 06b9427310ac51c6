"""Run configuration: the reference's Lua config keys mapped onto hydro_sim_desc.

The reference reads `size`, `xmin`, `xmax`, `boundaryMethods`, `solverName`,
`integratorName`, `slopeLimiter`, `cfl`, `useFixedDT`, `fixedDT` and the `defs`
table from res/config.lua (src/HydroGPUApp.cpp:164-230, src/Solver/Solver.cpp:178-204,
285-307).  `make_desc` accepts the same keys and fills the C struct declared in
include/hydro_capi.h.
"""
import ctypes

# include/hydro_capi.h enums -------------------------------------------------

EQUATIONS = {"Euler": 0, "SRHD": 1, "Maxwell": 2}
SOLVERS = {"Roe": 0, "HLL": 1, "HLLC": 2, "Burgers": 3}

# res/SlopeLimiter.cl:5-68, in order
LIMITERS = [
    "DonorCell", "LaxWendroff", "BeamWarming", "Fromm", "CHARM", "HCUS", "HQUICK",
    "Koren", "MinMod", "Oshker", "Ospre", "Smart", "Sweby", "UMIST", "VanAlbada1",
    "VanAlbada2", "VanLeer", "MonotizedCentral", "Superbee", "BarthJespersen",
]

# src/Solver/Solver.cpp:182-194 registry names (BackwardEulerConjugateGradient
# is unfinished in the reference and is not offered)
INTEGRATORS = [
    "ForwardEuler", "RungeKutta2", "RungeKutta2Heun", "RungeKutta2Ralston", "RungeKutta3",
    "RungeKutta4", "RungeKutta4_3_8thsRule", "RungeKutta2TVD", "RungeKutta2NonTVD",
    "RungeKutta3TVD", "RungeKutta4TVD", "RungeKutta4NonTVD",
]

# include/HydroGPU/Boundary/Boundary.h:8-14
BOUNDARY = {"NONE": -1, "PERIODIC": 0, "MIRROR": 1, "FREEFLOW": 2}

PRECISION = {"fp64": 0, "fp32_mixed": 1}

FLAG_REF_DX_6DIGITS = 0x1
FLAG_GRAVITY_IDENTITY_PASS = 0x2
FLAG_KAHAN = 0x4
FLAG_CLIP_TO_TEND = 0x8
FLAG_REFERENCE = 0x3

# solverName -> (equation, solver)   (src/HydroGPUApp.cpp:90-119)
SOLVER_NAMES = {
    "EulerRoe": ("Euler", "Roe"),
    "EulerHLL": ("Euler", "HLL"),
    "EulerHLLC": ("Euler", "HLLC"),
    "EulerBurgers": ("Euler", "Burgers"),
    "SRHDRoe": ("SRHD", "Roe"),
    "MaxwellRoe": ("Maxwell", "Roe"),
}


class HydroSimDesc(ctypes.Structure):
    """Mirror of `hydro_sim_desc` (include/hydro_capi.h)."""

    _fields_ = [
        ("dim", ctypes.c_int),
        ("size", ctypes.c_int * 3),
        ("xmin", ctypes.c_double * 3),
        ("xmax", ctypes.c_double * 3),
        ("equation", ctypes.c_int),
        ("solver", ctypes.c_int),
        ("limiter", ctypes.c_int),
        ("integrator", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("boundary", (ctypes.c_int * 2) * 3),
        ("cfl", ctypes.c_double),
        ("use_fixed_dt", ctypes.c_int),
        ("fixed_dt", ctypes.c_double),
        ("t_end", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("srhd_solve_prim_max_iter", ctypes.c_int),
        ("srhd_solve_prim_stop_epsilon", ctypes.c_double),
        ("srhd_solve_prim_vel_epsilon", ctypes.c_double),
        ("srhd_solve_prim_pmin_epsilon", ctypes.c_double),
        ("srhd_rho_min", ctypes.c_double),
        ("srhd_rho_max", ctypes.c_double),
        ("srhd_eint_max", ctypes.c_double),
        ("srhd_D_min", ctypes.c_double),
        ("srhd_D_max", ctypes.c_double),
        ("srhd_tau_min", ctypes.c_double),
        ("srhd_tau_max", ctypes.c_double),
        ("maxwell_permittivity", ctypes.c_double),
        ("maxwell_permeability", ctypes.c_double),
        ("maxwell_conductivity", ctypes.c_double),
        ("flags", ctypes.c_uint),
        ("use_gravity", ctypes.c_int),
        ("gauss_seidel_max_iter", ctypes.c_int),
        ("self_grav_gravitational_constant", ctypes.c_double),
    ]


# res/config.lua:70-93
DEFAULT_DEFS = {
    "idealGas_heatCapacityRatio": 1.4,
    "srhd_solvePrimMaxIter": 1000,
    "srhd_solvePrimStopEpsilon": 1e-7,
    "srhd_solvePrimVelEpsilon": 1e-15,
    "srhd_solvePrimPMinEpsilon": 1e-16,
    "srhd_rhoMin": 1e-15,
    "srhd_rhoMax": 1e20,
    "srhd_eIntMax": 1e20,
    "srhd_DMin": 1e-15,
    "srhd_DMax": 1e20,
    "srhd_tauMin": 1e-15,
    "srhd_tauMax": 1e20,
    "maxwell_permittivity": 1.0,
    "maxwell_permeability": 1.0,
    "maxwell_conductivity": 1.0,
    "selfGrav_gravitationalConstant": 1.0,
}


def lua_number(v):
    """Lua's tostring of a number ("%.14g"), which is how `defs` reach the
    kernel source (src/Solver/Solver.cpp:289-307)."""
    return float("%.14g" % float(v))


def make_desc(size, xmin=None, xmax=None, solverName="EulerRoe", integratorName="ForwardEuler",
              slopeLimiter="Superbee", boundaryMethods=None, cfl=None, useFixedDT=False,
              fixedDT=0.001, defs=None, precision="fp64", flags=FLAG_REFERENCE, t_end=0.0,
              useGravity=False, gaussSeidelMaxIter=20):
    """Build a hydro_sim_desc from reference-style config keys.

    `size` includes the 2+2 ghost cells per axis, as in the reference.
    `boundaryMethods` is a list of {"min": name, "max": name} per axis.
    Defaults follow res/config.lua (xmin=-.5, xmax=.5, FREEFLOW, cfl = 1/#size,
    limiter Superbee -- the limiter key the reference reads is `slopeLimiter`,
    src/Solver/Solver.cpp:285-287).
    """
    size = list(size)
    dim = len(size)
    if not 1 <= dim <= 3:
        raise ValueError("size must have 1..3 entries")
    d = HydroSimDesc()
    d.dim = dim
    for a in range(3):
        d.size[a] = size[a] if a < dim else 1
    xmin = list(xmin) if xmin is not None else [-0.5] * 3
    xmax = list(xmax) if xmax is not None else [0.5] * 3
    for a in range(3):
        d.xmin[a] = xmin[a] if a < len(xmin) else -0.5
        d.xmax[a] = xmax[a] if a < len(xmax) else 0.5
    if solverName not in SOLVER_NAMES:
        raise ValueError("unknown solverName %r" % solverName)
    eqn, sol = SOLVER_NAMES[solverName]
    d.equation = EQUATIONS[eqn]
    d.solver = SOLVERS[sol]
    if slopeLimiter not in LIMITERS:
        raise ValueError("unknown slopeLimiter %r" % slopeLimiter)
    d.limiter = LIMITERS.index(slopeLimiter)
    if integratorName not in INTEGRATORS:
        raise ValueError("failed to find integrator named %s" % integratorName)
    d.integrator = INTEGRATORS.index(integratorName)
    d.precision = PRECISION[precision]
    bm = boundaryMethods or [{"min": "FREEFLOW", "max": "FREEFLOW"}] * 3
    for a in range(3):
        entry = bm[a] if a < len(bm) else {"min": "FREEFLOW", "max": "FREEFLOW"}
        d.boundary[a][0] = BOUNDARY[entry["min"]]
        d.boundary[a][1] = BOUNDARY[entry["max"]]
    d.cfl = float(cfl) if cfl is not None else 1.0 / dim
    d.use_fixed_dt = 1 if useFixedDT else 0
    d.fixed_dt = float(fixedDT)
    d.t_end = float(t_end)
    dd = dict(DEFAULT_DEFS)
    if defs:
        dd.update(defs)
    d.gamma = lua_number(dd["idealGas_heatCapacityRatio"])
    d.srhd_solve_prim_max_iter = int(dd["srhd_solvePrimMaxIter"])
    d.srhd_solve_prim_stop_epsilon = lua_number(dd["srhd_solvePrimStopEpsilon"])
    d.srhd_solve_prim_vel_epsilon = lua_number(dd["srhd_solvePrimVelEpsilon"])
    d.srhd_solve_prim_pmin_epsilon = lua_number(dd["srhd_solvePrimPMinEpsilon"])
    d.srhd_rho_min = lua_number(dd["srhd_rhoMin"])
    d.srhd_rho_max = lua_number(dd["srhd_rhoMax"])
    d.srhd_eint_max = lua_number(dd["srhd_eIntMax"])
    d.srhd_D_min = lua_number(dd["srhd_DMin"])
    d.srhd_D_max = lua_number(dd["srhd_DMax"])
    d.srhd_tau_min = lua_number(dd["srhd_tauMin"])
    d.srhd_tau_max = lua_number(dd["srhd_tauMax"])
    d.maxwell_permittivity = lua_number(dd["maxwell_permittivity"])
    d.maxwell_permeability = lua_number(dd["maxwell_permeability"])
    d.maxwell_conductivity = lua_number(dd["maxwell_conductivity"])
    d.flags = int(flags)
    # src/HydroGPUApp.cpp:77-78, 207-208 (useGravity false, 20 sweeps by default)
    d.use_gravity = 1 if useGravity else 0
    d.gauss_seidel_max_iter = int(gaussSeidelMaxIter)
    d.self_grav_gravitational_constant = lua_number(dd["selfGrav_gravitationalConstant"])
    return d


def num_states(desc):
    if desc.equation == EQUATIONS["Maxwell"]:
        return 6
    return desc.dim + 2


def volume(desc):
    return desc.size[0] * desc.size[1] * desc.size[2]


def interior_cells(desc):
    n = 1
    for a in range(desc.dim):
        n *= desc.size[a] - 4
    return n
