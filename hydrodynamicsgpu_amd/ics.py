"""Initial conditions in the reference's conventions.

Cell centres follow Solver::resetState (src/Solver/Solver.cpp:357-366):
    pos = (xmax - xmin) * (i + .5) / size + xmin   over ALL cells, ghosts included.
States follow buildStateEuler (res/util.lua:36-80): rho, rho*v, E with
    E = rho * (|v|^2/2 + eInt),  eInt = p / ((gamma - 1) * rho)
and Euler::readStateCell's channel map (src/Equation/Euler.cpp:78-89).
Random perturbations use a counter hash (SplitMix64 of seed ^ cell index) so
that every language reproduces the same field; the reference's crand() is
Lua's math.random and is not reproducible across interpreters.
"""
import numpy as np

from .config import num_states

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    """SplitMix64 finaliser of uint64 array `x` (Steele, Lea, Flood 2014)."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform01(seed, cells, stream=0):
    """U in [0,1) per cell: top 53 bits of splitmix64(seed ^ cell ^ stream<<56)."""
    key = np.asarray(cells, dtype=np.uint64) ^ np.uint64(seed) ^ (np.uint64(stream) << np.uint64(56))
    return (splitmix64(key) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cell_centers(desc):
    """Per-axis 1D arrays of cell centres (ghosts included)."""
    out = []
    for a in range(3):
        n = desc.size[a]
        i = np.arange(n, dtype=np.float64)
        out.append((desc.xmax[a] - desc.xmin[a]) * (i + 0.5) / float(n) + desc.xmin[a])
    return out


def grid(desc):
    """x, y, z arrays shaped (SZ, SY, SX) and flat cell indices."""
    cx, cy, cz = cell_centers(desc)
    z, y, x = np.meshgrid(cz, cy, cx, indexing="ij")
    cells = np.arange(x.size, dtype=np.uint64).reshape(x.shape)
    return x, y, z, cells


def build_state_euler(desc, rho, vx, vy, vz, p, gamma=None):
    """AoS state[s + N*cell] from primitive fields (arrays shaped (SZ,SY,SX))."""
    g = desc.gamma if gamma is None else gamma
    dim = desc.dim
    vy = vy if dim > 1 else np.zeros_like(rho)
    vz = vz if dim > 2 else np.zeros_like(rho)
    eint = p / ((g - 1.0) * rho)
    ekin = 0.5 * (vx * vx + vy * vy + vz * vz)
    E = rho * (ekin + eint)
    comps = [rho, rho * vx]
    if dim > 1:
        comps.append(rho * vy)
    if dim > 2:
        comps.append(rho * vz)
    comps.append(E)
    N = num_states(desc)
    assert len(comps) == N
    return np.stack([c.reshape(-1) for c in comps], axis=1).reshape(-1).copy()


def sod(desc, gamma=None):
    """res/initConds.lua:62-77: rho,p = 1,1 where x<=0 (and y<=0, z<=0), else .125,.1."""
    x, y, z, _ = grid(desc)
    inside = x <= 0
    if desc.dim > 1:
        inside &= y <= 0
    if desc.dim > 2:
        inside &= z <= 0
    rho = np.where(inside, 1.0, 0.125)
    p = np.where(inside, 1.0, 0.1)
    zero = np.zeros_like(rho)
    return build_state_euler(desc, rho, zero, zero, zero, p, gamma)


def refined(desc, factor=2):
    """A grid `factor` times finer over the same interior region: the
    interior cells of `desc` split factor^dim ways, 2 ghost cells per side
    of the fine size (so the periodic period and the mirror walls stay
    where they are).  Returns a copy of the descriptor."""
    import copy

    f = copy.deepcopy(desc)
    for a in range(desc.dim):
        n = desc.size[a]
        dx = (desc.xmax[a] - desc.xmin[a]) / n
        f.size[a] = factor * (n - 4) + 4
        f.xmin[a] = desc.xmin[a] + 2 * dx - 2 * dx / factor
        f.xmax[a] = desc.xmax[a] - 2 * dx + 2 * dx / factor
    return f


def parent_cells(coarse, fine, factor=2):
    """Flat index of the coarse cell containing each fine cell (grid shape of
    `fine`), for grids related by `refined`; fine ghost cells map to the
    nearest coarse cell."""
    idx = []
    for a in range(3):
        nf = fine.size[a]
        i = np.arange(nf)
        if a < fine.dim:
            i = np.clip(2 + np.floor_divide(i - 2, factor), 0, coarse.size[a] - 1)
        idx.append(i)
    iz, iy, ix = np.meshgrid(idx[2], idx[1], idx[0], indexing="ij")
    return ((iz * coarse.size[1] + iy) * coarse.size[0] + ix).astype(np.uint64)


def shear_kh(desc, seed=1234, cells=None):
    """2D Kelvin-Helmholtz shear layer: rho=2, vx=+.5 for |y|<.25 else rho=1,
    vx=-.5; p=2.5; vy = .01 sin(4 pi x) + 1e-3 (2U-1).  `cells` (grid shape)
    names the hash counter of each cell, by default the cell's own index;
    a refined grid passes `parent_cells` so that it carries the same
    piecewise-constant noise as the grid it refines."""
    x, y, z, own = grid(desc)
    inside = np.abs(y) < 0.25
    rho = np.where(inside, 2.0, 1.0)
    vx = np.where(inside, 0.5, -0.5)
    u = uniform01(seed, own if cells is None else cells)
    vy = 0.01 * np.sin(4.0 * np.pi * x) + 1e-3 * (2.0 * u - 1.0)
    p = np.full_like(rho, 2.5)
    return build_state_euler(desc, rho, vx, vy, np.zeros_like(rho), p)


def blast(desc, seed=4321, cells=None):
    """3D blast: rho = 1 + 1e-4 (2U-1), u=0, p=10 inside r<.2 else .1
    (`cells` as for shear_kh)."""
    x, y, z, own = grid(desc)
    r = np.sqrt(x * x + y * y + z * z)
    u = uniform01(seed, own if cells is None else cells)
    rho = 1.0 + 1e-4 * (2.0 * u - 1.0)
    p = np.where(r < 0.2, 10.0, 0.1)
    zero = np.zeros_like(rho)
    return build_state_euler(desc, rho, zero, zero, zero, p)


def srhd_quadrant_states(seed=777, w_min=3.0, w_max=10.0, spread=0.05):
    """Per-quadrant (rho, p, W, vx, vy) of `srhd_quadrants`, quadrant index
    (x > centre) + 2 (y > centre)."""
    qs = np.arange(4, dtype=np.uint64)
    u0 = uniform01(seed, qs, 1)
    u1 = uniform01(seed, qs, 2)
    u2 = uniform01(seed, qs, 3)
    u3 = uniform01(seed, qs, 4)
    g = uniform01(seed, np.arange(2, dtype=np.uint64), 5)
    rho_q = 10.0 ** (-1.0 + u0)            # [0.1, 1]
    p_q = 10.0 ** (-2.0 + 2.0 * u1)         # [0.01, 1]
    W = w_min + (w_max - w_min) * u2        # Lorentz factor per quadrant
    fast = min(int(4.0 * g[0]), 3)
    W[fast] = w_max * (0.9 + 0.1 * u2[fast])
    speed = np.sqrt(1.0 - 1.0 / (W * W))
    ang = 2.0 * np.pi * g[1] + spread * (2.0 * u3 - 1.0)
    return rho_q, p_q, W, speed * np.cos(ang), speed * np.sin(ang)


def srhd_quadrants(desc, seed=777, w_min=3.0, w_max=10.0, spread=0.05):
    """Four-quadrant relativistic Riemann problem split at the domain centre.
    Each quadrant draws rho in [.1,1], p in [.01,1] (log-uniform) and a
    Lorentz factor uniform in [w_min, w_max]; one quadrant, picked by the
    hash, draws it in [.9 w_max, w_max], so the fastest flow has W >= 9 with
    the defaults (W = 9.14 for seed 777).  The velocities share one seeded
    direction, each quadrant turned from it by at most `spread` radians, so
    neighbouring quadrants move at up to ~.9 c relative to each other
    (relative Lorentz factor <= 1.4).  Quadrants that collide head-on at
    these speeds drive the reference's Roe scheme (arithmetic-mean
    linearisation, res/SRHDRoe.cl:21-312) to the density floor within its
    first step, in the FP64 oracle as on the GPU, which is why the directions
    are shared.  Returns the Newtonian (rho, rho v, E) the reference's
    initVariables converts (res/SRHDCommon.cl:24-92)."""
    x, y, z, _ = grid(desc)
    cxm = 0.5 * (desc.xmin[0] + desc.xmax[0])
    cym = 0.5 * (desc.xmin[1] + desc.xmax[1])
    q = (x > cxm).astype(np.int64) + 2 * (y > cym).astype(np.int64)
    rho_q, p_q, _, vx_q, vy_q = srhd_quadrant_states(seed, w_min, w_max, spread)
    rho = rho_q[q]
    p = p_q[q]
    vx = vx_q[q]
    vy = vy_q[q]
    return build_state_euler(desc, rho, vx, vy, np.zeros_like(rho), p)


def riemann(desc, left, right):
    """Two-state Riemann problem split at the x-centre; left/right are
    (rho, u, p) tuples (u along x)."""
    x, y, z, _ = grid(desc)
    xc = 0.5 * (desc.xmin[0] + desc.xmax[0])
    inside = x <= xc
    rho = np.where(inside, left[0], right[0])
    vx = np.where(inside, left[1], right[1])
    p = np.where(inside, left[2], right[2])
    zero = np.zeros_like(rho)
    return build_state_euler(desc, rho, vx, zero, zero, p)


def maxwell_1(desc):
    """res/initConds.lua:779-795 (Maxwell-1): E=(0,0,eps), B=(0,0,+-1)."""
    x, y, z, _ = grid(desc)
    inside = x <= 0
    if desc.dim > 1:
        inside &= y <= 0
    if desc.dim > 2:
        inside &= z <= 0
    eps = desc.maxwell_permittivity
    zero = np.zeros(x.shape)
    comps = [zero, zero, np.full(x.shape, eps), zero, zero, np.where(inside, 1.0, -1.0)]
    return np.stack([c.reshape(-1) for c in comps], axis=1).reshape(-1).copy()


def primitives_euler(desc, aos):
    """(rho, v[dim], p) from an AoS Euler state."""
    N = num_states(desc)
    q = np.asarray(aos).reshape(-1, N)
    rho = q[:, 0]
    v = q[:, 1:1 + desc.dim] / rho[:, None]
    E = q[:, desc.dim + 1]
    p = (desc.gamma - 1.0) * (E - 0.5 * rho * (v * v).sum(axis=1))
    return rho, v, p


def riemann_batch(count=512, seed=2024, min_1d=256, max_1d=65536, min_2d=64, max_2d=512,
                  max_shock=3.0, min_rarefaction=0.05, **desc_kw):
    """The uneven batch mix: `count` independent Euler grids, each a 1D line
    with length log-uniform in [min_1d, max_1d] or a 2D square with side
    log-uniform in [min_2d, max_2d] (a fair coin per member), holding a
    Riemann problem split at the x-centre with rho, p log-uniform in
    [0.1, 10] and |u| < 1 on each side.  Draws whose exact solution has a
    star pressure above `max_shock` or below `min_rarefaction` times the
    smaller side pressure are drawn again: the strongest collisions drive the
    Roe/Superbee/RK2 scheme (in the reference as here) to negative pressures
    (with a bound of 8, one member of the default 512 fails at step 59 in
    FP64 and at step 40 in FP32; with 3, every member stays positive for the
    100 FP32 steps checked).  Everything comes from
    the SplitMix64 counter hash of `seed`, so the mix is the same in every
    language.  Returns a list of (desc, AoS state); `desc_kw` go to
    config.make_desc (solver, integrator, limiter, precision, ...)."""
    from . import config
    from .riemann import star_state

    gamma = desc_kw.get("defs", {}).get("idealGas_heatCapacityRatio", config.DEFAULT_DEFS["idealGas_heatCapacityRatio"])
    lg = lambda x: float(np.exp(np.log(0.1) + x * (np.log(10.0) - np.log(0.1))))  # noqa: E731
    out = []
    for m in range(count):
        for attempt in range(64):
            r = uniform01(seed, np.arange(8, dtype=np.uint64) + np.uint64(8 * m), stream=attempt)
            left = (lg(r[2]), 2.0 * r[3] - 1.0, lg(r[4]))
            right = (lg(r[5]), 2.0 * r[6] - 1.0, lg(r[7]))
            ps, _ = star_state(left, right, gamma)
            pmin = min(left[2], right[2])
            if min_rarefaction * pmin <= ps <= max_shock * pmin:
                break
        if r[0] < 0.5:
            n = int(round(np.exp(np.log(min_1d) + r[1] * (np.log(max_1d) - np.log(min_1d)))))
            size = [n]
        else:
            n = int(round(np.exp(np.log(min_2d) + r[1] * (np.log(max_2d) - np.log(min_2d)))))
            size = [n, n]
        kw = dict(desc_kw)
        kw.setdefault("boundaryMethods", [{"min": "FREEFLOW", "max": "FREEFLOW"}] * 3)
        d = config.make_desc(size, xmin=[-0.5] * len(size), xmax=[0.5] * len(size), **kw)
        out.append((d, riemann(d, left, right)))
    return out
