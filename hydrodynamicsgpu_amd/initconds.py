"""The reference's catalogue of initial conditions, by name.

res/initConds.lua lists initial conditions as `{name, equations, setup}`;
`setup()` may change run globals (gamma, cfl, boundaryMethods, xmin/xmax,
size) and installs `initState(x, y, z)`, which Solver::resetState calls at
every cell centre, ghosts included (src/Solver/Solver.cpp:351-392).  Euler
and SRHD states come from buildStateEuler (res/util.lua:36-80), whose ten
outputs Euler::readStateCell / SRHD::readStateCell reduce to
(density, momentum[dim], total energy) (src/Equation/Euler.cpp:78-89,
src/Equation/SRHD.cpp:78-93); total energy includes the magnetic term
|B|^2 / (2 mu0) whenever an entry sets a field, as there.

Entries served: every Euler/SRHD/Maxwell entry:
  Constant, 1D Advect Wave, Sphere, Sod, Configuration 1..6, Sedov,
  Brio-Wu, Orszag-Tang, Colella-Woodward, Kelvin-Helmholtz,
  Double Mach Reflection, Square Cavity, Shock Bubble Interaction,
  Flow Around Cylinder and Forward Facing Step (solid cells), Spiral
  Implosion, Rayleigh-Taylor (an external potential), self-gravitation
  tests 1, 1 spinning, 2 and 4 (useGravity), Relativistic Blast Wave
  Interaction, Marti & Muller 2008 Problem #1/#2, Relativistic Jet,
  Maxwell-1.
Left out: the numerical-relativity entries (ADM/BSSNOK equations).
"Double Mach Reflection" is the second entry of that name: the reference's
by-name map is filled in list order (res/initConds.lua, end of file), so the
later one wins.

Usage: `setup(name, cfg)` applies an entry's setup to a config dict (keys as
config.make_desc takes them) and returns (cfg, state_fn); `state_fn(desc)`
gives the AoS state.  luaconfig runs the same setups on a Lua config's
globals.  Rayleigh-Taylor and the self-gravitation tests draw crand() noise
(seeded here, see _crand); no other entry draws random numbers.
`channels_function(name)` also gives the potential and solid channels, and
a setup may set cfg["useGravity"].
"""
import math
import zlib

import numpy as np

from .config import num_states

MHD_MU0 = 1.0  # defs.mhd_vacuumPermeability (res/config.lua)


# ---------------------------------------------------------------- state helpers
def _centres(desc):
    """Cell-centre coordinate arrays (SZ, SY, SX) of every axis, unused axes
    included: they hold one cell, whose centre the entries also test."""
    cs = []
    for a in range(3):
        n = desc.size[a]
        i = np.arange(n, dtype=np.float64)
        cs.append((desc.xmax[a] - desc.xmin[a]) * (i + 0.5) / float(n) + desc.xmin[a])
    z, y, x = np.meshgrid(cs[2], cs[1], cs[0], indexing="ij")
    return x, y, z


def _euler_state(desc, gamma, density, pressure=None, eint=None, vx=0.0, vy=0.0, vz=0.0, bx=0.0, by=0.0, bz=0.0):
    """buildStateEuler + readStateCell: AoS (density, momentum[dim], energy)."""
    shape = (desc.size[2], desc.size[1], desc.size[0])
    f = lambda v: np.broadcast_to(np.asarray(v, dtype=np.float64), shape)  # noqa: E731
    rho = f(density)
    vx = f(vx)
    vy = f(vy) if desc.dim > 1 else f(0.0)
    vz = f(vz) if desc.dim > 2 else f(0.0)
    ke = 0.5 * (vx * vx + vy * vy + vz * vz)
    if eint is None:
        eint = f(pressure) / ((gamma - 1.0) * rho)
    else:
        eint = f(eint)
    bx, by, bz = f(bx), f(by), f(bz)
    E = rho * (ke + eint) + 0.5 * (bx * bx + by * by + bz * bz) / MHD_MU0
    comps = [rho, rho * vx]
    if desc.dim > 1:
        comps.append(rho * vy)
    if desc.dim > 2:
        comps.append(rho * vz)
    comps.append(E)
    assert len(comps) == num_states(desc)
    return np.stack([np.ascontiguousarray(c).reshape(-1) for c in comps], axis=1).reshape(-1).copy()


def _quadrants(x, y, q1, q2, q3, q4):
    """buildStateEulerQuadrant (res/util.lua:82-96): y > 0 ? (x > 0 ? q1 : q2)
    : (x < 0 ? q3 : q4), per field."""
    out = {}
    for k in q1:
        out[k] = np.where(y > 0, np.where(x > 0, q1[k], q2[k]), np.where(x < 0, q3[k], q4[k]))
    return out


def _lower_octant(x, y, z):
    return (x <= 0) & (y <= 0) & (z <= 0)


# ---------------------------------------------------------------- setups
# Each setup takes the config dict and changes it as the entry's Lua setup
# changes the globals; each state function takes (desc, gamma).

def _set_bm(cfg, *pairs):
    cfg["boundaryMethods"] = [{"min": a, "max": b} for a, b in pairs]


def _set_gamma(cfg, g):
    cfg.setdefault("defs", {})["idealGas_heatCapacityRatio"] = g


FREE = ("FREEFLOW", "FREEFLOW")


def _st_constant(desc, gamma):
    return _euler_state(desc, gamma, 1.0, pressure=1.0, vx=1.0, vy=1.0, vz=1.0, bx=1.0, by=1.0, bz=1.0)


def _st_advect(desc, gamma):
    x, y, z = _centres(desc)
    rsq = x * x + y * y + z * z
    return _euler_state(desc, gamma, np.exp(-100 * rsq) + 1, pressure=1.0, vx=1.0)


def _st_sphere(desc, gamma):
    x, y, z = _centres(desc)
    inside = x * x + y * y + z * z <= .2 * .2
    return _euler_state(desc, gamma, np.where(inside, 1.0, .1), pressure=np.where(inside, 1.0, .1))


def _su_sod(cfg):
    _set_gamma(cfg, 7 / 5)


def _st_sod(desc, gamma):
    inside = _lower_octant(*_centres(desc))
    return _euler_state(desc, gamma, np.where(inside, 1.0, .125), pressure=np.where(inside, 1.0, .1))


# Kurganov & Tadmor 2D Riemann configurations: (rho, p, vx, vy) per quadrant
_CONFIGS = {
    1: ((1, 1, 0, 0), (.5197, .4, -.7259, 0), (.1072, .0439, -.7259, -1.4045), (.2579, .15, 0, -1.4045)),
    2: ((1, 1, 0, 0), (.5197, .4, -.7259, 0), (1, 1, -.7259, -.7259), (.5197, .4, 0, -.7259)),
    3: ((1.5, 1.5, 0, 0), (.5323, .3, 1.206, 0), (.138, .029, 1.206, 1.206), (.5323, .3, 0, 1.206)),
    4: ((1.1, 1.1, 0, 0), (.5065, .35, .8939, 0), (1.1, 1.1, .8939, .8939), (.5065, .35, 0, .8939)),
    5: ((1, 1, -.75, -.5), (2, 1, -.75, .5), (1, 1, .75, .5), (3, 1, .75, -.5)),
    6: ((1, 1, .75, -.5), (2, 1, .75, .5), (1, 1, -.75, .5), (3, 1, -.75, -.5)),
}


def _su_config(cfg):
    cfg["cfl"] = .475
    _set_bm(cfg, FREE, FREE, FREE)


def _st_config(k):
    def st(desc, gamma):
        x, y, _ = _centres(desc)
        qs = [dict(zip(("density", "pressure", "vx", "vy"), map(float, q))) for q in _CONFIGS[k]]
        f = _quadrants(x, y, *qs)
        return _euler_state(desc, gamma, f["density"], pressure=f["pressure"], vx=f["vx"], vy=f["vy"])
    return st


def _st_sedov(desc, gamma):
    x, y, z = _centres(desc)
    mid = [(desc.xmax[a] + desc.xmin[a]) * .5 for a in range(3)]
    dx = [(desc.xmax[a] - desc.xmin[a]) / (desc.size[a] if a < desc.dim else 1) for a in range(3)]
    inside = (np.abs(x - mid[0]) < 1.5 * dx[0]) & (np.abs(y - mid[1]) < 1.5 * dx[1]) & \
             (np.abs(z - mid[2]) < 1.5 * dx[2])
    return _euler_state(desc, gamma, 1.0, pressure=np.where(inside, 1e5, 1e-5))


def _su_briowu(cfg):
    _set_gamma(cfg, 2)


def _st_briowu(desc, gamma):
    lhs = _lower_octant(*_centres(desc))
    return _euler_state(desc, gamma, np.where(lhs, 1.0, .125), pressure=np.where(lhs, 1.0, .1), bx=.75,
                        by=np.where(lhs, 1.0, -1.0), bz=0.0)


def _su_orszag(cfg):
    _set_gamma(cfg, 5 / 3)


def _st_orszag(desc, gamma):
    x, y, _ = _centres(desc)
    b0 = 1 / math.sqrt(4 * math.pi)
    return _euler_state(desc, gamma, 25 / (36 * math.pi), pressure=5 / (12 * math.pi),
                        vx=-np.sin(2 * math.pi * (y + .5)), vy=np.sin(2 * math.pi * (x + .5)), vz=0.0,
                        bx=-b0 * np.sin(2 * math.pi * (y + .5)), by=b0 * np.sin(4 * math.pi * (x + .5)), bz=0.0)


def _su_free(cfg):
    _set_bm(cfg, FREE, FREE, FREE)


def _st_colella(desc, gamma):
    x, _, _ = _centres(desc)
    p = np.where(x < -.4, 1000.0, np.where(x < .4, .01, 100.0))
    return _euler_state(desc, gamma, 1.0, pressure=p, vx=0.0, vy=0.0, vz=0.0)


def _su_kh(cfg):
    _set_bm(cfg, ("PERIODIC", "PERIODIC"), ("PERIODIC", "PERIODIC"), ("PERIODIC", "PERIODIC"))


def _st_kh(desc, gamma):
    x, y, z = _centres(desc)
    inside = (y > -.25) & (y < .25)
    theta = (x - desc.xmin[0]) / (desc.xmax[0] - desc.xmin[0]) * 2 * math.pi
    if desc.dim >= 3:
        theta = theta * (z - desc.xmin[2]) / (desc.xmax[2] - desc.xmin[2])
    noise = desc.size[0] * 2e-5
    return _euler_state(desc, gamma, np.where(inside, 2.0, 1.0), pressure=2.5,
                        vx=np.cos(theta) * noise + np.where(inside, -.5, .5), vy=np.sin(theta) * noise,
                        vz=np.sin(theta) * noise)


def _su_dmr(cfg):
    _set_bm(cfg, FREE, ("MIRROR", "FREEFLOW"), ("MIRROR", "MIRROR"))
    cfg["xmin"] = [0.0, 0.0, 0.0]
    cfg["xmax"] = [4.0, 1.0, 1.0]
    size = list(cfg["size"])
    if len(size) < 2:
        raise ValueError("Double Mach Reflection needs a 2D or 3D size (it halves size[2])")
    size[0] = size[0] * 2
    size[1] = size[1] / 2
    if size[1] != int(size[1]):
        raise ValueError("Double Mach Reflection halves size[2], which must be even")
    cfg["size"] = [int(s) for s in size]
    _set_gamma(cfg, 1.4)


def _st_dmr(desc, gamma):
    x, y, _ = _centres(desc)
    inside = x < y * math.sqrt(1 / 3)
    c30, s30 = math.cos(math.radians(30)), math.sin(math.radians(30))
    return _euler_state(desc, gamma, np.where(inside, 8.0, 1.4), pressure=np.where(inside, 116.5, 1.0),
                        vx=np.where(inside, 8.25 * c30, 0.0), vy=np.where(inside, -8.25 * s30, 0.0))


def _su_cavity(cfg):
    _set_bm(cfg, ("MIRROR", "MIRROR"), ("MIRROR", "NONE"), ("MIRROR", "MIRROR"))


def _st_cavity(desc, gamma):
    _, y, _ = _centres(desc)
    return _euler_state(desc, gamma, 1.0, pressure=1.0, vx=np.where(y > .45, 1.0, 0.0))


def _su_bubble(cfg):
    _set_bm(cfg, FREE, FREE, FREE)
    cfg["xmin"] = [-1.0, -.5, -.5]
    cfg["xmax"] = [1.0, .5, .5]
    size = list(cfg["size"])
    size[0] = size[0] * 2
    cfg["size"] = size


def _st_bubble(desc, gamma):
    x, y, z = _centres(desc)
    wave = x < -.45
    in_bubble = x * x + y * y + z * z < .2 * .2
    return _euler_state(desc, gamma, np.where(wave, 1.0, np.where(in_bubble, .1, 1.0)),
                        pressure=np.where(wave, 1.0, .1), vx=np.where(wave, 0.0, -.5))


def _su_gamma75(cfg):
    _set_gamma(cfg, 7 / 5)


def _st_rel_blast(desc, gamma):
    xs = _centres(desc)
    lhs = np.ones(xs[0].shape, dtype=bool)
    rhs = np.ones(xs[0].shape, dtype=bool)
    for i in range(desc.dim):
        lhs &= xs[i] <= .9 * desc.xmin[i] + .1 * desc.xmax[i]
        rhs &= xs[i] >= .1 * desc.xmin[i] + .9 * desc.xmax[i]
    return _euler_state(desc, gamma, 1.0, pressure=np.where(lhs, 1e3, np.where(rhs, 1e2, 1e-2)))


def _su_gamma53(cfg):
    _set_gamma(cfg, 5 / 3)


def _st_mm1(desc, gamma):
    inside = _lower_octant(*_centres(desc))
    return _euler_state(desc, gamma, np.where(inside, 10.0, 1.0), eint=np.where(inside, 2.0, 1e-6))


def _st_mm2(desc, gamma):
    inside = _lower_octant(*_centres(desc))
    return _euler_state(desc, gamma, 1.0, pressure=np.where(inside, 1e3, 1e-2))


def _st_jet(desc, gamma):
    x, y, _ = _centres(desc)
    xmin, xmax = desc.xmin, desc.xmax
    height = .05
    width = (xmax[0] - xmin[0]) / desc.size[0] * 10
    jet = (x < xmin[0] + width) & (y > (.5 + height) * xmin[1] + (.5 - height) * xmax[1]) & \
          (y < (.5 - height) * xmin[1] + (.5 + height) * xmax[1])
    dx = (x - xmin[0]) / width
    dy = (y - .5 * (xmax[1] + xmin[1])) / (height * (xmax[1] - xmin[1]))
    coeff = np.exp(-(dx * dx + dy * dy))
    return _euler_state(desc, gamma, 1.0, pressure=np.where(jet, 1 + 10 * coeff, 1.0),
                        vx=np.where(jet, .99 * coeff, 0.0))


def _st_maxwell1(desc, gamma):
    inside = _lower_octant(*_centres(desc))
    eps = desc.maxwell_permittivity
    shape = inside.shape
    zero = np.zeros(shape)
    comps = [zero, zero, np.full(shape, 1.0 * eps), zero, zero, np.where(inside, 1.0, -1.0)]
    return np.stack([c.reshape(-1) for c in comps], axis=1).reshape(-1).copy()


# ---------------------------------------------------------------- solid cells, potentials, self-gravity
# These entries also fill the potential and solid channels of buildStateEuler
# (res/util.lua:63-79); their state functions return (state, potential or
# None, solid or None).  crand() (res/util.lua:6, Lua's unseeded
# math.random) is drawn here from a generator seeded by the entry's name,
# so the noise is reproducible but not the draws of any reference run.

def _crand(name, shape, count):
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    return [rng.random(shape) * 2 - 1 for _ in range(count)]


def _su_cylinder(cfg):
    _set_bm(cfg, ("NONE", "FREEFLOW"), FREE, FREE)


def _st_cylinder(desc, gamma):
    x, y, z = _centres(desc)
    solid = (x * x + y * y + z * z < .25 * .25).astype(np.uint8)
    return _euler_state(desc, gamma, 1.0, pressure=1.0, vx=.2), None, solid.reshape(-1)


def _su_step(cfg):
    _set_bm(cfg, FREE, ("MIRROR", "MIRROR"), ("MIRROR", "MIRROR"))


def _st_step(desc, gamma):
    x, y, _ = _centres(desc)
    xmid = [.5 * (desc.xmin[a] + desc.xmax[a]) for a in range(3)]
    inside = (x > xmid[0]) & (y < xmid[1])
    return _euler_state(desc, gamma, 1.4, pressure=1.0, vx=3.0), None, inside.astype(np.uint8).reshape(-1)


def _su_mirror(cfg):
    _set_bm(cfg, ("MIRROR", "MIRROR"), ("MIRROR", "MIRROR"), ("MIRROR", "MIRROR"))


def _st_spiral(desc, gamma):
    # the entry's `useGravity = true` runs inside initState, after the app
    # has read useGravity (src/HydroGPUApp.cpp:207), so it changes nothing
    x, y, z = _centres(desc)
    s = np.where(np.sqrt(x * x + y * y + z * z) > .25, -1.0, 1.0)
    return _euler_state(desc, gamma, 1.1 - np.exp(-x * x - y * y), eint=1.0, vx=-y * s, vy=x * s)


def _su_rt(cfg):
    size = [int(v) for v in cfg["size"]]
    k = len(size) - 1  # the last axis of the size table (`#size`)
    xmin = list(cfg.get("xmin") or [-.5, -.5, -.5])
    xmax = list(cfg.get("xmax") or [.5, .5, .5])
    mid = .5 * (xmin[k] + xmax[k])
    xmin[k] = mid + (xmin[k] - mid) * 3
    xmax[k] = mid + (xmax[k] - mid) * 3
    size[k] *= 3
    cfg["xmin"], cfg["xmax"], cfg["size"] = xmin, xmax, size
    bm = [("PERIODIC", "PERIODIC")] * 3
    bm[k] = ("MIRROR", "MIRROR")
    _set_bm(cfg, *bm)


def _st_rt(desc, gamma):
    xs = _centres(desc)
    k = desc.dim - 1
    xmid = .5 * (desc.xmin[k] + desc.xmax[k])
    top = xs[k] > xmid
    force = (0.0, 1.0, 0.0)
    pot = np.zeros(xs[0].shape)
    for a in range(desc.dim):
        pot = pot + (xs[a] - desc.xmin[a]) * force[a]
    rho = np.where(top, 2.0, 1.0)
    noise = [.001 * c for c in _crand("Rayleigh-Taylor", xs[0].shape, 3)]
    st = _euler_state(desc, gamma, rho, pressure=2.5 - rho * pot, vx=noise[0], vy=noise[1], vz=noise[2])
    return st, pot.reshape(-1), None


def _su_selfgrav(cfg):
    _set_bm(cfg, FREE, FREE, FREE)
    cfg["useGravity"] = True


def _selfgrav_fields(desc, name, sources):
    """buildSelfGravitationState (res/util.lua:98-134) per cell: the first
    source (in order) that is nearer than the ones before and contains the
    cell; else the nearest.  Returns (index of the source, inside, offset
    from its centre, noise velocities)."""
    x, y, z = _centres(desc)
    best = np.full(x.shape, np.inf)
    which = np.zeros(x.shape, dtype=int)
    inside = np.zeros(x.shape, dtype=bool)
    done = np.zeros(x.shape, dtype=bool)
    for i, (c, r) in enumerate(sources):
        d2 = (c[0] - x) ** 2 + (c[1] - y) ** 2 + (c[2] - z) ** 2
        nearer = ~done & (d2 < best)
        best = np.where(nearer, d2, best)
        which = np.where(nearer, i, which)
        hit = nearer & (d2 < r * r)
        inside |= hit
        done |= hit
    cx = np.array([s[0][0] for s in sources])[which]
    cy = np.array([s[0][1] for s in sources])[which]
    cz = np.array([s[0][2] for s in sources])[which]
    dx, dy, dz = x - cx, y - cy, z - cz
    noise = np.exp(-100 * (dx * dx + dy * dy + dz * dz))
    vel = [.01 * noise * c for c in _crand(name, x.shape, 3)]
    return inside, (dx, dy, dz), vel


def _st_selfgrav(name, sources, spin=None):
    def st(desc, gamma):
        inside, (dx, dy, dz), vel = _selfgrav_fields(desc, name, sources)
        rho = np.where(inside, 1.0, .1)
        vx, vy, vz = vel
        if spin is not None:  # the source's own state inside it
            vx = np.where(inside, -spin * dy, vx)
            vy = np.where(inside, spin * dx, vy)
            vz = np.where(inside, 0.0, vz)
        return _euler_state(desc, gamma, rho, pressure=1.0, vx=vx, vy=vy, vz=vz), None, None
    return st


def _st_selfgrav2(desc, gamma):
    # two sources whose inside state is (rho 1, p 1, at rest); the entry then
    # replaces the momentum by (-5 rho y, 5 rho x) and keeps the energy
    # (its eight returned values leave the potential and solid channels nil)
    sources = [((-.25, 0, 0), .1), ((.25, 0, 0), .1)]
    inside, _, vel = _selfgrav_fields(desc, "self-gravitation test 2", sources)
    x, y, z = _centres(desc)
    rho = np.where(inside, 1.0, .1)
    vx, vy, vz = (np.where(inside, 0.0, v) for v in vel)
    st = _euler_state(desc, gamma, rho, pressure=1.0, vx=vx, vy=vy, vz=vz).reshape(-1, desc.dim + 2)
    st[:, 1] = (-5 * rho * y).reshape(-1)
    if desc.dim > 1:
        st[:, 2] = (5 * rho * x).reshape(-1)
    return st.reshape(-1).copy(), None, None


_EULER = ("Euler", "MHD", "SRHD")

# name -> (equations, setup(cfg) or None, state(desc, gamma))
CATALOG = {
    "Constant": (_EULER, None, _st_constant),
    "1D Advect Wave": (_EULER, None, _st_advect),
    "Sphere": (_EULER, None, _st_sphere),
    "Sod": (_EULER, _su_sod, _st_sod),
    **{"Configuration %d" % k: (_EULER, _su_config, _st_config(k)) for k in range(1, 7)},
    "Sedov": (_EULER, None, _st_sedov),
    "Brio-Wu": (_EULER, _su_briowu, _st_briowu),
    "Orszag-Tang": (_EULER, _su_orszag, _st_orszag),
    "Colella-Woodward": (_EULER, _su_free, _st_colella),
    "Kelvin-Helmholtz": (_EULER, _su_kh, _st_kh),
    "Double Mach Reflection": (_EULER, _su_dmr, _st_dmr),
    "Square Cavity": (_EULER, _su_cavity, _st_cavity),
    "Shock Bubble Interaction": (_EULER, _su_bubble, _st_bubble),
    "Relativistic Blast Wave Interaction": (_EULER, _su_gamma75, _st_rel_blast),
    "Marti & Muller 2008 Problem #1": (_EULER, _su_gamma53, _st_mm1),
    "Marti & Muller 2008 Problem #2": (_EULER, _su_gamma53, _st_mm2),
    "Relativistic Jet": (_EULER, _su_gamma53, _st_jet),
    "Maxwell-1": (("Maxwell",), None, _st_maxwell1),
    "Flow Around Cylinder": (_EULER, _su_cylinder, _st_cylinder),
    "Forward Facing Step": (_EULER, _su_step, _st_step),
    "Spiral Implosion": (_EULER, _su_mirror, _st_spiral),
    "Rayleigh-Taylor": (_EULER, _su_rt, _st_rt),
    "self-gravitation test 1": (_EULER, _su_selfgrav, _st_selfgrav("self-gravitation test 1", [((0, 0, 0), .2)])),
    "self-gravitation test 1 spinning": (_EULER, _su_selfgrav,
                                         _st_selfgrav("self-gravitation test 1 spinning", [((0, 0, 0), .2)],
                                                      spin=10.0)),
    "self-gravitation test 2": (_EULER, _su_selfgrav, _st_selfgrav2),
    "self-gravitation test 4": (_EULER, _su_selfgrav,
                                _st_selfgrav("self-gravitation test 4",
                                             [((.25, .25, 0), .1), ((-.25, .25, 0), .1), ((.25, -.25, 0), .1),
                                              ((-.25, -.25, 0), .1)])),
}

UNSUPPORTED = {
    "NR Gauge Shock Waves": "the ADM/BSSNOK equations",
    "NR Alcubierre Warp Bubble": "the ADM/BSSNOK equations",
    "NR Schwarzschild Black Hole": "the ADM/BSSNOK equations",
    "NR Stellar": "the ADM/BSSNOK equations",
}


def _entry(name):
    if name in CATALOG:
        return CATALOG[name]
    if name in UNSUPPORTED:
        raise ValueError("initial condition %r needs %s, which this build does not serve" % (name, UNSUPPORTED[name]))
    raise ValueError("unknown initial condition %r" % name)


def setup(name, cfg):
    """Apply entry `name`'s setup to a copy of the config dict `cfg` (keys of
    config.make_desc: size, xmin, xmax, boundaryMethods, cfl, defs, ...).
    Returns (cfg, state_fn) with state_fn(desc) -> AoS state."""
    eqs, su, st = _entry(name)
    out = dict(cfg)
    if "defs" in out:
        out["defs"] = dict(out["defs"])
    if su is not None:
        su(out)
    gamma = (out.get("defs") or {}).get("idealGas_heatCapacityRatio")
    return out, state_function(name, gamma)


def state_function(name, gamma=None):
    """state_fn(desc) -> AoS state of entry `name`.  `gamma` is the ratio of
    specific heats as the config holds it (the Lua value, e.g. 5/3); the
    kernels see it rounded to 14 significant digits (config.lua_number), and
    buildStateEuler uses the unrounded value, as here when given."""
    ch = channels_function(name, gamma)

    def fn(desc):
        return ch(desc)[0]
    return fn


def channels_function(name, gamma=None):
    """fn(desc) -> (state, potential or None, solid or None) of entry `name`:
    the state and the potential and solid channels that
    SelfGravitationBehavior::Converter reads (one value per cell)."""
    _, _, st = _entry(name)

    def fn(desc):
        out = st(desc, float(gamma) if gamma is not None else desc.gamma)
        return out if isinstance(out, tuple) else (out, None, None)
    return fn


def lua_setups():
    """name -> setup(globals, args) acting on a Lua config's globals
    (luaconfig.evaluate)."""
    from .luaconfig import LuaTable, from_py, to_py

    def wrap(name):
        def run(g, args):
            _, su, _ = _entry(name)
            if su is None:
                return
            cfg = {}
            for k in ("size", "xmin", "xmax", "boundaryMethods", "cfl", "defs"):
                v = g.get(k)
                if v is not None:
                    cfg[k] = to_py(v)
            if "size" in cfg and isinstance(cfg["size"], list):
                cfg["size"] = [int(s) for s in cfg["size"]]
            before = {k: cfg.get(k) for k in cfg}
            su(cfg)
            for k, v in cfg.items():
                if k == "defs":
                    defs = g.get("defs")
                    if not isinstance(defs, LuaTable):
                        defs = LuaTable()
                        g.set("defs", defs)
                    for dk, dv in v.items():
                        defs.set(dk, dv)
                elif k == "useGravity":
                    g.set("useGravity", bool(v))
                elif before.get(k) != v:
                    g.set(k, from_py([float(x) for x in v] if k in ("size", "xmin", "xmax") else v))
        return run

    return {name: wrap(name) for name in list(CATALOG) + list(UNSUPPORTED)}
