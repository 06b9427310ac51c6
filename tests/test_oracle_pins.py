"""Closed-form pins of the oracle's wave speeds, timestep and limited flux.

Every expected value here comes from a textbook formula evaluated on
primitive variables (rho, u, p), a route independent of the oracle's
conserved-variable arithmetic, or from a hand-computed number:

  * slope limiters: values at chosen r worked out by hand from the formulas
    of res/SlopeLimiter.cl:5-68 (Waterson & Deconinck 2007 list the textbook
    forms; the reference's Ospre carries .5 where the literature has 1.5);
  * HLL: Davis wave-speed estimates bounded by the Roe-averaged speeds
    (Toro, Riemann Solvers, 3rd ed., sec. 10.5.1, eqs. 10.48 and 10.52),
    res/EulerHLL.cl:100-103;
  * HLLC: the PVRS pressure estimate and the pressure-based wave speeds
    (Toro eqs. 10.67-10.69), res/EulerHLLC.cl:87-94, on shock, rarefaction,
    mixed and vacuum-clamped states;
  * Roe calcCellTimestep (res/Roe.cl:4-65): on a uniform state every face
    carries u-c, u, u+c, so dt = cfl dx / (2c + 1e-9) subsonic and
    cfl dx / (|u| + c + 1e-9) supersonic; a non-uniform line separates the
    left-face / right-face pairing (max speed from the left face, min speed
    from the right face);
  * HLL calcCellTimestep (res/EulerHLL.cl:142-175): the same without fabs
    and without 1e-9;
  * the r-dependent limited Roe flux (res/Roe.cl:272-298): a density wave
    carried by uniform u and p is a pure contact wave, so one forward-Euler
    step is the Sweby flux-limited scheme for linear advection,
    F = u rho_up + |u|/2 (1 - |nu|) phi(r) (rho_R - rho_L), r the upwind
    jump over this jump (Sweby 1984; LeVeque, Finite Volume Methods, 6.40).
"""
import numpy as np
import pytest

import oracle
from hydrodynamicsgpu_amd import config, ics

LIM = {n: i for i, n in enumerate(config.LIMITERS)}
GAMMA = 1.4
ROE, HLL, HLLC = 0, 1, 2


def cons(rho, v, p, gamma=GAMMA):
    v = np.atleast_1d(np.asarray(v, dtype=float))
    E = p / (gamma - 1) + 0.5 * rho * float(v @ v)
    return np.concatenate([[rho], rho * v, [E]])


# ---------------------------------------------------------------------------
# slope limiters: hand-computed points


LIMITER_POINTS = {
    # phi(r) = (1 + r)/2
    "Fromm": [(0.0, 0.5), (1.0, 1.0), (3.0, 2.0), (-1.0, 0.0), (0.5, 0.75)],
    # phi(r) = max(0, r) (3r + 1)/(r + 1)^2
    "CHARM": [(1.0, 1.0), (2.0, 14.0 / 9.0), (0.5, 5.0 / 9.0), (-2.0, 0.0), (3.0, 30.0 / 16.0)],
    # phi(r) = max(0, 1.5 (r + |r|)/(r + 2))
    "HCUS": [(1.0, 1.0), (2.0, 1.5), (0.5, 0.6), (-1.0, 0.0), (4.0, 2.0)],
    # phi(r) = max(0, 2 (r + |r|)/(r + 3))
    "HQUICK": [(1.0, 1.0), (3.0, 2.0), (0.5, 4.0 / 7.0), (-2.0, 0.0), (5.0, 2.5)],
    # phi(r) = max(0, min(2r, (1 + 2r)/3, 2))
    "Koren": [(0.1, 0.2), (1.0, 1.0), (2.0, 5.0 / 3.0), (4.0, 2.0), (-1.0, 0.0), (0.25, 0.5)],
    # phi(r) = max(0, min(r, 1.5))   (Osher with beta = 1.5)
    "Oshker": [(0.5, 0.5), (1.0, 1.0), (2.0, 1.5), (-1.0, 0.0), (1.25, 1.25)],
    # phi(r) = .5 (r^2 + r)/(r^2 + r + 1)   (the reference's coefficient)
    "Ospre": [(1.0, 1.0 / 3.0), (2.0, 3.0 / 7.0), (0.0, 0.0), (-1.0, 0.0), (-2.0, 1.0 / 3.0)],
    # phi(r) = max(0, min(2r, .25 + .75 r, 4))
    "Smart": [(0.1, 0.2), (1.0, 1.0), (3.0, 2.5), (10.0, 4.0), (-1.0, 0.0)],
    # phi(r) = max(0, min(1.5 r, 1), min(r, 1.5))   (Sweby with beta = 1.5)
    "Sweby": [(0.5, 0.75), (1.0, 1.0), (1.2, 1.2), (3.0, 1.5), (-1.0, 0.0), (0.25, 0.375)],
    # phi(r) = max(0, min(2r, .75 + .25 r, .25 + .75 r, 2))
    "UMIST": [(0.1, 0.2), (1.0, 1.0), (3.0, 1.5), (10.0, 2.0), (-1.0, 0.0), (0.5, 0.625)],
    # phi(r) = 2r/(r^2 + 1)
    "VanAlbada2": [(1.0, 1.0), (2.0, 0.8), (0.5, 0.8), (0.0, 0.0), (-1.0, -1.0)],
    # phi(r) = (r + 1)/2 min(1, 4r/(r + 1), 4/(r + 1))
    "BarthJespersen": [(1.0, 1.0), (0.2, 0.4), (5.0, 2.0), (0.0, 0.0), (3.0, 2.0)],
    # phi(r) = (r^2 + r)/(r^2 + 1)
    "VanAlbada1": [(1.0, 1.0), (2.0, 1.2), (0.0, 0.0), (-1.0, 0.0), (3.0, 1.2)],
}


@pytest.mark.parametrize("name", sorted(LIMITER_POINTS))
def test_limiter_hand_computed_points(name):
    for r, expect in LIMITER_POINTS[name]:
        got = oracle.slope_limiter(LIM[name], r)
        assert got == pytest.approx(expect, rel=1e-15, abs=1e-300), (name, r, got, expect)


def test_every_limiter_is_pinned_at_closed_form_points():
    # the limiters pinned elsewhere (test_oracle_kernels.py) plus the table here
    # cover every entry of res/SlopeLimiter.cl
    elsewhere = {"DonorCell", "LaxWendroff", "BeamWarming", "MinMod", "MonotizedCentral", "Superbee", "VanLeer"}
    assert set(config.LIMITERS) == elsewhere | set(LIMITER_POINTS)


# ---------------------------------------------------------------------------
# HLL / HLLC wave speeds


def _prims(rho, u, p, gamma=GAMMA):
    a = np.sqrt(gamma * p / rho)
    H = a * a / (gamma - 1) + 0.5 * float(np.dot(u, u))
    return a, H


def roe_speeds(L, R, side, gamma=GAMMA):
    """Roe-averaged normal velocity and sound speed from primitives
    (Toro eqs. 11.60): weights sqrt(rho), H = a^2/(gamma-1) + |u|^2/2."""
    (rl, ul, pl), (rr, ur, pr) = L, R
    ul, ur = np.atleast_1d(ul).astype(float), np.atleast_1d(ur).astype(float)
    al, hl = _prims(rl, ul, pl, gamma)
    ar, hr = _prims(rr, ur, pr, gamma)
    wl, wr = np.sqrt(rl), np.sqrt(rr)
    ut = (wl * ul + wr * ur) / (wl + wr)
    Ht = (wl * hl + wr * hr) / (wl + wr)
    at = np.sqrt((gamma - 1) * (Ht - 0.5 * float(ut @ ut)))
    return ut[side], at, ul[side] - al, ur[side] + ar, al, ar


def davis_bounded(L, R, side):
    """Toro eq. 10.52 with the Roe speeds (10.48): SL = min(uL - aL, u~ - a~),
    SR = max(uR + aR, u~ + a~)."""
    ut, at, minl, maxr, _, _ = roe_speeds(L, R, side)
    return min(minl, ut - at), max(maxr, ut + at), ut, (minl, ut - at, maxr, ut + at)


def pvrs(L, R, side, gamma=GAMMA):
    """Toro eqs. 10.67-10.69: p* = max(0, (pL+pR)/2 - (uR-uL) rho_bar a_bar / 2),
    q_K = 1 if p* <= p_K else sqrt(1 + (gamma+1)/(2 gamma) (p*/p_K - 1)),
    SL = uL - aL qL, SR = uR + aR qR."""
    (rl, ul, pl), (rr, ur, pr) = L, R
    uln, urn = np.atleast_1d(ul)[side], np.atleast_1d(ur)[side]
    al, ar = np.sqrt(gamma * pl / rl), np.sqrt(gamma * pr / rr)
    ps = max(0.0, 0.5 * (pl + pr) - 0.5 * (urn - uln) * 0.5 * (rl + rr) * 0.5 * (al + ar))
    ql = 1.0 if ps <= pl else np.sqrt(1 + (gamma + 1) / (2 * gamma) * (ps / pl - 1))
    qr = 1.0 if ps <= pr else np.sqrt(1 + (gamma + 1) / (2 * gamma) * (ps / pr - 1))
    return uln - al * ql, urn + ar * qr, ps, ql, qr


# (left, right) primitive pairs.  Each Davis branch is taken by some pair:
# the left-state bound below the Roe bound and the reverse, likewise on the
# right (asserted below so the pins stay discriminating).
DAVIS_CASES = [
    ((1.0, [0.0], 1.0), (4.0, [0.0], 1.0)),       # light | heavy
    ((4.0, [0.0], 1.0), (1.0, [0.0], 1.0)),       # heavy | light
    ((1.0, [0.75], 1.0), (0.125, [-0.3], 0.1)),    # colliding
    ((0.5, [-0.6], 0.4), (2.0, [0.9], 3.0)),       # receding, unequal
]


def test_hll_davis_bounded_speeds_match_toro():
    branches = set()
    for L, R in DAVIS_CASES:
        sl, sr, ut, (minl, minroe, maxr, maxroe) = davis_bounded(L, R, 0)
        branches.add(("L" if minl < minroe else "roe", "R" if maxr > maxroe else "roe"))
        ev = oracle.hll_speeds(HLL, 1, 0, GAMMA, cons(*L), cons(*R))
        np.testing.assert_allclose(ev, [sl, ut, sr], rtol=1e-14, atol=1e-15)
    # both arms of both bounds occur, so dropping either side of min/max fails
    assert {b[0] for b in branches} == {"L", "roe"}
    assert {b[1] for b in branches} == {"R", "roe"}


@pytest.mark.parametrize("dim", [2, 3])
def test_hll_davis_speeds_rotate_with_the_side(dim):
    # transverse velocities enter the Roe-averaged |u|^2 (hence a~) but the
    # speeds are taken along the face normal
    rng = np.random.default_rng(11)
    for _ in range(6):
        L = (float(rng.uniform(0.2, 3)), rng.uniform(-1, 1, dim), float(rng.uniform(0.1, 4)))
        R = (float(rng.uniform(0.2, 3)), rng.uniform(-1, 1, dim), float(rng.uniform(0.1, 4)))
        for side in range(dim):
            sl, sr, ut, _ = davis_bounded(L, R, side)
            ev = oracle.hll_speeds(HLL, dim, side, GAMMA, cons(*L), cons(*R))
            np.testing.assert_allclose(ev, [sl] + [ut] * dim + [sr], rtol=1e-13, atol=1e-14)


def test_hll_davis_hand_computed():
    # rho 1 | 4 at rest, p = 1 both sides: aL = sqrt(1.4), aR = sqrt(.35);
    # Roe weights 1, 2 -> u~ = 0, H~ = (1*3.5 + 2*.875)/3 = 1.75,
    # a~ = sqrt(.4 * 1.75) = sqrt(.7).  SL = min(-sqrt(1.4), -sqrt(.7)),
    # SR = max(sqrt(.35), sqrt(.7)).
    ev = oracle.hll_speeds(HLL, 1, 0, GAMMA, cons(1.0, [0.0], 1.0), cons(4.0, [0.0], 1.0))
    np.testing.assert_allclose(ev, [-np.sqrt(1.4), 0.0, np.sqrt(0.7)], rtol=1e-15, atol=1e-16)


PVRS_CASES = [
    ((1.0, [1.0], 1.0), (1.0, [-1.0], 1.0)),        # two shocks
    ((1.0, [-0.3], 1.0), (1.0, [0.3], 1.0)),        # two rarefactions
    ((1.0, [0.0], 1.0), (0.125, [0.0], 0.1)),        # Sod: rarefaction | shock
    ((1.0, [-3.0], 0.4), (1.0, [3.0], 0.4)),        # p_pvrs < 0: clamped to 0
    ((0.3, [0.4], 2.0), (2.5, [-0.2], 0.05)),        # shock into the right state only
    ((2.5, [0.2], 0.05), (0.3, [-0.4], 2.0)),        # shock into the left state only
]


def test_hllc_pvrs_speeds_match_toro():
    kinds = set()
    for L, R in PVRS_CASES:
        sl, sr, ps, ql, qr = pvrs(L, R, 0)
        kinds.add((ql > 1, qr > 1, ps == 0.0))
        ut = roe_speeds(L, R, 0)[0]
        ev = oracle.hll_speeds(HLLC, 1, 0, GAMMA, cons(*L), cons(*R))
        np.testing.assert_allclose(ev, [sl, ut, sr], rtol=1e-14, atol=1e-15)
    # shock and rarefaction branches on both sides, and the p* = 0 clamp
    assert (True, True, False) in kinds and (False, False, False) in kinds
    assert (False, True, False) in kinds and (True, False, False) in kinds
    assert (False, False, True) in kinds


def test_hllc_pvrs_hand_computed():
    # rho = p = 1 both sides, u = +-1, gamma 1.4: a = sqrt(1.4),
    # p* = 1 + sqrt(1.4), q = sqrt(1 + (2.4/2.8) sqrt(1.4)),
    # SL = 1 - sqrt(1.4) q, SR = -1 + sqrt(1.4) q
    a = np.sqrt(1.4)
    q = np.sqrt(1 + (2.4 / 2.8) * a)
    ev = oracle.hll_speeds(HLLC, 1, 0, GAMMA, cons(1.0, [1.0], 1.0), cons(1.0, [-1.0], 1.0))
    np.testing.assert_allclose(ev, [1 - a * q, 0.0, -1 + a * q], rtol=1e-14, atol=1e-15)
    # the shock factor uses (gamma + 1)/(2 gamma): with (gamma - 1) the
    # speeds would move by more than 10 %
    q_wrong = np.sqrt(1 + (0.4 / 2.8) * a)
    assert abs((1 - a * q_wrong) - ev[0]) > 0.1 * abs(ev[0])


@pytest.mark.parametrize("dim", [2, 3])
def test_hllc_pvrs_speeds_rotate_with_the_side(dim):
    rng = np.random.default_rng(5)
    for _ in range(6):
        L = (float(rng.uniform(0.2, 3)), rng.uniform(-1, 1, dim), float(rng.uniform(0.1, 4)))
        R = (float(rng.uniform(0.2, 3)), rng.uniform(-1, 1, dim), float(rng.uniform(0.1, 4)))
        for side in range(dim):
            sl, sr, _, _, _ = pvrs(L, R, side)
            ut = roe_speeds(L, R, side)[0]
            ev = oracle.hll_speeds(HLLC, dim, side, GAMMA, cons(*L), cons(*R))
            np.testing.assert_allclose(ev, [sl] + [ut] * dim + [sr], rtol=1e-13, atol=1e-14)


# ---------------------------------------------------------------------------
# calcCellTimestep + findMinTimestep


def _uniform_sim(solverName, size, rho, v, p, cfl=0.5, xmax=None):
    dim = len(size)
    xmax = xmax or [float(s) * 0.5 for s in size]   # dx = 0.5 exactly on every axis
    d = config.make_desc(size, xmin=[0.0] * dim, xmax=xmax, solverName=solverName, cfl=cfl,
                         flags=config.FLAG_REFERENCE)
    shape = (1,) * (3 - dim) + tuple(reversed(size))
    f = lambda x: np.full(shape, float(x))  # noqa: E731
    vv = list(v) + [0.0] * (3 - len(v))
    q = ics.build_state_euler(d, f(rho), f(vv[0]), f(vv[1]), f(vv[2]), f(p))
    s = oracle.OracleSim(d)
    s.set_state(q)
    s.boundary()
    s.init_step()   # HLL freezes its wave speeds here (src/Solver/HLL.cpp:31-33)
    return s, d


@pytest.mark.parametrize("u", [0.3, -0.3, 2.5, -2.5])
def test_roe_cfl_uniform_state_1d(u):
    # rho = 1, p = 1: c = sqrt(1.4) ~ 1.183; |u| = 2.5 is supersonic
    c = np.sqrt(1.4)
    s, _ = _uniform_sim("EulerRoe", [24], 1.0, [u], 1.0)
    dt = s.calc_timestep()
    speed = 2 * c if abs(u) < c else abs(u) + c
    assert dt == pytest.approx(0.5 * 0.5 / (speed + 1e-9), rel=1e-14)


@pytest.mark.parametrize("u", [0.3, -0.3, 2.5, -2.5])
def test_hll_cfl_uniform_state_1d(u):
    c = np.sqrt(1.4)
    for name in ("EulerHLL", "EulerHLLC"):
        s, _ = _uniform_sim(name, [24], 1.0, [u], 1.0)
        dt = s.calc_timestep()
        speed = 2 * c if abs(u) < c else abs(u) + c
        # no 1e-9 in the HLL form: the relative difference would be ~4e-10
        assert dt == pytest.approx(0.5 * 0.5 / speed, rel=1e-14), name


def test_cfl_takes_the_smallest_side_2d():
    # dx = .5, dy = .25; v = (.2, -.6): side x gives .5/(2c), side y .25/(2c)
    c = np.sqrt(1.4)
    for name, eps in (("EulerRoe", 1e-9), ("EulerHLL", 0.0), ("EulerHLLC", 0.0)):
        s, _ = _uniform_sim(name, [12, 16], 1.0, [0.2, -0.6], 1.0, xmax=[6.0, 4.0])
        assert s.calc_timestep() == pytest.approx(0.5 * 0.25 / (2 * c + eps), rel=1e-14), name
    # supersonic in y only: .25/(|vy| + c) against .5/(2c)
    for name, eps in (("EulerRoe", 1e-9), ("EulerHLL", 0.0)):
        s, _ = _uniform_sim(name, [12, 16], 1.0, [0.2, -2.0], 1.0, xmax=[6.0, 4.0])
        assert s.calc_timestep() == pytest.approx(0.5 * 0.25 / (2.0 + c + eps), rel=1e-14), name


def _cfl_line(pairing, solver, rho, u, p, dx, cfl):
    """Global CFL dt of a 1D line by res/Roe.cl:4-65 (Roe, +1e-9, fabs) or
    res/EulerHLL.cl:142-175 (HLL speeds), with the max speed taken from
    the left face of each cell ("ref") or the right face ("swapped").
    Face i lies between cells i-1 and i; cells 2..S-3 are active."""
    S = rho.size
    lo, hi = {}, {}
    for i in range(2, S - 1):
        L, R = (rho[i - 1], [u[i - 1]], p[i - 1]), (rho[i], [u[i]], p[i])
        if solver == "roe":
            ut, at = roe_speeds(L, R, 0)[:2]
            lo[i], hi[i] = ut - at, ut + at
        else:
            lo[i], hi[i] = davis_bounded(L, R, 0)[:2]
    best = np.inf
    for c in range(2, S - 2):
        fmax, fmin = (c, c + 1) if pairing == "ref" else (c + 1, c)
        mx, mn = max(0.0, hi[fmax]), min(0.0, lo[fmin])
        best = min(best, dx / (abs(mx - mn) + 1e-9) if solver == "roe" else dx / (mx - mn))
    return cfl * best


@pytest.mark.parametrize("name,solver", [("EulerRoe", "roe"), ("EulerHLL", "hll")])
def test_cfl_face_pairing(name, solver):
    # a line whose speeds vary strongly from face to face: the min over cells
    # of dx/(max(0, lambda_max(left face)) - min(0, lambda_min(right face)))
    n = 28
    x = np.arange(n)
    rho = 1.0 + 0.6 * (x % 3 == 0) + 0.1 * np.sin(x)
    u = np.where(x < n // 2, 1.1, -0.9) + 0.4 * np.cos(1.7 * x)
    p = 0.5 + 0.3 * (x % 4 == 1)
    d = config.make_desc([n], xmin=[0.0], xmax=[n * 0.5], solverName=name, cfl=0.5, flags=config.FLAG_REFERENCE)
    q = ics.build_state_euler(d, rho.reshape(1, 1, -1), u.reshape(1, 1, -1), None, None, p.reshape(1, 1, -1))
    s = oracle.OracleSim(d)
    s.set_state(q)
    s.init_step()
    dt = s.calc_timestep()
    ref = _cfl_line("ref", solver, rho, u, p, 0.5, 0.5)
    swapped = _cfl_line("swapped", solver, rho, u, p, 0.5, 0.5)
    assert abs(ref - swapped) > 1e-3 * ref   # the line tells the pairings apart
    assert dt == pytest.approx(ref, rel=1e-13)


# ---------------------------------------------------------------------------
# the r-dependent limited Roe flux: Sweby's scheme on a contact wave


def sweby_step(rho, u, nu, limiter, dtdx):
    """One forward-Euler step of the flux-limited scheme on the whole line
    (ghosts included).  Face i (between cells i-1 and i) is computed for
    i = 2..S-2; the jumps at faces 1 and S-1 are never written by the
    reference and read as zero, so the limiter sees r = 0 there."""
    S = rho.size
    jump = np.zeros(S + 1)
    for i in range(2, S - 1):
        jump[i] = rho[i] - rho[i - 1]
    F = np.zeros(S + 1)
    for i in range(2, S - 1):
        up = jump[i - 1] if u >= 0 else jump[i + 1]
        r = up / jump[i]
        phi = limiter(r)
        F[i] = u * (rho[i - 1] if u >= 0 else rho[i]) + 0.5 * abs(u) * (1 - abs(nu)) * phi * jump[i]
    out = rho.copy()
    for c in range(2, S - 2):
        out[c] = rho[c] - dtdx * (F[c + 1] - F[c])
    return out


SWEBY_LIMITERS = {
    "MonotizedCentral": lambda r: max(0.0, min(2.0, 0.5 * (1 + r), 2 * r)),
    "Superbee": lambda r: max(0.0, min(1.0, 2 * r), min(2.0, r)),
    "MinMod": lambda r: max(0.0, min(r, 1.0)),
}


@pytest.mark.parametrize("limiter", sorted(SWEBY_LIMITERS))
@pytest.mark.parametrize("u", [0.6, -0.6])
def test_limited_flux_is_sweby_on_a_contact_wave(limiter, u):
    n = 64
    S = n + 4
    d = config.make_desc([S], xmin=[0.0], xmax=[S * 0.5], integratorName="ForwardEuler", slopeLimiter=limiter,
                         boundaryMethods=[{"min": "PERIODIC", "max": "PERIODIC"}], cfl=0.8,
                         flags=config.FLAG_REFERENCE)
    x = (np.arange(n) + 0.5) / n
    interior = 1.0 + 0.2 * np.sin(2 * np.pi * x) + 0.3 * ((x > 0.3) & (x < 0.6)) + 0.01 * x
    rho = np.concatenate([interior[-2:], interior, interior[:2]])
    p = np.ones(S)
    q0 = ics.build_state_euler(d, rho.reshape(1, 1, -1), np.full((1, 1, S), u), None, None, p.reshape(1, 1, -1))
    s = oracle.OracleSim(d)
    s.set_state(q0)
    dt = s.update()                  # boundary, Roe basis, CFL, one forward-Euler step
    dx = 0.5
    nu = u * dt / dx
    assert 0.1 < abs(nu) < 1
    got = s.get_state().reshape(S, 3)[:, 0]
    expect = sweby_step(rho, u, nu, SWEBY_LIMITERS[limiter], dt / dx)
    np.testing.assert_allclose(got[2:S - 2], expect[2:S - 2], rtol=0, atol=1e-12)
    # the pin separates the upwind ratio from the downwind one by eight
    # orders of magnitude more than its tolerance
    swapped = _sweby_swapped(rho, u, nu, SWEBY_LIMITERS[limiter], dt / dx)
    assert np.max(np.abs(swapped[2:S - 2] - expect[2:S - 2])) > 1e-4


def _sweby_swapped(rho, u, nu, limiter, dtdx):
    """The same scheme with r taken from the downwind jump (a plausible slip)."""
    S = rho.size
    jump = np.zeros(S + 1)
    for i in range(2, S - 1):
        jump[i] = rho[i] - rho[i - 1]
    F = np.zeros(S + 1)
    for i in range(2, S - 1):
        dn = jump[i + 1] if u >= 0 else jump[i - 1]
        phi = limiter(dn / jump[i])
        F[i] = u * (rho[i - 1] if u >= 0 else rho[i]) + 0.5 * abs(u) * (1 - abs(nu)) * phi * jump[i]
    out = rho.copy()
    for c in range(2, S - 2):
        out[c] = rho[c] - dtdx * (F[c + 1] - F[c])
    return out
