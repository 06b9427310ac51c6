"""Exact solution of the 1D Euler Riemann problem for an ideal gas.

Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics", ch. 4:
Newton iteration for p* (eqs. 4.5-4.9), then sampling (section 4.5).
Used to pin the oracle and the HIP path against the mathematics (Sod tube),
and by the batch IC generator to keep its random Riemann problems inside the
regime the Roe scheme survives.
"""
import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        f = (p - pk) * np.sqrt(A / (p + B))
        df = np.sqrt(A / (B + p)) * (1.0 - 0.5 * (p - pk) / (B + p))
    else:  # rarefaction
        f = 2.0 * ck / (g - 1.0) * ((p / pk) ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rho * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(left, right, g):
    rl, ul, pl = left
    rr, ur, pr = right
    cl = np.sqrt(g * pl / rl)
    cr = np.sqrt(g * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dfl = _f(p, rl, pl, cl, g)
        fr, dfr = _f(p, rr, pr, cr, g)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-12)
        if abs(pn - p) / (0.5 * (pn + p)) < 1e-14:
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(left, right, g, xi):
    """(rho, u, p) at similarity coordinates xi = (x - x0) / t."""
    rl, ul, pl = left
    rr, ur, pr = right
    cl = np.sqrt(g * pl / rl)
    cr = np.sqrt(g * pr / rr)
    ps, us = star_state(left, right, g)
    xi = np.atleast_1d(np.asarray(xi, dtype=np.float64))
    rho = np.empty_like(xi)
    u = np.empty_like(xi)
    p = np.empty_like(xi)
    gm = (g - 1.0) / (g + 1.0)
    for n, s in enumerate(xi):
        if s <= us:  # left of contact
            if ps > pl:  # left shock
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s <= sl:
                    rho[n], u[n], p[n] = rl, ul, pl
                else:
                    rho[n] = rl * (ps / pl + gm) / (gm * ps / pl + 1.0)
                    u[n], p[n] = us, ps
            else:  # left rarefaction
                shl = ul - cl
                cml = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - cml
                if s <= shl:
                    rho[n], u[n], p[n] = rl, ul, pl
                elif s >= stl:
                    rho[n] = rl * (ps / pl) ** (1.0 / g)
                    u[n], p[n] = us, ps
                else:
                    c = 2.0 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    u[n] = 2.0 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    rho[n] = rl * (c / cl) ** (2.0 / (g - 1))
                    p[n] = pl * (c / cl) ** (2.0 * g / (g - 1))
        else:
            if ps > pr:  # right shock
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s >= sr:
                    rho[n], u[n], p[n] = rr, ur, pr
                else:
                    rho[n] = rr * (ps / pr + gm) / (gm * ps / pr + 1.0)
                    u[n], p[n] = us, ps
            else:
                shr = ur + cr
                cmr = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + cmr
                if s >= shr:
                    rho[n], u[n], p[n] = rr, ur, pr
                elif s <= str_:
                    rho[n] = rr * (ps / pr) ** (1.0 / g)
                    u[n], p[n] = us, ps
                else:
                    c = 2.0 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    u[n] = 2.0 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    rho[n] = rr * (c / cr) ** (2.0 / (g - 1))
                    p[n] = pr * (c / cr) ** (2.0 * g / (g - 1))
    return rho, u, p
