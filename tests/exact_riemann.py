"""Exact solution of the 1D Euler Riemann problem for an ideal gas.

Test aid only, written independently of the solver: the pressure function
of Toro, "Riemann Solvers and Numerical Methods for Fluid Dynamics" (3rd ed.,
ch. 4.2-4.5), solved by Newton iteration, and the self-similar sampling of
the wave pattern.  Used to pin the oracle's approximate Riemann solver on
weak waves and the Sod problem.
"""
import math


def _f(p, rho, pk, ck, g):
    """Toro (4.6)-(4.7): velocity change across a shock or rarefaction."""
    if p > pk:
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        return (p - pk) * math.sqrt(A / (p + B)), math.sqrt(A / (B + p)) * (1.0 - (p - pk) / (2.0 * (B + p)))
    r = p / pk
    return 2.0 * ck / (g - 1.0) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0), (1.0 / (rho * ck)) * r ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g=1.4, tol=1e-15):
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dl = _f(p, rl, pl, cl, g)
        fr, dr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dl + dr)
        pn = max(1e-14, p - dp)
        if abs(pn - p) <= tol * 0.5 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(xi, rl, ul, pl, rr, ur, pr, g=1.4):
    """State (rho, u, p) at x/t = xi."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    gm = (g - 1.0) / (g + 1.0)
    if xi <= us:  # left of the contact
        if ps > pl:  # left shock
            sl = ul - cl * math.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
            if xi <= sl:
                return rl, ul, pl
            return rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps
        shl = ul - cl
        if xi <= shl:
            return rl, ul, pl
        cs = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
        stl = us - cs
        if xi > stl:
            return rl * (ps / pl) ** (1.0 / g), us, ps
        c = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * (ul - xi))
        u = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * ul + xi)
        rho = rl * (c / cl) ** (2.0 / (g - 1.0))
        return rho, u, pl * (c / cl) ** (2.0 * g / (g - 1.0))
    if ps > pr:  # right shock
        sr = ur + cr * math.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
        if xi >= sr:
            return rr, ur, pr
        return rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps
    shr = ur + cr
    if xi >= shr:
        return rr, ur, pr
    cs = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
    str_ = us + cs
    if xi <= str_:
        return rr * (ps / pr) ** (1.0 / g), us, ps
    c = 2.0 / (g + 1.0) * (cr - (g - 1.0) / 2.0 * (ur - xi))
    u = 2.0 / (g + 1.0) * (-cr + (g - 1.0) / 2.0 * ur + xi)
    rho = rr * (c / cr) ** (2.0 / (g - 1.0))
    return rho, u, pr * (c / cr) ** (2.0 * g / (g - 1.0))


def godunov_flux(rl, ul, pl, rr, ur, pr, g=1.4, ut1=(0.0, 0.0), ut2=(0.0, 0.0)):
    """Exact Godunov flux (mass, normal momentum, transverse momenta,
    energy) at x/t = 0; transverse velocities are advected."""
    rho, u, p = sample(0.0, rl, ul, pl, rr, ur, pr, g)
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    t1 = ut1[0] if us >= 0 else ut1[1]
    t2 = ut2[0] if us >= 0 else ut2[1]
    E = p / (g - 1.0) + 0.5 * rho * (u * u + t1 * t1 + t2 * t2)
    return [rho * u, rho * u * u + p, rho * u * t1, rho * u * t2, u * (E + p)]
