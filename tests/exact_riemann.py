"""Exact solution of the 1-D Euler Riemann problem (Toro, "Riemann Solvers and
Numerical Methods for Fluid Dynamics", 3rd ed., ch. 4): Newton iteration for
the star pressure, then sampling of the self-similar solution."""
import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        a = 2.0 / ((g + 1.0) * rho)
        b = (g - 1.0) / (g + 1.0) * pk
        s = np.sqrt(a / (p + b))
        return (p - pk) * s, s * (1.0 - 0.5 * (p - pk) / (b + p))
    pr = p / pk  # rarefaction
    return (2.0 * ck / (g - 1.0)) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0), (1.0 / (rho * ck)) * pr ** (-(g + 1.0) / (2.0 * g))


def star_state(rl, ul, pl, rr, ur, pr, g=1.4):
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dfl = _f(p, rl, pl, cl, g)
        fr, dfr = _f(p, rr, pr, cr, g)
        dp = (fl + fr + (ur - ul)) / (dfl + dfr)
        p_new = max(1e-14, p - dp)
        if abs(p_new - p) < 1e-15 * p:
            p = p_new
            break
        p = p_new
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, rl, ul, pl, rr, ur, pr, g=1.4):
    """(rho, u, p) at positions x and time t for a discontinuity at x0."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl, cr = np.sqrt(g * pl / rl), np.sqrt(g * pr / rr)
    out = np.zeros((3, len(x)))
    for k, xi in enumerate(x):
        s = (xi - x0) / t
        if s <= us:  # left of contact
            if ps > pl:
                sl = ul - cl * np.sqrt((g + 1) / (2 * g) * ps / pl + (g - 1) / (2 * g))
                if s < sl:
                    w = (rl, ul, pl)
                else:
                    r = rl * (ps / pl + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pl + 1)
                    w = (r, us, ps)
            else:
                shl = ul - cl
                cs = cl * (ps / pl) ** ((g - 1) / (2 * g))
                stl = us - cs
                if s < shl:
                    w = (rl, ul, pl)
                elif s > stl:
                    w = (rl * (ps / pl) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (cl + (g - 1) / 2 * ul + s)
                    c = 2 / (g + 1) * (cl + (g - 1) / 2 * (ul - s))
                    r = rl * (c / cl) ** (2 / (g - 1))
                    w = (r, u, pl * (c / cl) ** (2 * g / (g - 1)))
        else:
            if ps > pr:
                sr = ur + cr * np.sqrt((g + 1) / (2 * g) * ps / pr + (g - 1) / (2 * g))
                if s > sr:
                    w = (rr, ur, pr)
                else:
                    r = rr * (ps / pr + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pr + 1)
                    w = (r, us, ps)
            else:
                shr = ur + cr
                cs = cr * (ps / pr) ** ((g - 1) / (2 * g))
                str_ = us + cs
                if s > shr:
                    w = (rr, ur, pr)
                elif s < str_:
                    w = (rr * (ps / pr) ** (1 / g), us, ps)
                else:
                    u = 2 / (g + 1) * (-cr + (g - 1) / 2 * ur + s)
                    c = 2 / (g + 1) * (cr - (g - 1) / 2 * (ur - s))
                    r = rr * (c / cr) ** (2 / (g - 1))
                    w = (r, u, pr * (c / cr) ** (2 * g / (g - 1)))
        out[:, k] = w
    return out
