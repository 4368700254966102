"""Exact solution of the 1-D Riemann problem for the ideal-gas Euler
equations (E. F. Toro, Riemann Solvers and Numerical Methods for Fluid
Dynamics, 3rd ed., ch. 4: pressure function 4.5-4.8, Newton iteration 4.44
from the two-rarefaction guess 4.46, sampling 4.52-4.63).  Test
infrastructure: the physical reference the solvers converge to."""
import math

import numpy as np


def _f(p, rho, pk, a, g):
    """Toro (4.6)/(4.7): f_K(p) and its derivative."""
    if p > pk:                                   # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + B))
    r = p / pk                                   # rarefaction
    return 2.0 * a / (g - 1.0) * (r ** ((g - 1.0) / (2.0 * g)) - 1.0), (1.0 / (rho * a)) * r ** (-(g + 1.0) / (2.0 * g))


def star(wl, wr, g=1.4, tol=1e-14):
    """(p*, u*) of left/right primitive states (rho, u, p)."""
    rl, ul, pl = wl
    rr, ur, pr = wr
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    if 2.0 * (al + ar) / (g - 1.0) <= ur - ul:
        raise ValueError("vacuum is generated")
    z = (g - 1.0) / (2.0 * g)
    p = ((al + ar - 0.5 * (g - 1.0) * (ur - ul)) / (al / pl ** z + ar / pr ** z)) ** (1.0 / z)   # (4.46)
    for _ in range(100):
        fl, dl = _f(p, rl, pl, al, g)
        fr, dr = _f(p, rr, pr, ar, g)
        pn = max(p - (fl + fr + ur - ul) / (dl + dr), 1e-14)
        done = 2.0 * abs(pn - p) / (pn + p) < tol
        p = pn
        if done:
            break
    fl, _ = _f(p, rl, pl, al, g)
    fr, _ = _f(p, rr, pr, ar, g)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(wl, wr, xs, t, x0=0.0, g=1.4):
    """Primitive (rho, u, p) at positions xs and time t > 0."""
    rl, ul, pl = wl
    rr, ur, pr = wr
    ps, us = star(wl, wr, g)
    al, ar = math.sqrt(g * pl / rl), math.sqrt(g * pr / rr)
    gm, gp = (g - 1.0) / (g + 1.0), (g + 1.0) / (2.0 * g)
    out = np.empty((len(xs), 3))
    for i, x in enumerate(xs):
        s = (x - x0) / t
        if s <= us:                              # left of the contact
            if ps > pl:                          # left shock
                sl = ul - al * math.sqrt(gp * ps / pl + (g - 1.0) / (2.0 * g))
                out[i] = (rl, ul, pl) if s <= sl else (rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps)
            else:                                # left rarefaction
                shl, stl = ul - al, us - al * (ps / pl) ** ((g - 1.0) / (2.0 * g))
                if s <= shl:
                    out[i] = (rl, ul, pl)
                elif s >= stl:
                    out[i] = (rl * (ps / pl) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) + gm / al * (ul - s)
                    out[i] = (rl * c ** (2.0 / (g - 1.0)), 2.0 / (g + 1.0) * (al + 0.5 * (g - 1.0) * ul + s),
                              pl * c ** (2.0 * g / (g - 1.0)))
        else:                                    # right of the contact
            if ps > pr:                          # right shock
                sr = ur + ar * math.sqrt(gp * ps / pr + (g - 1.0) / (2.0 * g))
                out[i] = (rr, ur, pr) if s >= sr else (rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps)
            else:                                # right rarefaction
                shr, stt = ur + ar, us + ar * (ps / pr) ** ((g - 1.0) / (2.0 * g))
                if s >= shr:
                    out[i] = (rr, ur, pr)
                elif s <= stt:
                    out[i] = (rr * (ps / pr) ** (1.0 / g), us, ps)
                else:
                    c = 2.0 / (g + 1.0) - gm / ar * (ur - s)
                    out[i] = (rr * c ** (2.0 / (g - 1.0)), 2.0 / (g + 1.0) * (-ar + 0.5 * (g - 1.0) * ur + s),
                              pr * c ** (2.0 * g / (g - 1.0)))
    return out
