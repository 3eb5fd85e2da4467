"""Exact solution of the 1D Euler Riemann problem for a gamma-law gas.

Textbook construction (Toro, "Riemann Solvers and Numerical Methods for Fluid
Dynamics", ch. 4): Newton iteration on the star-region pressure, then sampling
of the self-similar fan.  Used to pin the oracle on the Sod shock tube.
"""
import numpy as np


def _fk(p, rk, pk, ck, g):
    if p > pk:   # shock
        ak = 2.0 / ((g + 1.0) * rk)
        bk = (g - 1.0) / (g + 1.0) * pk
        q = np.sqrt(ak / (p + bk))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (bk + p))
    # rarefaction
    pr = p / pk
    f = 2.0 * ck / (g - 1.0) * (pr ** ((g - 1.0) / (2.0 * g)) - 1.0)
    df = 1.0 / (rk * ck) * pr ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl = np.sqrt(g * pl / rl)
    cr = np.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dfl = _fk(p, rl, pl, cl, g)
        fr, dfr = _fk(p, rr, pr, cr, g)
        dp = (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(1e-14, p - dp)
        if abs(pn - p) < 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, cl, g)
    fr, _ = _fk(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x, t, x0, left, right, g):
    """Return (rho, u, p) arrays at positions x and time t > 0."""
    rl, ul, pl = left
    rr, ur, pr = right
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl = np.sqrt(g * pl / rl)
    cr = np.sqrt(g * pr / rr)
    s = (np.asarray(x, dtype=float) - x0) / t
    rho = np.empty_like(s)
    u = np.empty_like(s)
    p = np.empty_like(s)
    gm = (g - 1.0) / (g + 1.0)
    for i, si in enumerate(s):
        if si <= us:   # left of contact
            if ps > pl:   # left shock
                rsl = rl * (ps / pl + gm) / (gm * ps / pl + 1.0)
                shl = ul - cl * np.sqrt((g + 1.0) / (2 * g) * ps / pl + (g - 1.0) / (2 * g))
                rho[i], u[i], p[i] = (rl, ul, pl) if si <= shl else (rsl, us, ps)
            else:   # left rarefaction
                csl = cl * (ps / pl) ** ((g - 1.0) / (2 * g))
                shl, stl = ul - cl, us - csl
                if si <= shl:
                    rho[i], u[i], p[i] = rl, ul, pl
                elif si >= stl:
                    rho[i], u[i], p[i] = rl * (ps / pl) ** (1.0 / g), us, ps
                else:
                    uu = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * ul + si)
                    cc = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * (ul - si))
                    rho[i] = rl * (cc / cl) ** (2.0 / (g - 1.0))
                    u[i] = uu
                    p[i] = pl * (cc / cl) ** (2.0 * g / (g - 1.0))
        else:
            if ps > pr:   # right shock
                rsr = rr * (ps / pr + gm) / (gm * ps / pr + 1.0)
                shr = ur + cr * np.sqrt((g + 1.0) / (2 * g) * ps / pr + (g - 1.0) / (2 * g))
                rho[i], u[i], p[i] = (rr, ur, pr) if si >= shr else (rsr, us, ps)
            else:
                csr = cr * (ps / pr) ** ((g - 1.0) / (2 * g))
                shr, stt = ur + cr, us + csr
                if si >= shr:
                    rho[i], u[i], p[i] = rr, ur, pr
                elif si <= stt:
                    rho[i], u[i], p[i] = rr * (ps / pr) ** (1.0 / g), us, ps
                else:
                    uu = 2.0 / (g + 1.0) * (-cr + (g - 1.0) / 2.0 * ur + si)
                    cc = 2.0 / (g + 1.0) * (cr - (g - 1.0) / 2.0 * (ur - si))
                    rho[i] = rr * (cc / cr) ** (2.0 / (g - 1.0))
                    u[i] = uu
                    p[i] = pr * (cc / cr) ** (2.0 * g / (g - 1.0))
    return rho, u, p


def cell_average(xlo, xhi, t, x0, left, right, g, sub=64):
    """Zone averages of (rho, u, p) over [xlo, xhi] by midpoint sub-sampling."""
    xs = xlo[:, None] + (np.arange(sub)[None, :] + 0.5) / sub * (xhi - xlo)[:, None]
    r, u, p = sample(xs.ravel(), t, x0, left, right, g)
    return (r.reshape(xs.shape).mean(1), u.reshape(xs.shape).mean(1), p.reshape(xs.shape).mean(1))
