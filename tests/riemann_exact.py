"""Exact solution of the 1-D Euler Riemann problem (Toro, "Riemann Solvers and Numerical
Methods for Fluid Dynamics", 3rd ed., ch. 4: pressure function, Newton iteration for p*,
sampling of the wave pattern). Used as an independent check of the HLLC oracle."""
import math

G = 1.4


def _fk(p, rk, pk, ck):
    if p > pk:  # shock
        Ak = 2.0 / ((G + 1.0) * rk)
        Bk = (G - 1.0) / (G + 1.0) * pk
        q = math.sqrt(Ak / (p + Bk))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (p + Bk))
    pr = p / pk  # rarefaction
    f = 2.0 * ck / (G - 1.0) * (pr ** ((G - 1.0) / (2.0 * G)) - 1.0)
    df = 1.0 / (rk * ck) * pr ** (-(G + 1.0) / (2.0 * G))
    return f, df


def star(rl, ul, pl, rr, ur, pr):
    cl, cr = math.sqrt(G * pl / rl), math.sqrt(G * pr / rr)
    p = max(1e-8, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dfl = _fk(p, rl, pl, cl)
        fr, dfr = _fk(p, rr, pr, cr)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-10)
        if abs(pn - p) < 1e-14 * (pn + p):
            p = pn
            break
        p = pn
    fl, _ = _fk(p, rl, pl, cl)
    fr, _ = _fk(p, rr, pr, cr)
    return p, 0.5 * (ul + ur) + 0.5 * (fr - fl)


def sample(rl, ul, pl, rr, ur, pr, s):
    """State (rho, u, p) at x/t = s."""
    ps, us = star(rl, ul, pl, rr, ur, pr)
    cl, cr = math.sqrt(G * pl / rl), math.sqrt(G * pr / rr)
    if s <= us:
        if ps > pl:
            sl = ul - cl * math.sqrt((G + 1) / (2 * G) * ps / pl + (G - 1) / (2 * G))
            if s <= sl:
                return rl, ul, pl
            return rl * ((ps / pl + (G - 1) / (G + 1)) / ((G - 1) / (G + 1) * ps / pl + 1)), us, ps
        shl = ul - cl
        if s <= shl:
            return rl, ul, pl
        cml = cl * (ps / pl) ** ((G - 1) / (2 * G))
        stl = us - cml
        if s > stl:
            return rl * (ps / pl) ** (1 / G), us, ps
        c = 2 / (G + 1) * (cl + (G - 1) / 2 * (ul - s))
        u = 2 / (G + 1) * (cl + (G - 1) / 2 * ul + s)
        r = rl * (c / cl) ** (2 / (G - 1))
        return r, u, pl * (c / cl) ** (2 * G / (G - 1))
    if ps > pr:
        sr = ur + cr * math.sqrt((G + 1) / (2 * G) * ps / pr + (G - 1) / (2 * G))
        if s >= sr:
            return rr, ur, pr
        return rr * ((ps / pr + (G - 1) / (G + 1)) / ((G - 1) / (G + 1) * ps / pr + 1)), us, ps
    shr = ur + cr
    if s >= shr:
        return rr, ur, pr
    cmr = cr * (ps / pr) ** ((G - 1) / (2 * G))
    str_ = us + cmr
    if s <= str_:
        return rr * (ps / pr) ** (1 / G), us, ps
    c = 2 / (G + 1) * (cr - (G - 1) / 2 * (ur - s))
    u = 2 / (G + 1) * (-cr + (G - 1) / 2 * ur + s)
    r = rr * (c / cr) ** (2 / (G - 1))
    return r, u, pr * (c / cr) ** (2 * G / (G - 1))
