"""Exact Riemann solver for the ideal-gas Euler equations (Toro, "Riemann
Solvers and Numerical Methods for Fluid Dynamics", 3rd ed., chapter 4:
pressure function f_K, Newton iteration for p*, sampling of the wave
pattern).  Test infrastructure: the textbook answer the 1D Sod run of the
oracle is pinned to."""
import math

import numpy as np


def _fK(p, rhoK, pK, cK, g):
    if p > pK:  # shock
        A = 2.0 / ((g + 1.0) * rhoK)
        B = (g - 1.0) / (g + 1.0) * pK
        f = (p - pK) * math.sqrt(A / (p + B))
        df = math.sqrt(A / (B + p)) * (1.0 - (p - pK) / (2.0 * (B + p)))
    else:  # rarefaction
        f = 2.0 * cK / (g - 1.0) * ((p / pK) ** ((g - 1.0) / (2.0 * g)) - 1.0)
        df = 1.0 / (rhoK * cK) * (p / pK) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rhoL, uL, pL, rhoR, uR, pR, g=1.4):
    cL = math.sqrt(g * pL / rhoL)
    cR = math.sqrt(g * pR / rhoR)
    p = max(1e-12, 0.5 * (pL + pR))
    for _ in range(100):
        fL, dL = _fK(p, rhoL, pL, cL, g)
        fR, dR = _fK(p, rhoR, pR, cR, g)
        dp = (fL + fR + uR - uL) / (dL + dR)
        pn = max(1e-12, p - dp)
        if abs(pn - p) < 1e-15 * (pn + p):
            p = pn
            break
        p = pn
    fL, _ = _fK(p, rhoL, pL, cL, g)
    fR, _ = _fK(p, rhoR, pR, cR, g)
    u = 0.5 * (uL + uR) + 0.5 * (fR - fL)
    return p, u


def sample(x, t, x0, rhoL, uL, pL, rhoR, uR, pR, g=1.4):
    """Exact (rho, u, p) at positions x, time t, initial jump at x0."""
    ps, us = star_state(rhoL, uL, pL, rhoR, uR, pR, g)
    cL = math.sqrt(g * pL / rhoL)
    cR = math.sqrt(g * pR / rhoR)
    out = np.zeros((3, len(x)))
    for n, xi in enumerate(x):
        S = (xi - x0) / t
        if S <= us:
            if ps > pL:  # left shock
                SL = uL - cL * math.sqrt((g + 1) / (2 * g) * ps / pL + (g - 1) / (2 * g))
                if S <= SL:
                    r, u, p = rhoL, uL, pL
                else:
                    r = rhoL * ((ps / pL + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pL + 1))
                    u, p = us, ps
            else:  # left rarefaction
                SHL = uL - cL
                cs = cL * (ps / pL) ** ((g - 1) / (2 * g))
                STL = us - cs
                if S <= SHL:
                    r, u, p = rhoL, uL, pL
                elif S >= STL:
                    r, u, p = rhoL * (ps / pL) ** (1 / g), us, ps
                else:
                    c = 2 / (g + 1) * (cL + (g - 1) / 2 * (uL - S))
                    u = 2 / (g + 1) * (cL + (g - 1) / 2 * uL + S)
                    r = rhoL * (c / cL) ** (2 / (g - 1))
                    p = pL * (c / cL) ** (2 * g / (g - 1))
        else:
            if ps > pR:  # right shock
                SR = uR + cR * math.sqrt((g + 1) / (2 * g) * ps / pR + (g - 1) / (2 * g))
                if S >= SR:
                    r, u, p = rhoR, uR, pR
                else:
                    r = rhoR * ((ps / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * ps / pR + 1))
                    u, p = us, ps
            else:  # right rarefaction
                SHR = uR + cR
                cs = cR * (ps / pR) ** ((g - 1) / (2 * g))
                STR = us + cs
                if S >= SHR:
                    r, u, p = rhoR, uR, pR
                elif S <= STR:
                    r, u, p = rhoR * (ps / pR) ** (1 / g), us, ps
                else:
                    c = 2 / (g + 1) * (cR - (g - 1) / 2 * (uR - S))
                    u = 2 / (g + 1) * (-cR + (g - 1) / 2 * uR + S)
                    r = rhoR * (c / cR) ** (2 / (g - 1))
                    p = pR * (c / cR) ** (2 * g / (g - 1))
        out[:, n] = (r, u, p)
    return out
