"""Exact solution of the 1D special-relativistic Riemann problem (ideal gas,
zero tangential velocity), after Marti & Mueller (1994, JFM 258, 317) and
their Living Review (2003, LRR 6, 7).  Test infrastructure: a closed-form
check for the CPU oracle.

Waves: a left and a right nonlinear wave (rarefaction if p* < p_K, shock if
p* > p_K) around a contact.  p* solves v*_L(p*) = v*_R(p*).
  rarefaction: the Riemann invariant  atanh(v) +- A(c)  with
      A(c) = 2/sqrt(g-1) * atanh(c / sqrt(g-1))
    is constant across the fan (+ for the left wave), isentropic p ~ rho^g;
  shock: Taub adiabat for h*, mass flux j, shock speed V_s, and
      v* = (h W v + W_s (p* - p)/j) / (h W + (p* - p)(W_s v / j + 1/(rho W))).
"""
import math

from scipy.optimize import brentq


def _h(rho, p, g):
    return 1.0 + g * p / ((g - 1.0) * rho)


def _cs(rho, p, g):
    return math.sqrt(g * p / (rho * _h(rho, p, g)))


def _A(c, g):
    s = math.sqrt(g - 1.0)
    return 2.0 / s * math.atanh(c / s)


def _shock(p, rho, v, ps, g, sign):
    """State behind a shock into (rho, p, v) with post-shock pressure ps.
    sign = -1 for the left wave, +1 for the right wave."""
    h = _h(rho, p, g)
    W = 1.0 / math.sqrt(1.0 - v * v)
    # Taub adiabat, quadratic in h*:
    a = 1.0 + (g - 1.0) * (p - ps) / (g * ps)
    b = -(g - 1.0) * (p - ps) / (g * ps)
    c = h * (p - ps) / rho - h * h
    hs = (-b + math.sqrt(b * b - 4.0 * a * c)) / (2.0 * a)
    rhos = g * ps / ((g - 1.0) * (hs - 1.0))
    j = sign * math.sqrt((ps - p) / (h / rho - hs / rhos))
    A = rho * rho * W * W
    Vs = (A * v + sign * abs(j) * math.sqrt(j * j + A * (1.0 - v * v))) / (A + j * j)
    Ws = 1.0 / math.sqrt(1.0 - Vs * Vs)
    vs = (h * W * v + Ws * (ps - p) / j) / (h * W + (ps - p) * (Ws * v / j + 1.0 / (rho * W)))
    return vs, rhos, Vs


def _rare(p, rho, v, ps, g, sign):
    """State behind a rarefaction; sign = -1 left wave, +1 right wave."""
    rhos = rho * (ps / p) ** (1.0 / g)
    c = _cs(rho, p, g)
    cst = _cs(rhos, ps, g)
    # left: atanh(v) + A(c) conserved; right: atanh(v) - A(c) conserved
    t = math.atanh(v) + sign * (_A(cst, g) - _A(c, g))
    return math.tanh(t), rhos


def _vstar(p, rho, v, ps, g, sign):
    if ps > p:
        return _shock(p, rho, v, ps, g, sign)[0]
    return _rare(p, rho, v, ps, g, sign)[0]


def star_state(left, right, g):
    """left/right = (rho, p, v).  Returns p*, v*, rho*_L, rho*_R."""
    rL, pL, vL = left
    rR, pR, vR = right
    f = lambda ps: _vstar(pL, rL, vL, ps, g, -1) - _vstar(pR, rR, vR, ps, g, +1)
    # v*_L decreases and v*_R increases with p*: f > 0 below the root
    lo, hi = 1e-12 * min(pL, pR) + 1e-300, max(pL, pR)
    while f(hi) > 0:
        hi *= 2.0
    ps = brentq(f, lo, hi, xtol=1e-300, rtol=1e-15, maxiter=500)
    vs = _vstar(pL, rL, vL, ps, g, -1)
    rhoL = _shock(pL, rL, vL, ps, g, -1)[1] if ps > pL else _rare(pL, rL, vL, ps, g, -1)[1]
    rhoR = _shock(pR, rR, vR, ps, g, +1)[1] if ps > pR else _rare(pR, rR, vR, ps, g, +1)[1]
    return ps, vs, rhoL, rhoR


def sample(left, right, g, x, t, x0=0.5):
    """(rho, p, v) at positions x (iterable) and time t.  Only the wave
    pattern left-rarefaction + right-shock (the shock-tube case) and its
    mirror are handled, which covers the shock tubes used in the tests."""
    rL, pL, vL = left
    rR, pR, vR = right
    ps, vs, rhoLs, rhoRs = star_state(left, right, g)
    out = []
    # left wave
    if ps <= pL:
        cL = _cs(rL, pL, g)
        head = (vL - cL) / (1.0 - vL * cL)
        cLs = _cs(rhoLs, ps, g)
        tail = (vs - cLs) / (1.0 - vs * cLs)
    else:
        head = tail = _shock(pL, rL, vL, ps, g, -1)[2]
    # right wave
    if ps > pR:
        rhead = rtail = _shock(pR, rR, vR, ps, g, +1)[2]
    else:
        cR = _cs(rR, pR, g)
        rhead = (vR + cR) / (1.0 + vR * cR)
        cRs = _cs(rhoRs, ps, g)
        rtail = (vs + cRs) / (1.0 + vs * cRs)
    for xi in x:
        xi_ = (xi - x0) / t
        if xi_ < head:
            out.append((rL, pL, vL))
        elif xi_ < tail:
            # inside the left fan: solve for the state whose C- characteristic
            # speed equals xi_, keeping the left invariant constant
            def fn(pp):
                v, rho = _rare(pL, rL, vL, pp, g, -1)
                c = _cs(rho, pp, g)
                return (v - c) / (1.0 - v * c) - xi_
            pp = brentq(fn, ps, pL, xtol=1e-14 * pL)
            v, rho = _rare(pL, rL, vL, pp, g, -1)
            out.append((rho, pp, v))
        elif xi_ < vs:
            out.append((rhoLs, ps, vs))
        elif xi_ < rtail:
            out.append((rhoRs, ps, vs))
        elif xi_ < rhead:
            def fn(pp):
                v, rho = _rare(pR, rR, vR, pp, g, +1)
                c = _cs(rho, pp, g)
                return (v + c) / (1.0 + v * c) - xi_
            pp = brentq(fn, ps, pR, xtol=1e-14 * pR)
            v, rho = _rare(pR, rR, vR, pp, g, +1)
            out.append((rho, pp, v))
        else:
            out.append((rR, pR, vR))
    return out
