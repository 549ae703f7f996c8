"""Independent FP64 formulas used only to pin the oracle.

Written separately from oracle/wb_oracle.c, with a different method:
Obara-Saika vertical/horizontal recursion for the 3-centre nuclear-attraction
type integrals (instead of McMurchie-Davidson), scipy's incomplete gamma for
the Boys function, and a literal VV10 double loop.
"""
import math

import numpy
from scipy.special import gamma, gammainc


def boys_closed(n, T):
    """F_n(T) = Gamma(n+1/2) P(n+1/2, T) / (2 T^{n+1/2})."""
    T = numpy.asarray(T, dtype=numpy.float64)
    out = numpy.empty_like(T)
    small = T < 1e-12
    out[small] = 1.0 / (2 * n + 1)
    Ts = T[~small]
    out[~small] = gamma(n + 0.5) * gammainc(n + 0.5, Ts) / (2.0 * Ts ** (n + 0.5))
    return out


def _cart(l):
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def os_esp_prim(la, lb, a, b, A, B, C, omega=0.0):
    """Cartesian primitive block [(la comps) x (lb comps)] of
    int x_A^.. e^{-a r_A^2} x_B^.. e^{-b r_B^2} v(|r-C|) dr by Obara-Saika.
    v = 1/r, or erf(omega r)/r through theta^(m+1/2) F_m(theta T)."""
    A, B, C = (numpy.asarray(v, dtype=numpy.float64) for v in (A, B, C))
    p = a + b
    P = (a * A + b * B) / p
    PA, PB, PC = P - A, P - B, P - C
    Kab = math.exp(-a * b / p * float(((A - B) ** 2).sum()))
    th = 1.0 if omega <= 0 else omega * omega / (omega * omega + p)
    T = th * p * float((PC ** 2).sum())
    L = la + lb
    Fm = [boys_closed(m, numpy.array([T]))[0] * th ** (m + 0.5) for m in range(L + 1)]
    memo = {}

    def vrr(ax, ay, az, m):
        # (a|0)^(m) by vertical recursion on the first centre
        key = (ax, ay, az, m)
        if key in memo:
            return memo[key]
        if ax < 0 or ay < 0 or az < 0:
            return 0.0
        if ax == ay == az == 0:
            v = 2 * math.pi / p * Kab * Fm[m]
        else:
            e = [ax, ay, az]
            i = 0 if ax > 0 else (1 if ay > 0 else 2)
            e1 = list(e); e1[i] -= 1
            e2 = list(e1); e2[i] -= 1
            v = PA[i] * vrr(*e1, m) - PC[i] * vrr(*e1, m + 1)
            if e1[i] > 0:
                v += e1[i] / (2 * p) * (vrr(*e2, m) - vrr(*e2, m + 1))
        memo[key] = v
        return v

    def hrr(ea, eb):
        # (a|b) = (a+1_i|b-1_i) + AB_i (a|b-1_i)
        if sum(eb) == 0:
            return vrr(*ea, 0)
        i = 0 if eb[0] > 0 else (1 if eb[1] > 0 else 2)
        ea1 = list(ea); ea1[i] += 1
        eb1 = list(eb); eb1[i] -= 1
        return hrr(ea1, eb1) + (A[i] - B[i]) * hrr(ea, eb1)

    ca, cb = _cart(la), _cart(lb)
    return numpy.array([[hrr(list(u), list(v)) for v in cb] for u in ca])


def os_esp_shellpair(la, exps_a, coef_a, A, lb, exps_b, coef_b, B, C, Ta, Tb, omega=0.0):
    """Contracted spherical block (2la+1, 2lb+1); Ta/Tb (2l+1, ncart)."""
    V = numpy.zeros((len(_cart(la)), len(_cart(lb))))
    for a, ca in zip(exps_a, coef_a):
        for b, cb in zip(exps_b, coef_b):
            V += ca * cb * os_esp_prim(la, lb, a, b, A, B, C, omega)
    return Ta @ V @ Tb.T


def vv10_literal(coords, q, W0, kappa):
    n = len(coords)
    F = numpy.zeros(n); U = numpy.zeros(n); W = numpy.zeros(n)
    for i in range(n):
        d = coords[i] - coords
        R2 = (d * d).sum(axis=1)
        g = R2 * W0[i] + kappa[i]
        gp = R2 * W0 + kappa
        gt = g + gp
        T = q / (g * gp * gt)
        F[i] = -1.5 * T.sum()
        Tu = T * (1.0 / g + 1.0 / gt)
        U[i] = Tu.sum()
        W[i] = (Tu * R2).sum()
    return F, U, W


def ssf_polynomial():
    """Coefficients (x, x^3, x^5, x^7) of the odd degree-7 polynomial g with
    g(1) = 1 and g'(1) = g''(1) = g'''(1) = 0 (the Stratmann-Scuseria-Frisch
    switching function), solved from those conditions."""
    pw = numpy.array([1, 3, 5, 7])
    rows = [numpy.ones(4), pw.astype(float), pw * (pw - 1.0), pw * (pw - 1.0) * (pw - 2.0)]
    return numpy.linalg.solve(numpy.array(rows), numpy.array([1.0, 0, 0, 0]))


def ssf_brute(coords, atoms, owner, a=0.64):
    """Owner's cell weight P_o / sum_a P_a over ALL atoms (no nearest-atom
    screening), P_a = prod_{b != a} (1 - g(clip(mu_ab / a))) / 2."""
    c = ssf_polynomial()
    out = numpy.empty(len(coords))
    natm = len(atoms)
    for i, r in enumerate(coords):
        d = [math.dist(r, A) for A in atoms]
        P = []
        for ia in range(natm):
            p = 1.0
            for ib in range(natm):
                if ib == ia:
                    continue
                mu = (d[ia] - d[ib]) / math.dist(atoms[ia], atoms[ib])
                x = min(1.0, max(-1.0, mu / a))
                p *= 0.5 * (1.0 - (c[0] * x + c[1] * x ** 3 + c[2] * x ** 5 + c[3] * x ** 7))
            P.append(p)
        out[i] = P[owner[i]] / sum(P)
    return out
