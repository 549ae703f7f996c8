"""Cartesian -> real-spherical transforms for l <= 3 (libcint conventions).

Cartesian components follow libcint's order (lx descending, then ly
descending, lz = l - lx - ly).  Row m of ``cart2sph(l)`` holds the
coefficients of r^l Y_lm in those monomials, Y_lm the real spherical
harmonics normalized to 1 on the unit sphere, m = -l..l; for l = 1 the
rows are (x, y, z) as in PySCF.  Combined with the radial normalization
int R^2 r^{2l+2} dr = 1 every spherical AO is normalized.

The matrices are stored transposed relative to PySCF's
``gto.mole.cart2sph(l)`` (which is (ncart, nsph)); ``cart2sph_pyscf``
returns that layout.
"""
import math

import numpy

LMAX = 3


def cart_components(l):
    return [(lx, ly, l - lx - ly)
            for lx in range(l, -1, -1)
            for ly in range(l - lx, -1, -1)]


def ncart(l):
    return (l + 1) * (l + 2) // 2


def _table():
    pi = math.pi
    s = math.sqrt
    t = {}
    t[0] = {0: {(0, 0, 0): 0.5 / s(pi)}}
    c1 = s(3.0 / (4 * pi))
    t[1] = {-1: {(1, 0, 0): c1}, 0: {(0, 1, 0): c1}, 1: {(0, 0, 1): c1}}
    a = s(15.0 / (4 * pi))
    b = s(5.0 / (16 * pi))
    c = s(15.0 / (16 * pi))
    t[2] = {-2: {(1, 1, 0): a},
            -1: {(0, 1, 1): a},
            0: {(2, 0, 0): -b, (0, 2, 0): -b, (0, 0, 2): 2 * b},
            1: {(1, 0, 1): a},
            2: {(2, 0, 0): c, (0, 2, 0): -c}}
    f3 = s(35.0 / (32 * pi))
    f2 = s(105.0 / (4 * pi))
    f1 = s(21.0 / (32 * pi))
    f0 = s(7.0 / (16 * pi))
    g2 = s(105.0 / (16 * pi))
    t[3] = {-3: {(2, 1, 0): 3 * f3, (0, 3, 0): -f3},
            -2: {(1, 1, 1): f2},
            -1: {(0, 1, 2): 4 * f1, (2, 1, 0): -f1, (0, 3, 0): -f1},
            0: {(0, 0, 3): 2 * f0, (2, 0, 1): -3 * f0, (0, 2, 1): -3 * f0},
            1: {(1, 0, 2): 4 * f1, (3, 0, 0): -f1, (1, 2, 0): -f1},
            2: {(2, 0, 1): g2, (0, 2, 1): -g2},
            3: {(3, 0, 0): f3, (1, 2, 0): -3 * f3}}
    return t


_T = _table()


def cart2sph(l):
    """(2l+1, ncart) matrix: sph_m = sum_c T[m, c] cart_c."""
    comps = cart_components(l)
    idx = {c: i for i, c in enumerate(comps)}
    T = numpy.zeros((2 * l + 1, len(comps)))
    # l = 1 uses PySCF's (x, y, z) row order, which the dict keys above
    # already encode as m = -1, 0, 1
    for m in range(-l, l + 1):
        for comp, v in _T[l][m].items():
            T[m + l, idx[comp]] = v
    return T


def cart2sph_pyscf(l):
    """(ncart, nsph), the layout of pyscf.gto.mole.cart2sph(l)."""
    return cart2sph(l).T.copy()
