"""Atom-centred Becke integration grids for the synthetic workloads.

PySCF's ``gen_grid`` (Lebedev angular rules) is not available offline, so
this builds its own grid of the same construction:

* radial: Treutler-Ahlrichs M4 mapping of Chebyshev (2nd kind) nodes,
  r = xi/ln2 (1+x)^0.6 ln(2/(1-x));
* angular: product rule, Gauss-Legendre in cos(theta) times a uniform
  trapezoid in phi, exact for spherical harmonics up to its degree; the
  degree is PySCF's Lebedev order minus 4, which gives about the Lebedev
  point count;
* pruning: lower angular degree inside 0.25 and 0.5 of the Bragg radius
  (the NWChem scheme's idea);
* Stratmann-Scuseria-Frisch partition (Becke cells with compact support);
  points with weight below ``w_cut`` are dropped.  ``Grids(..., device=ctx)``
  builds the points, the partition and the cut on the GPU (gpu/grids.py);
  the host version here is the default and the test reference.

Radial/angular sizes per level follow PySCF's ``RAD_GRIDS``/``ANG_ORDER``
tables for periods 1-3.  ``Grids`` mirrors the attributes the hot path reads
from a PySCF grids object (``coords``, ``weights``, ``level``).
"""
import numpy

from .mole import BOHR

# PySCF gen_grid tables, periods 1..3 (levels 0..9)
_RAD = [(10, 15, 20), (30, 40, 50), (40, 60, 65), (50, 75, 80), (60, 90, 95),
        (70, 105, 110), (80, 120, 125), (90, 135, 140), (100, 150, 155),
        (200, 200, 200)]
_ANG = [(11, 15, 17), (17, 23, 23), (23, 29, 29), (29, 29, 35), (35, 41, 41),
        (41, 47, 47), (47, 53, 53), (53, 59, 59), (59, 59, 59), (65, 65, 65)]
_PERIOD = {1: 0, 6: 1, 7: 1, 8: 1, 9: 1, 16: 2, 17: 2}
# Treutler-Ahlrichs xi
_XI = {1: 0.8, 6: 1.1, 7: 0.9, 8: 0.9, 9: 0.9, 16: 1.0, 17: 0.9}
# Bragg radii (Angstrom)
_BRAGG = {1: 0.35, 6: 0.70, 7: 0.65, 8: 0.60, 9: 0.50, 16: 1.00, 17: 1.00}


def radial_treutler(n, xi):
    i = numpy.arange(1, n + 1)
    x = numpy.cos(i * numpy.pi / (n + 1))
    s = numpy.sin(i * numpy.pi / (n + 1))
    wcheb = numpy.pi / (n + 1) * s * s / numpy.sqrt(1.0 - x * x)
    ln2 = xi / numpy.log(2.0)
    r = ln2 * (1 + x) ** 0.6 * numpy.log(2.0 / (1 - x))
    dr = ln2 * (0.6 * (1 + x) ** -0.4 * numpy.log(2.0 / (1 - x))
                + (1 + x) ** 0.6 / (1 - x))
    return r, wcheb * dr * r * r


def angular_product(degree):
    """Unit-sphere points and weights (sum 4 pi), exact to `degree`."""
    nt = (degree + 2) // 2
    nphi = degree + 1
    ct, wt = numpy.polynomial.legendre.leggauss(nt)
    phi = 2.0 * numpy.pi * (numpy.arange(nphi) + 0.5) / nphi
    st = numpy.sqrt(1.0 - ct * ct)
    x = (st[:, None] * numpy.cos(phi)[None, :]).ravel()
    y = (st[:, None] * numpy.sin(phi)[None, :]).ravel()
    z = numpy.repeat(ct, nphi)
    w = numpy.repeat(wt, nphi) * (2.0 * numpy.pi / nphi)
    return numpy.stack([x, y, z], axis=1), w


_SSF_A = 0.64


def _ssf_partition(coords, atoms, owner, chunk=8192):
    """Stratmann-Scuseria-Frisch partition weight of each point's owner.

    s(mu) = (1 - g(mu))/2 with g the SSF polynomial on |mu| < a and +-1
    outside; an atom c has a non-zero cell weight at r only if
    |r - c| - |r - n| < a |c - n| for the nearest atom n, which bounds
    the work."""
    natm = len(atoms)
    rij = numpy.linalg.norm(atoms[:, None] - atoms[None], axis=2)
    inv_r = numpy.where(rij > 0, 1.0 / numpy.where(rij > 0, rij, 1.0), 0.0)
    out = numpy.empty(len(coords))
    for i0 in range(0, len(coords), chunk):
        c = coords[i0:i0 + chunk]
        own = owner[i0:i0 + chunk]
        d = numpy.linalg.norm(c[:, None, :] - atoms[None, :, :], axis=2)
        near = d.argmin(axis=1)
        dn = d[numpy.arange(len(c)), near]
        # P_a = 0 unless mu(a, nearest) < a_ssf, i.e. d_a - d_n < a_ssf R
        rel = (d - dn[:, None]) < _SSF_A * rij[near] + 1e-12
        P = numpy.zeros((len(c), natm))
        for a in range(natm):
            sel = numpy.flatnonzero(rel[:, a])
            if sel.size == 0:
                continue
            da = d[sel]
            mu = (da[:, a, None] - da) * inv_r[a][None, :]
            nu = numpy.clip(mu / _SSF_A, -1.0, 1.0)
            g = (35 * nu - 35 * nu ** 3 + 21 * nu ** 5 - 5 * nu ** 7) / 16.0
            s = 0.5 * (1.0 - g)
            s[:, a] = 1.0
            P[sel, a] = numpy.prod(s, axis=1)
        tot = P.sum(axis=1)
        out[i0:i0 + chunk] = P[numpy.arange(len(c)), own] / tot
    return out


class Grids:
    """Becke grid with PySCF-style attributes (coords/weights in bohr)."""

    def __init__(self, mol, level=3, prune=True, w_cut=1e-15, device=None):
        self.mol = mol
        self.device = device
        self.level = int(level)
        self.prune = prune
        self.w_cut = w_cut
        self.coords = None
        self.weights = None

    def shell_tables(self):
        """The grid's radial shells and angular rules: per shell its atom,
        radius, radial weight and angular rule index; the rules as
        ((n, 3) unit vectors, (n,) weights), in first-use order."""
        charges = self.mol.atom_charges()
        sh_atom, sh_r, sh_w, sh_rule = [], [], [], []
        rules, rule_of = [], {}
        for ia, z in enumerate(charges):
            per = _PERIOD[int(z)]
            nrad = _RAD[self.level][per]
            # product-rule degree giving about the point count of the
            # Lebedev rule of PySCF's order
            deg = max(7, _ANG[self.level][per] - 4)
            r, wr = radial_treutler(nrad, _XI[int(z)])
            rb = _BRAGG[int(z)] / BOHR
            for ri, wri in zip(r, wr):
                d = deg
                if self.prune:
                    if ri < 0.25 * rb:
                        d = min(deg, 7)
                    elif ri < 0.5 * rb:
                        d = min(deg, 11)
                if d not in rule_of:
                    rule_of[d] = len(rules)
                    rules.append(angular_product(d))
                sh_atom.append(ia)
                sh_r.append(ri)
                sh_w.append(wri)
                sh_rule.append(rule_of[d])
        return (numpy.asarray(sh_atom, dtype=numpy.int32), numpy.asarray(sh_r, dtype=float),
                numpy.asarray(sh_w, dtype=float), numpy.asarray(sh_rule, dtype=numpy.int32), rules)

    def raw_points(self):
        """All atom-centred points before the partition: coords (n, 3),
        radial x angular weights (n,), owner atom (n,)."""
        atoms = self.mol.atom_coords()
        sh_atom, sh_r, sh_w, sh_rule, rules = self.shell_tables()
        pts, wts, owner = [], [], []
        for ia, ri, wri, k in zip(sh_atom, sh_r, sh_w, sh_rule):
            u, wa = rules[k]
            pts.append(atoms[ia] + ri * u)
            wts.append(wri * wa)
            owner.append(numpy.full(len(wa), ia))
        return numpy.concatenate(pts), numpy.concatenate(wts), numpy.concatenate(owner)

    def build(self, **kwargs):
        """coords and weights; with device=ctx the points, the partition and
        the weight cut are made on the GPU (gpu.grids.build_grid), bit for
        bit the host's points and cut with the device partition."""
        atoms = self.mol.atom_coords()
        if self.device is not None:
            from ..gpu.grids import build_grid
            self.coords, self.weights = build_grid(self.device, atoms, *self.shell_tables(),
                                                   w_cut=self.w_cut)
            return self
        coords, w, owner = self.raw_points()
        w = w * _ssf_partition(coords, atoms, owner)
        keep = w > self.w_cut
        self.coords = numpy.ascontiguousarray(coords[keep])
        self.weights = numpy.ascontiguousarray(w[keep])
        return self

    @property
    def size(self):
        return 0 if self.weights is None else self.weights.size
