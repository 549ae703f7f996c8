"""A minimal molecule object with PySCF's ``Mole`` data layout.

The host layer reads a basis through the accessors PySCF's ``gto.Mole``
exposes (``_atm``/``_bas``/``_env`` in libcint slot order, ``bas_*``,
``_libcint_ctr_coeff``, ``ao_loc_nr``, ``with_range_coulomb``).  This class
provides the same accessors for the synthetic workloads, so the same
extraction code serves a real PySCF molecule and a synthetic one.

Normalization follows PySCF's ``gto.mole.make_bas_env`` /
``_nomalize_contracted_ao``: each primitive coefficient is multiplied by
``gto_norm(l, alpha)`` and every contracted column is then scaled so the
radial function is normalized, ``int_0^inf R(r)^2 r^{2l+2} dr = 1``.
"""
import contextlib
import math

import numpy

from . import basis_data

BOHR = 0.52917721092          # Angstrom per bohr (PySCF's param.BOHR)

# libcint slot layout (cint.h)
CHARGE_OF, PTR_COORD, NUC_MOD_OF, PTR_ZETA = 0, 1, 2, 3
ATM_SLOTS = 6
ATOM_OF, ANG_OF, NPRIM_OF, NCTR_OF, KAPPA_OF, PTR_EXP, PTR_COEFF = range(7)
BAS_SLOTS = 8
PTR_RANGE_OMEGA = 8
PTR_ENV_START = 20


def gaussian_int(n, alpha):
    """int_0^inf x^n exp(-alpha x^2) dx."""
    n1 = (n + 1) * 0.5
    return math.gamma(n1) / (2.0 * alpha ** n1)


def gto_norm(l, alpha):
    return 1.0 / math.sqrt(gaussian_int(2 * l + 2, 2.0 * alpha))


def normalize_contracted(l, exps, coefs):
    """exps (nprim,), coefs (nprim, nctr) raw -> libcint-normalized."""
    exps = numpy.asarray(exps, dtype=numpy.float64)
    cs = numpy.asarray(coefs, dtype=numpy.float64).reshape(len(exps), -1)
    cs = cs * numpy.array([gto_norm(l, a) for a in exps])[:, None]
    ee = exps[:, None] + exps[None, :]
    n1 = l + 1.5
    ovl = numpy.array([[math.gamma(n1) / (2.0 * e ** n1) for e in row]
                       for row in ee])
    s1 = 1.0 / numpy.sqrt(numpy.einsum('pi,pq,qi->i', cs, ovl, cs))
    return cs * s1[None, :]


class Mole:
    """Molecule with a segmented spherical Gaussian basis.

    atoms: [(symbol, (x, y, z))]; coordinates in Angstrom unless
    unit='bohr'.  basis: a name in basis_data or a dict of the same form.
    """

    def __init__(self, atoms, basis='def2-svp', unit='angstrom', charge=0):
        self.cart = False
        self.unit = unit
        self.charge = int(charge)
        self.verbose = 0
        self.basis = basis
        bset = basis_data.get_basis(basis) if isinstance(basis, str) else basis
        scale = 1.0 / BOHR if unit.lower().startswith('a') else 1.0
        env = [0.0] * PTR_ENV_START
        atm, bas = [], []
        self._symbols = []
        for ia, (sym, xyz) in enumerate(atoms):
            if sym not in bset:
                raise KeyError('no %r shells in the basis' % sym)
            ptr = len(env)
            env.extend([float(c) * scale for c in xyz])
            env.append(0.0)                               # zeta slot
            atm.append([basis_data.CHARGES[sym], ptr, 1, ptr + 3, 0, 0])
            self._symbols.append(sym)
            for l, prims in bset[sym]:
                exps = numpy.array([p[0] for p in prims])
                coefs = numpy.array([p[1:] for p in prims])
                cs = normalize_contracted(l, exps, coefs)
                pe = len(env)
                env.extend(exps.tolist())
                pc = len(env)
                env.extend(cs.T.ravel().tolist())         # (nctr, nprim)
                bas.append([ia, l, len(exps), cs.shape[1], 0, pe, pc, 0])
        self._atm = numpy.asarray(atm, dtype=numpy.int32)
        self._bas = numpy.asarray(bas, dtype=numpy.int32)
        self._env = numpy.asarray(env, dtype=numpy.float64)

    # -- PySCF-compatible accessors -----------------------------------------
    @property
    def natm(self):
        return len(self._atm)

    @property
    def nbas(self):
        return len(self._bas)

    @property
    def omega(self):
        return float(self._env[PTR_RANGE_OMEGA])

    @omega.setter
    def omega(self, value):
        self._env[PTR_RANGE_OMEGA] = float(value)

    @contextlib.contextmanager
    def with_range_coulomb(self, omega):
        old = self.omega
        self.omega = omega
        try:
            yield self
        finally:
            self.omega = old

    def atom_symbol(self, ia):
        return self._symbols[ia]

    def atom_charges(self):
        return self._atm[:, CHARGE_OF].copy()

    def atom_coords(self):
        ptr = self._atm[:, PTR_COORD]
        return numpy.stack([self._env[p:p + 3] for p in ptr]).reshape(-1, 3)

    def bas_atom(self, ib):
        return int(self._bas[ib, ATOM_OF])

    def bas_angular(self, ib):
        return int(self._bas[ib, ANG_OF])

    def bas_nprim(self, ib):
        return int(self._bas[ib, NPRIM_OF])

    def bas_nctr(self, ib):
        return int(self._bas[ib, NCTR_OF])

    def bas_kappa(self, ib):
        return int(self._bas[ib, KAPPA_OF])

    def bas_exp(self, ib):
        p = self._bas[ib, PTR_EXP]
        return self._env[p:p + self._bas[ib, NPRIM_OF]].copy()

    def bas_coord(self, ib):
        return self.atom_coords()[self._bas[ib, ATOM_OF]]

    def _libcint_ctr_coeff(self, ib):
        """(nprim, nctr) normalized coefficients as stored for libcint."""
        nprim, nctr = self._bas[ib, NPRIM_OF], self._bas[ib, NCTR_OF]
        p = self._bas[ib, PTR_COEFF]
        return self._env[p:p + nprim * nctr].reshape(nctr, nprim).T.copy()

    def ao_loc_nr(self):
        l = self._bas[:, ANG_OF]
        dims = (2 * l + 1) * self._bas[:, NCTR_OF]
        loc = numpy.zeros(self.nbas + 1, dtype=numpy.int64)
        loc[1:] = numpy.cumsum(dims)
        return loc

    def nao_nr(self):
        return int(self.ao_loc_nr()[-1])

    @property
    def nao(self):
        return self.nao_nr()

    @property
    def nelectron(self):
        return int(self.atom_charges().sum()) - self.charge

    @property
    def nelec(self):
        n = self.nelectron
        return (n // 2, n - n // 2)
