"""Basis packing: a PySCF-style molecule -> the flat shell arrays of the C ABI.

Reads the accessors PySCF's ``gto.Mole`` exposes (``bas_angular``,
``bas_exp``, ``_libcint_ctr_coeff``, ``bas_coord``, ``bas_kappa``, ``cart``),
as the reference's ShellPack does (gpu/ao_eval.py:57-131).  A generally
contracted shell (nctr > 1) becomes one segmented shell per contraction
column over the same primitives; AO order is unchanged (PySCF orders the
functions of such a shell contraction-major, gpu/ao_eval.py:25-27).
"""
import numpy

from .errors import Unsupported

LMAX = 3


class ShellArrays:
    def __init__(self, l, nprim, center, exps, coefs):
        self.l = numpy.asarray(l, dtype=numpy.int32)
        self.nprim = numpy.asarray(nprim, dtype=numpy.int32)
        self.center = numpy.asarray(center, dtype=numpy.float64).reshape(-1, 3)
        self.exps = numpy.asarray(exps, dtype=numpy.float64)
        self.coefs = numpy.asarray(coefs, dtype=numpy.float64)
        self.nao = int((2 * self.l + 1).sum())


def pack_mol(mol):
    if getattr(mol, 'cart', False):
        raise Unsupported('GPU path supports spherical basis functions only')
    l, nprim, cen, exps, coefs = [], [], [], [], []
    for ib in range(mol.nbas):
        ang = int(mol.bas_angular(ib))
        if hasattr(mol, 'bas_kappa') and int(mol.bas_kappa(ib)) != 0:
            raise Unsupported('spinor shells are not supported')
        if ang > LMAX:
            raise Unsupported('shell %d has l = %d > %d' % (ib, ang, LMAX))
        e = numpy.asarray(mol.bas_exp(ib), dtype=numpy.float64)
        c = numpy.asarray(mol._libcint_ctr_coeff(ib), dtype=numpy.float64)
        if c.ndim == 1:
            c = c[:, None]
        for k in range(c.shape[1]):
            l.append(ang)
            nprim.append(len(e))
            cen.append(numpy.asarray(mol.bas_coord(ib), dtype=numpy.float64))
            exps.append(e)
            coefs.append(c[:, k])
    return ShellArrays(l, nprim, numpy.asarray(cen), numpy.concatenate(exps),
                       numpy.concatenate(coefs))
