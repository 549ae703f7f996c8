"""A generally contracted test basis and its fully uncontracted twin.

A shell with several contraction columns over one primitive set (nctr > 1,
PySCF orders its functions contraction-major) spans functions that are
linear combinations of the primitives.  With the primitives as shells of
their own (each normalized), the contracted functions are chi = phi C, so
every quantity linear in the basis functions obeys a closed relation:

    X_gc = X_prim C,     K_gc[D] = C^T K_prim[C D C^T] C,
    vmat_gc[D] = C^T vmat_prim[C D C^T] C,

which pins the handling of general contraction (column order, AO order,
per-column normalization) independently of how either side splits shells.
The exponents and contraction pattern are those of the correlation-
consistent double-zeta sets for H and O (general contraction of the s and
p primitives); any such pattern would do.
"""
import numpy

from pyscf_wb97mv_fast_amd.systems import mole

GC_BASIS = {
    'H': [(0, [(13.01, 0.019685, 0.0), (1.962, 0.137977, 0.0), (0.4446, 0.478148, 0.0),
               (0.1220, 0.501240, 1.0)]),
          (1, [(0.727, 1.0)])],
    'O': [(0, [(11720.0, 0.000710, -0.000160, 0.0), (1759.0, 0.005470, -0.001263, 0.0),
               (400.8, 0.027837, -0.006267, 0.0), (113.7, 0.104800, -0.025716, 0.0),
               (37.03, 0.283062, -0.070924, 0.0), (13.27, 0.448719, -0.165411, 0.0),
               (5.025, 0.270952, -0.116955, 0.0), (1.013, 0.015458, 0.557368, 0.0),
               (0.3023, -0.002585, 0.572759, 1.0)]),
          (1, [(17.70, 0.043018, 0.0), (3.854, 0.228913, 0.0), (1.046, 0.508728, 0.0),
               (0.2753, 0.460531, 1.0)]),
          (2, [(1.185, 1.0)])],
}


def prim_basis(bset):
    """Every primitive of every shell as a one-primitive shell, same order."""
    out = {}
    for sym, shells in bset.items():
        out[sym] = [(l, [(p[0], 1.0)]) for l, prims in shells for p in prims]
    return out


def contraction_matrix(mol_gc, mol_prim):
    """C (nao_prim x nao_gc) with chi_gc = phi_prim C: column k of a shell
    with exponents a_p holds c_pk / N(l, a_p) on the m-th function of
    primitive p, N the primitive normalization both molecules use."""
    nao_p = int(sum(2 * mol_prim.bas_angular(i) + 1 for i in range(mol_prim.nbas)))
    nao_g = int(sum((2 * mol_gc.bas_angular(i) + 1) * mol_gc.bas_nctr(i)
                    for i in range(mol_gc.nbas)))
    C = numpy.zeros((nao_p, nao_g))
    ip = ig = 0                      # AO offsets
    sp = 0                           # primitive shell index
    for ib in range(mol_gc.nbas):
        l = mol_gc.bas_angular(ib)
        d = 2 * l + 1
        exps = numpy.asarray(mol_gc.bas_exp(ib))
        cs = numpy.asarray(mol_gc._libcint_ctr_coeff(ib)).reshape(len(exps), -1)
        for p, a in enumerate(exps):
            assert mol_prim.bas_angular(sp + p) == l
            assert abs(mol_prim.bas_exp(sp + p)[0] - a) == 0.0
            norm = mole.gto_norm(l, a)
            for k in range(cs.shape[1]):
                for m in range(d):
                    C[ip + p * d + m, ig + k * d + m] = cs[p, k] / norm
        ip += len(exps) * d
        ig += cs.shape[1] * d
        sp += len(exps)
    return C


def molecules(atoms):
    """(generally contracted molecule, uncontracted twin, C)."""
    mg = mole.Mole(atoms, GC_BASIS)
    mp = mole.Mole(atoms, prim_basis(GC_BASIS))
    return mg, mp, contraction_matrix(mg, mp)
