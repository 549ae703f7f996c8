"""Closed-shell guess density for the synthetic workloads.

PySCF's minao guess needs its atomic basis library, so the workloads use an
extended-Hueckel guess instead: H_uu = shell energy, H_uv = 1.75 S_uv (H_uu +
H_vv)/2, solve H C = S C e and doubly occupy the lowest nelec/2 orbitals.
The result is symmetric and idempotent in the S metric (D S D = 2 D), with
the inter-atomic structure of a real density -- what the COSX screening and
the exchange energy are sensitive to.

The overlap matrix is analytic (Obara-Saika 1-D recursion for Cartesian
Gaussians, then the cart -> sph transform).
"""
import math

import numpy
import scipy.linalg

from ..cart2sph import cart2sph, cart_components


def _overlap_1d(la, lb, a, b, xa, xb):
    """s[i, j], i <= la, j <= lb, for primitive exponents a, b (vectors)."""
    p = a + b
    P = (a * xa + b * xb) / p
    s = numpy.zeros((la + 1, lb + 1) + numpy.shape(p))
    s[0, 0] = numpy.sqrt(numpy.pi / p) * numpy.exp(-a * b / p * (xa - xb) ** 2)
    pa, pb, h = P - xa, P - xb, 0.5 / p
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            if i > 0:
                v = pa * s[i - 1, j]
                if i > 1:
                    v = v + h * (i - 1) * s[i - 2, j]
                if j > 0:
                    v = v + h * j * s[i - 1, j - 1]
            else:
                v = pb * s[i, j - 1]
                if j > 1:
                    v = v + h * (j - 1) * s[i, j - 2]
            s[i, j] = v
    return s


def overlap_matrix(mol):
    """Analytic AO overlap, shell pairs of one kind (angular momenta, primitive
    and contraction counts) evaluated together as arrays."""
    ao_loc = mol.ao_loc_nr()
    nao = int(ao_loc[-1])
    S = numpy.zeros((nao, nao))
    nb = mol.nbas
    ls = [int(mol.bas_angular(ib)) for ib in range(nb)]
    exps = [numpy.asarray(mol.bas_exp(ib), dtype=float) for ib in range(nb)]
    coefs = [numpy.asarray(mol._libcint_ctr_coeff(ib), dtype=float).reshape(len(exps[ib]), -1)
             for ib in range(nb)]
    xyz = numpy.array([mol.bas_coord(ib) for ib in range(nb)], dtype=float)
    kind = [(ls[ib], len(exps[ib]), coefs[ib].shape[1]) for ib in range(nb)]
    groups = {}
    for i in range(nb):
        for j in range(i + 1):
            groups.setdefault((kind[i], kind[j]), []).append((i, j))
    for ((li, npi, nci), (lj, npj, ncj)), pairs in groups.items():
        I = numpy.array([p[0] for p in pairs])
        J = numpy.array([p[1] for p in pairs])
        a = numpy.stack([exps[i] for i in I])[:, :, None]          # (k, npi, 1)
        b = numpy.stack([exps[j] for j in J])[:, None, :]          # (k, 1, npj)
        ci = numpy.stack([coefs[i] for i in I])                    # (k, npi, nci)
        cj = numpy.stack([coefs[j] for j in J])                    # (k, npj, ncj)
        s1 = [_overlap_1d(li, lj, a, b, xyz[I, d][:, None, None], xyz[J, d][:, None, None])
              for d in range(3)]
        comps_i, comps_j = cart_components(li), cart_components(lj)
        cart = numpy.empty((len(comps_i), len(comps_j), len(I), nci, ncj))
        for u, (x1, y1, z1) in enumerate(comps_i):
            for v, (x2, y2, z2) in enumerate(comps_j):
                prim = s1[0][x1, x2] * s1[1][y1, y2] * s1[2][z1, z2]
                cart[u, v] = numpy.einsum('kpc,kpq,kqd->kcd', ci, prim, cj)
        blk = numpy.einsum('mu,nv,uvkcd->kcmdn', cart2sph(li), cart2sph(lj), cart)
        di, dj = nci * (2 * li + 1), ncj * (2 * lj + 1)
        blk = blk.reshape(len(I), di, dj)
        rows = ao_loc[I][:, None] + numpy.arange(di)[None, :]
        cols = ao_loc[J][:, None] + numpy.arange(dj)[None, :]
        S[rows[:, :, None], cols[:, None, :]] = blk
        S[cols[:, :, None], rows[:, None, :]] = blk.transpose(0, 2, 1)
    return S


# valence-state ionization energies (eV) of extended Hueckel theory
_VSIP = {1: {'s': -13.6},
         6: {'core': -290.0, 's': -21.4, 'p': -11.4},
         7: {'core': -400.0, 's': -26.0, 'p': -13.4},
         8: {'core': -540.0, 's': -32.3, 'p': -14.8},
         9: {'core': -690.0, 's': -40.0, 'p': -18.1},
         16: {'core': -230.0, 's': -20.0, 'p': -13.3},
         17: {'core': -270.0, 's': -26.3, 'p': -14.2}}


def _shell_energies(mol):
    """Diagonal Hamiltonian per AO: the most compact s shell of a heavy atom
    is the core, the next s and p shells the valence, everything else
    (diffuse / polarization) a weakly bound -5 eV."""
    diag = []
    seen = {}
    for ib in range(mol.nbas):
        ia = mol.bas_atom(ib)
        z = int(mol.atom_charges()[ia])
        l = mol.bas_angular(ib)
        key = (ia, l)
        rank = seen.get(key, 0)
        seen[key] = rank + 1
        e = _VSIP[z]
        if l == 0:
            if z > 2 and rank == 0:
                val = e['core']
            elif (z > 2 and rank == 1) or (z <= 2 and rank == 0):
                val = e['s']
            else:
                val = -5.0
        elif l == 1 and z > 2 and rank == 0:
            val = e['p']
        else:
            val = -5.0
        diag += [val] * ((2 * l + 1) * mol.bas_nctr(ib))
    return numpy.asarray(diag) / 27.211386


def huckel_hamiltonian(mol, k=1.75):
    """(H, S): the extended-Hueckel Hamiltonian and the analytic overlap."""
    S = overlap_matrix(mol)
    h = _shell_energies(mol)
    H = k * S * (h[:, None] + h[None, :]) * 0.5
    numpy.fill_diagonal(H, h)
    return H, S


def huckel_density(mol, k=1.75):
    H, S = huckel_hamiltonian(mol, k)
    return huckel_density_from(H, S, mol.nelectron // 2)


def huckel_density_from(H, S, nocc):
    e, C = scipy.linalg.eigh(H, S)
    Cocc = C[:, :nocc]
    D = 2.0 * Cocc @ Cocc.T
    return 0.5 * (D + D.T)
