"""Grid construction on the MI355X: the atomic cell partition, and the
quadrature overlap of the AOs on a grid.

Building an atom-centred grid is cheap except for the cell partition, which
weighs every point against every nearby atom pair (O(points x atoms^2) in the
worst case; 97% of a (H2O)27 grid build on the host).  ``build_grid`` makes
the whole grid on the device (``wb_grid_build``: points from the radial
shells and angular rules, the partition, the weight cut, in one pass with
no raw points crossing PCIe); ``ssf_partition`` runs the partition alone
(``wb_grid_partition``), both one thread per point in FP64, for the
Stratmann-Scuseria-Frisch cells the package's grids use (systems/grids.py).  The reference builds its SGX and XC grids on the host
through PySCF (grids.build(with_non0tab=True), gpu/install.py:164;
with_df.build(level=...), staging/schedule.py:391).
"""
import ctypes

import numpy

from .. import lib


def ssf_partition(ctx, coords, atoms, owner):
    """Partition weight of each point's owner atom (host arrays in and out).

    coords (n, 3) and atoms (natm, 3) in bohr, owner (n,) atom indices."""
    L = lib.load()
    coords = lib.f64(coords).reshape(-1, 3)
    atoms = lib.f64(atoms).reshape(-1, 3)
    owner = numpy.ascontiguousarray(owner, dtype=numpy.int32).ravel()
    if len(owner) != len(coords):
        raise ValueError('coords / owner size mismatch')
    out = numpy.empty(len(coords))
    lib.check(L.wb_grid_partition(ctx.h, len(atoms), lib.ptr(atoms), len(coords),
                                  lib.ptr(coords), lib.ptr(owner), lib.ptr(out)))
    return out


def build_grid(ctx, atoms, shell_atom, shell_r, shell_w, shell_rule, rules, w_cut=1e-15):
    """An atom-centred grid made on the device (wb_grid_build): radial shells
    (atom index, radius, radial weight, angular rule index per shell) times
    their angular rules ((unit vectors (n, 3), weights (n,)) per rule),
    weighted by the cell partition, points of weight <= w_cut dropped.
    Returns host (coords (n, 3), weights (n,))."""
    atoms = lib.f64(atoms).reshape(-1, 3)
    shell_atom = numpy.ascontiguousarray(shell_atom, dtype=numpy.int32)
    shell_rule = numpy.ascontiguousarray(shell_rule, dtype=numpy.int32)
    shell_r, shell_w = lib.f64(shell_r), lib.f64(shell_w)
    npts_rule = numpy.array([len(w) for _, w in rules], dtype=numpy.int32)
    table = lib.f64(numpy.concatenate([numpy.column_stack([u, w]) for u, w in rules]))
    cap = int(npts_rule[shell_rule].sum()) if len(shell_rule) else 0
    coords = numpy.empty((max(cap, 1), 3))
    weights = numpy.empty(max(cap, 1))
    n = ctypes.c_longlong()
    lib.check(lib.load().wb_grid_build(ctx.h, len(atoms), lib.ptr(atoms), len(shell_atom),
                                       lib.ptr(shell_atom), lib.ptr(shell_r), lib.ptr(shell_w),
                                       lib.ptr(shell_rule), len(rules), lib.ptr(npts_rule),
                                       lib.ptr(table), float(w_cut), cap, lib.ptr(coords),
                                       lib.ptr(weights), ctypes.byref(n)))
    n = int(n.value)
    return numpy.ascontiguousarray(coords[:n]), numpy.ascontiguousarray(weights[:n])


def grid_overlap(ctx, basis, grid):
    """S_g = X^T diag(w) X on the device (FP64): the quadrature overlap that
    the overlap-fitted COSX projector is made from (PySCF's fit_ovlp; the
    reference takes the finished projector from PySCF, gpu/sgx_k.py:698-704)."""
    out = numpy.empty((basis.nao, basis.nao))
    lib.check(lib.load().wb_grid_overlap(ctx.h, basis.h, grid.h, lib.ptr(out)))
    return out


def pjs_projector(ctx, basis, grid, S):
    """The overlap-fitted COSX projector Q = S_g^{-1} S on the device: S_g the
    grid overlap, S the analytic overlap (PySCF's int1e_ovlp), LU with partial
    pivoting as PySCF's scipy.linalg.solve(S_g, S) (the matrix PySCF keeps as
    sgx._pjs_data._overlap_correction_matrix and the reference reads at
    gpu/sgx_k.py:698-704).  K built from Q dm integrates the AO products
    against the grid as the analytic overlap does."""
    S = lib.f64(S)
    if S.shape != (basis.nao, basis.nao):
        raise ValueError('S must be (nao, nao)')
    out = numpy.empty((basis.nao, basis.nao))
    lib.check(lib.load().wb_pjs_projector(ctx.h, basis.h, grid.h, lib.ptr(S), lib.ptr(out)))
    return out
