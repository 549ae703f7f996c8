"""Grid construction on the MI355X: the atomic cell partition, and the
quadrature overlap of the AOs on a grid.

Building an atom-centred grid is cheap except for the cell partition, which
weighs every point against every nearby atom pair (O(points x atoms^2) in the
worst case; 97% of a (H2O)27 grid build on the host).  ``ssf_partition``
runs it on the device (``wb_grid_partition``, one thread per point, FP64),
for the Stratmann-Scuseria-Frisch cells the package's grids use
(systems/grids.py).  The reference builds its SGX and XC grids on the host
through PySCF (grids.build(with_non0tab=True), gpu/install.py:164;
with_df.build(level=...), staging/schedule.py:391).
"""
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
