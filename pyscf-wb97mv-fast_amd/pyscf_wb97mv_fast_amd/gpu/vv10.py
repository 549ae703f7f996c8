"""VV10 on the MI355X: drop-in for NumInt.nr_nlc_vxc (RKS).

``gpu_nr_nlc_vxc`` keeps the signature and return convention
``(nelec, excsum, vmat)`` of ``dft.numint.NumInt.nr_nlc_vxc`` as the
reference mirrors it (gpu/vv10.py:268-342), restating ``numint._vv10nlc``:
rho >= 1e-8 threshold, per-point W0 / kappa / derivatives in FP64, the
O(N^2) kernel sums, exc = Beta + F/2, v0 = Beta + F + 1.5 (U dK/dR +
W dW0/dR), v1 = 1.5 W dW0/dG, the GGA potential and the Fock build.  The
whole flow is one call of the C ABI (wb_nlc_vxc): AO values and gradients,
density and the Fock GEMMs in FP64 on MFMA, the pair sums in FP32 (mixed)
or FP64.

``Vv10Kernel.kernel_sum`` is the reference's kernel interface
(gpu/vv10.py:167-240): F (already x -1.5), U, W for given points.
"""
import numpy

from .. import lib
from ..basis import pack_mol
from ..errors import Unsupported

THRESH = 1e-8


class Vv10Kernel:
    def __init__(self, ctx, tile_tol=0.0, precision='mixed'):
        self.ctx = ctx
        self.tile_tol = float(tile_tol)
        self.precision = precision

    def kernel_sum(self, coords, q, W0, kappa):
        coords = lib.f64(coords).reshape(-1, 3)
        n = len(coords)
        q, W0, kappa = (lib.f64(a).ravel() for a in (q, W0, kappa))
        F, U, W = numpy.empty(n), numpy.empty(n), numpy.empty(n)
        if n:
            lib.check(lib.load().wb_vv10_sum(self.ctx.h, n, lib.ptr(coords), lib.ptr(q),
                                             lib.ptr(W0), lib.ptr(kappa), self.tile_tol,
                                             lib.PREC[self.precision], lib.ptr(F), lib.ptr(U),
                                             lib.ptr(W)))
        return F, U, W


def _prepare_dm(dm, hermi):
    dm = numpy.asarray(dm, dtype=numpy.float64)
    if dm.ndim == 3 and dm.shape[0] == 1:
        dm = dm[0]
    if dm.ndim != 2 or dm.shape[0] != dm.shape[1]:
        raise Unsupported('GPU VV10 supports a single RKS density matrix')
    if hermi != 1:
        dm = (dm + dm.T) * 0.5
    return numpy.ascontiguousarray(dm)


class NlcState:
    """Per-mol device basis and per-grids device grid for the VV10 flow."""

    def __init__(self, ctx, mol, pair_cut=1e-14):
        self.ctx = ctx
        self.mol = mol
        self.basis = lib.DeviceBasis(ctx, pack_mol(mol), pair_cut)
        self._grids = {}

    def grid(self, grids):
        key = (id(grids), int(numpy.asarray(grids.weights).size))
        st = self._grids.get(key)
        if st is None or st[0] is not grids:
            st = (grids, lib.DeviceGrid(self.ctx, grids.coords, grids.weights))
            self._grids[key] = st
            while len(self._grids) > 4:
                self._grids.pop(next(iter(self._grids)))
        return st[1]


def gpu_nr_nlc_vxc(state, ni, mol, grids, xc_code, dm, relativity=0, hermi=1,
                   max_memory=2000, verbose=None, tile_tol=0.0, precision='mixed'):
    """(nelec, excsum, vmat) of the VV10 term; ``ni.nlc_coeff(xc_code)``
    supplies [((b, C), fac), ...] as in PySCF."""
    if relativity not in (0, None):
        raise Unsupported('relativistic NLC not supported')
    if mol is not state.mol:
        raise Unsupported('NLC state belongs to another molecule')
    dm = _prepare_dm(dm, hermi)
    if dm.shape[0] != state.basis.nao:
        raise Unsupported('density matrix does not match the basis')
    nlc_coefs = ni.nlc_coeff(xc_code)
    if not nlc_coefs:
        raise Unsupported('%r has no NLC parameter set' % (xc_code,))
    grid = state.grid(grids)
    nelec, excsum = 0.0, 0.0
    vmat = numpy.zeros_like(dm)
    for nlc_pars, fac in nlc_coefs:
        b, C = (float(v) for v in nlc_pars)
        out2 = numpy.zeros(2)
        vm = numpy.empty_like(dm)
        lib.check(lib.load().wb_nlc_vxc(state.ctx.h, state.basis.h, grid.h, lib.ptr(dm), b, C,
                                        float(tile_tol), lib.PREC[precision], lib.ptr(out2),
                                        lib.ptr(vm)))
        nelec = out2[0]
        excsum += fac * out2[1]
        vmat += fac * vm
    return nelec, excsum, vmat
