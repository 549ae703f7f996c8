"""COSX exchange K on the MI355X: drop-in for pyscf.sgx.sgx_jk.get_k_only.

``GpuKBuilder.get_k`` keeps the contract of the reference's builder
(gpu/sgx_k.py:663-751): one symmetric RKS density matrix, hermi = 1, the
grid and the PJS matrices taken from the SGX object exactly as
``sgx_jk.get_k_only`` sets them up (Q = ``_overlap_correction_matrix``,
``sym_ovlp``), and its post-processing K = Q^T K if sym_ovlp, then
(K + K^T)/2.  Everything between QD = Q dm and the raw K runs in one call
of the C ABI on the device (wb_k_build):

    X = AO(grid)  FP64 -> F = X (QD)  FP64 MFMA GEMM, rounded once to FP32
    -> screening per (64-point sub-block, shell pair)
    -> fused McMurchie-Davidson kernel per (la, lb) class: G = A F
       (FP32 in the mixed form, FP64 in the FP64 form; A never stored)
    -> K = X^T (w o G)  FP64 MFMA GEMM

Device state (basis / shell pairs and, per grids object, the sorted grid)
is cached; a rebuilt grids object gets a fresh state (stale-grid rule of
gpu/sgx_k.py:856-864).  ``get_k_only_gpu`` reads omega from sgx.mol.omega
(gpu/sgx_k.py:983-991).

Companion operators: an RSH functional asks for K(1/r) and K(erf(w r)/r)
of the same density on the same grid in two consecutive get_k_only calls.
Given the operators to expect (``companions``), the first call evaluates
all of them in one device call -- the exchange kernels of the light
classes then run both operators in one pass over the tasks -- and keeps
the raw K of the others; a later call with the same grid and the same
Q dm (compared exactly) returns the kept one.
"""
import time

import numpy

from .. import lib
from ..basis import pack_mol
from ..errors import Unsupported

_STATE_CACHE = 4


class GpuKBuilder:
    def __init__(self, mol, ctx=None, tile_tol=1e-11, precision='mixed', pair_cut=1e-14,
                 device=0):
        if precision not in lib.PREC:
            raise ValueError('precision must be one of %s' % sorted(lib.PREC))
        self.mol = mol
        self.tile_tol = float(tile_tol)
        self.precision = precision
        self.pair_cut = float(pair_cut)
        self._ctx = ctx
        self._device = device
        self._basis = None
        self._grids = {}
        self._kcache = None
        self.last_stats = {}

    # -- device state ---------------------------------------------------------
    @property
    def ctx(self):
        if self._ctx is None:
            self._ctx = lib.Context(self._device)
        return self._ctx

    def _build(self):
        if self._basis is None:
            self._basis = lib.DeviceBasis(self.ctx, pack_mol(self.mol), self.pair_cut)
        return self._basis

    @property
    def nao(self):
        return self._build().nao

    def grid_state(self, grids):
        key = (id(grids), int(numpy.asarray(grids.weights).size))
        st = self._grids.get(key)
        if st is None or st[0] is not grids:
            st = (grids, lib.DeviceGrid(self.ctx, grids.coords, grids.weights))
            self._grids[key] = st
            while len(self._grids) > _STATE_CACHE:
                self._grids.pop(next(iter(self._grids)))
        return st[1]

    # -- K ----------------------------------------------------------------------
    def k_raw(self, grids, qd, omegas):
        """Raw K (nop, nao, nao) = X^T (w o A_op (X qd)) for each omega."""
        basis = self._build()
        qd = lib.f64(qd)
        nao = basis.nao
        if qd.shape != (nao, nao):
            raise Unsupported('density matrix does not match the basis')
        grid = self.grid_state(grids)
        om = lib.f64(numpy.atleast_1d(omegas))
        out = numpy.empty((len(om), nao, nao))
        t0 = time.perf_counter()
        lib.check(lib.load().wb_k_build(self.ctx.h, basis.h, grid.h, lib.ptr(qd), len(om),
                                        lib.ptr(om), self.tile_tol, lib.PREC[self.precision],
                                        lib.ptr(out)))
        st = numpy.zeros(2, dtype=numpy.int64)
        lib.check(lib.load().wb_k_stats(self.ctx.h, grid.h, lib.ptr(st)))
        self.last_stats = {'pairs_total': int(st[0]), 'pairs_kept': int(st[1]),
                           'wall': time.perf_counter() - t0, 'kernel': self.ctx.last_ms * 1e-3}
        return out

    def k_raw_shared(self, grids, qd, omega, companions=()):
        """Raw K for omega; with companions, all of them in one device call,
        the others kept for a later request with the same grids and qd."""
        c = self._kcache
        if (c is not None and c['grids'] is grids and omega in c['K']
                and c['qd'].shape == qd.shape and numpy.array_equal(c['qd'], qd)):
            self.last_stats = dict(c['stats'], cached=True)
            return c['K'].pop(omega)
        oms = [float(omega)] + [float(w) for w in companions if float(w) != float(omega)]
        Ks = self.k_raw(grids, qd, oms)
        self._kcache = {'grids': grids, 'qd': numpy.array(qd, copy=True),
                        'K': dict(zip(oms[1:], Ks[1:])), 'stats': dict(self.last_stats)}
        return Ks[0]

    def get_k(self, sgx, dm, omega, hermi=1, direct_scf_tol=1e-13, companions=()):
        """K (nao, nao) FP64 for one RKS density matrix on sgx's grid."""
        if hermi != 1:
            raise Unsupported('GPU COSX K supports hermi=1 only')
        dm = numpy.asarray(dm)
        if dm.ndim == 3 and dm.shape[0] == 1:
            dm = dm[0]
        if dm.ndim != 2 or dm.shape[0] != dm.shape[1]:
            raise Unsupported('GPU COSX K supports a single square density matrix')
        if not numpy.allclose(dm, dm.T, rtol=0, atol=1e-12):
            raise Unsupported('GPU COSX K needs a symmetric density matrix')
        mol, grids = sgx.mol, sgx.grids
        pd = getattr(sgx, '_pjs_data', None)
        if (pd is None or pd.mol is not mol or pd.grids is not grids
                or getattr(pd, '_itol', direct_scf_tol) != direct_scf_tol):
            sgx._build_pjs(direct_scf_tol)
            pd = sgx._pjs_data
        proj = numpy.asarray(pd._overlap_correction_matrix, dtype=numpy.float64)
        sym = bool(pd.sym_ovlp)
        K = self.k_raw_shared(grids, proj @ dm, omega, companions)
        if sym:
            K = proj.T @ K
        if hermi == 1:
            K = (K + K.T) * 0.5
        return K


def get_k_only_gpu(builder, sgx, dm, hermi=1, direct_scf_tol=1e-13, companions=()):
    """pyscf.sgx.sgx_jk.get_k_only's signature and contract for one RKS dm;
    omega from sgx.mol.omega (set by mol.with_range_coulomb)."""
    omega = float(getattr(sgx.mol, 'omega', 0.0) or 0.0)
    if omega < 0:
        raise Unsupported('short-range (erfc, omega < 0) exchange is not on the GPU path')
    return builder.get_k(sgx, dm, omega, hermi=hermi, direct_scf_tol=direct_scf_tol,
                         companions=companions)


def rsh_companions(mf):
    """The exchange operators an SCF step of mf asks for: the full-range
    operator of mf.with_df, the range-separated copies it holds
    (_rsh_df keys, '%.6f' % omega as core/sgx_patch.py:39-48 makes them)
    and, before those exist, the omega of mf's functional; the GPU path
    evaluates 1/r and erf(omega r)/r, omega > 0, only."""
    oms = {0.0}
    df = getattr(mf, 'with_df', None)
    for key in (getattr(df, '_rsh_df', None) or {}):
        try:
            omega = float(key)
        except (TypeError, ValueError):
            continue
        if omega > 0:                     # erfc operators (omega < 0) are not evaluated here
            oms.add(omega)
    ni = getattr(mf, '_numint', None)
    try:
        omega = float(ni.rsh_coeff(mf.xc)[0])
        if omega > 0:
            oms.add(omega)
    except Exception:                     # noqa: BLE001 -- no RSH information
        pass
    return tuple(sorted(oms))
