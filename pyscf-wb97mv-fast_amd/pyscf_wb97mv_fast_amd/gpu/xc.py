"""Semilocal XC on the MI355X: drop-in for NumInt.nr_rks (RKS ground state).

``gpu_nr_rks`` keeps the signature and return convention ``(nelec, excsum,
vmat)`` of ``dft.numint.NumInt.nr_rks`` as the reference mirrors it
(gpu/xc.py:130-201), for LDA, GGA and meta-GGA without the laplacian (the
wB97M-V case):

    device (wb_xc_rho):   AO values (and gradients for meta-GGA) -> T = D X on
                          FP64 MFMA -> rho, grad rho (gradients on the fly),
                          tau = 1/2 sum_d (D dX_d) . dX_d
    host:                 the functional ``ni.eval_xc(xc_code, rho, spin=0,
                          deriv=1)`` (PySCF's libxc, as in the reference) and
                          the chain rule v = (v_rho, 2 v_sigma grad rho, v_tau)
    device (wb_xc_vmat):  vmat = sum_g w [v_rho phi phi^T + v_d (d phi phi^T +
                          phi d phi^T) + 1/2 v_tau grad phi . grad phi^T]

nelec and excsum are summed in FP64 on the host in point order.  The
reference runs the density and Fock GEMMs in FP32 (gpu_nr_rks, gemm_dtype);
here everything is FP64.
"""
import numpy

from .. import lib
from ..errors import Unsupported

SUPPORTED_XCTYPE = ('LDA', 'GGA', 'MGGA')


def _prepare_dm(dms, hermi):
    dm = numpy.asarray(dms, dtype=numpy.float64)
    if dm.ndim == 3 and dm.shape[0] == 1:
        dm = dm[0]
    if dm.ndim != 2 or dm.shape[0] != dm.shape[1]:
        raise Unsupported('GPU XC path supports a single RKS density matrix')
    if hermi != 1:
        dm = (dm + dm.T) * 0.5        # what _gen_rho_evaluator does first
    return numpy.ascontiguousarray(dm)


def chain_rule(xctype, rho, vrho, vsigma=None, vtau=None):
    """(5, n) chain-ruled potential from libxc's (vrho, vsigma, vtau):
    v0 = vrho, v1..3 = 2 vsigma grad rho, v4 = vtau (the transform of
    PySCF's xc_deriv.transform_vxc for spin 0, as the reference's
    xc_flow._chain_rule, gpu/xc.py:14-17)."""
    n = rho.shape[-1]
    v = numpy.zeros((5, n))
    v[0] = vrho
    if xctype in ('GGA', 'MGGA'):
        v[1:4] = 2.0 * vsigma * rho[1:4]
    if xctype == 'MGGA':
        v[4] = vtau
    return v


def xc_rho(state, grids, dm, mgga):
    """(5, n) rho, d rho/dx, /dy, /dz, tau on grids (tau = 0 unless mgga)."""
    grid = state.grid(grids)
    rho = numpy.empty((5, grid.npts))
    lib.check(lib.load().wb_xc_rho(state.ctx.h, state.basis.h, grid.h, lib.ptr(dm),
                                   1 if mgga else 0, lib.ptr(rho)))
    return rho


def xc_vmat(state, grids, v, mgga):
    """vmat of the chain-ruled potential v (5, n) on grids."""
    grid = state.grid(grids)
    nao = state.basis.nao
    v = numpy.ascontiguousarray(v, dtype=numpy.float64)
    if v.shape != (5, grid.npts):
        raise ValueError('potential shape (5, %d) expected' % grid.npts)
    vmat = numpy.empty((nao, nao))
    lib.check(lib.load().wb_xc_vmat(state.ctx.h, state.basis.h, grid.h, lib.ptr(v),
                                    1 if mgga else 0, lib.ptr(vmat)))
    return vmat


def gpu_nr_rks(state, ni, mol, grids, xc_code, dms, relativity=0, hermi=1,
               max_memory=2000, verbose=None):
    """(nelec, excsum, vmat) of the semilocal XC term (``state``: the
    per-mol device basis and grids, vv10.NlcState)."""
    xctype = ni._xc_type(xc_code)
    if xctype not in SUPPORTED_XCTYPE:
        raise Unsupported('xctype %r not supported on the GPU path' % xctype)
    if relativity not in (0, None):
        raise Unsupported('relativistic XC not supported')
    dm = _prepare_dm(dms, hermi)
    if dm.shape[0] != state.basis.nao:
        raise Unsupported('density matrix does not match the basis')
    mgga = xctype == 'MGGA'
    rho = xc_rho(state, grids, dm, mgga)
    arg = rho[0] if xctype == 'LDA' else (rho[:4] if xctype == 'GGA' else rho)
    exc, vxc = ni.eval_xc(xc_code, arg, spin=0, deriv=1)[:2]
    vsigma = vxc[1] if len(vxc) > 1 else None
    vtau = vxc[3] if len(vxc) > 3 else None
    v = chain_rule(xctype, rho, vxc[0], vsigma, vtau)
    w = numpy.asarray(grids.weights, dtype=numpy.float64)
    den = rho[0] * w
    nelec = float(den.sum())
    excsum = float(numpy.dot(den, numpy.asarray(exc, dtype=numpy.float64).ravel()))
    vmat = xc_vmat(state, grids, v, mgga)
    return nelec, excsum, vmat
