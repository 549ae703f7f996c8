"""Stand-ins for the PySCF objects the hooks talk to (no PySCF offline).

``SyntheticSGX`` carries what ``sgx_jk.get_k_only`` reads from a PySCF SGX
object (``mol``, ``grids``, ``_pjs_data`` with ``_overlap_correction_matrix``
and ``sym_ovlp``, ``_build_pjs``, ``_rsh_df``) and PySCF 2.14's ``get_jk``,
including the range-separated branch as released (a shallow per-omega copy
called positionally without ``vhfopt``; core.sgx_patch repairs it).  Its PJS
matrix is the identity (Q = I) unless one is given; the production path
takes PySCF's.  J comes from ``j_builder(sgx, dm)`` (none offline: tests
give one).  ``sgx_jk`` stands for the module ``pyscf.sgx.sgx_jk`` (its
``get_k_only`` is what the K hook replaces).

``SyntheticNumInt.nlc_coeff`` returns omega-B97M-V's VV10 parameters
(b = 6.0, C = 0.01) in PySCF's [((b, C), factor)] form and ``rsh_coeff`` its
(omega, alpha, hyb) = (0.3, 1.0, 0.15); ``SyntheticRKS`` holds the attributes
install_gpu patches and a ``get_veff`` that combines the terms as PySCF's
``rks.get_veff`` does for an RSH functional with COSX.
"""
import copy
import types

import numpy


class PJSData:
    def __init__(self, mol, grids, itol, Q, sym_ovlp):
        self.mol = mol
        self.grids = grids
        self._itol = itol
        self._overlap_correction_matrix = Q
        self.sym_ovlp = sym_ovlp


def _no_get_k_only(sgx, dm, hermi=1, direct_scf_tol=1e-13):
    raise NotImplementedError('no CPU SGX offline: the K hook (or a test) provides K')


# stands for the module pyscf.sgx.sgx_jk
sgx_jk = types.SimpleNamespace(get_k_only=_no_get_k_only)


class SyntheticSGX:
    """Q: a fixed PJS projector; fit_ovlp=ctx: the overlap-fitted projector
    S_g^-1 S built on that device context at _build_pjs (gpu.grids.pjs_projector,
    as PySCF builds it); neither: the identity."""

    def __init__(self, mol, grids, Q=None, sym_ovlp=False, j_builder=None, fit_ovlp=None):
        self.mol = mol
        self.grids = grids
        self._Q = Q
        self.fit_ovlp = fit_ovlp
        self._sym = sym_ovlp
        self._pjs_data = None
        self._rsh_df = {}
        self._vjopt = None
        self._overlap_correction_matrix = None
        self.pjs = True
        self.j_builder = j_builder
        self.j_calls = 0

    def _build_pjs(self, direct_scf_tol=1e-13):
        Q = numpy.eye(self.mol.nao) if self._Q is None else self._Q
        if self._Q is None and self.fit_ovlp is not None:
            from .. import lib
            from ..basis import pack_mol
            from ..gpu.grids import pjs_projector
            from .guess import overlap_matrix
            ctx = self.fit_ovlp
            Q = pjs_projector(ctx, lib.DeviceBasis(ctx, pack_mol(self.mol)),
                              lib.DeviceGrid(ctx, self.grids.coords, self.grids.weights),
                              overlap_matrix(self.mol))
        self._pjs_data = PJSData(self.mol, self.grids, direct_scf_tol, Q, self._sym)

    def copy(self):
        """A shallow copy, as PySCF's; the copy counts its own J builds."""
        twin = copy.copy(self)
        twin.j_calls = 0
        return twin

    def get_jk(self, dm, hermi=1, vhfopt=None, with_j=True, with_k=True,
               direct_scf_tol=1e-13, omega=None):
        """PySCF 2.14's SGX.get_jk, its range-separated branch as released."""
        if omega is not None:
            key = '%.6f' % omega
            rsh_df = self._rsh_df.get(key)
            if rsh_df is None:
                rsh_df = self.copy()          # shallow: shares _pjs_data
                rsh_df._rsh_df = None
                self._rsh_df[key] = rsh_df
            with rsh_df.mol.with_range_coulomb(omega):
                # the released call: vhfopt missing, arguments one slot off
                return rsh_df.get_jk(dm, hermi, with_j, with_k, direct_scf_tol)
        vj = vk = None
        if with_j:
            if self.j_builder is None:
                raise NotImplementedError('no J builder offline')
            self.j_calls += 1
            vj = self.j_builder(self, dm)
        if with_k:
            vk = sgx_jk.get_k_only(self, dm, hermi, direct_scf_tol)
        return vj, vk


WB97MV_NLC = ((6.0, 0.01), 1.0)


WB97MV_RSH = (0.3, 1.0, 0.15)      # (omega, alpha, hyb): E_x = hyb K + (alpha - hyb) K_lr


def _is_wb97mv(xc_code):
    return 'wb97m' in str(xc_code).lower().replace('_', '').replace('-', '')


class SyntheticNumInt:
    libxc = types.SimpleNamespace(is_nlc=_is_wb97mv)

    def rsh_coeff(self, xc_code):
        if not _is_wb97mv(xc_code):
            return (0.0, 0.0, 0.0)
        return WB97MV_RSH

    def nlc_coeff(self, xc_code):
        if not _is_wb97mv(xc_code):
            return []
        return [WB97MV_NLC]

    def nr_nlc_vxc(self, mol, grids, xc_code, dm, relativity=0, hermi=1, max_memory=2000,
                   verbose=None):
        raise NotImplementedError('no CPU NumInt offline')

    def nr_rks(self, mol, grids, xc_code, dms, relativity=0, hermi=1, max_memory=2000,
               verbose=None):
        raise NotImplementedError('no CPU NumInt offline')

    def _xc_type(self, xc_code):
        return 'MGGA' if _is_wb97mv(xc_code) else 'GGA'

    def eval_xc(self, xc_code, rho, spin=0, relativity=0, deriv=1, omega=None, verbose=None):
        raise NotImplementedError('no libxc offline: the functional is PySCF\'s')


class SyntheticRKS:
    def __init__(self, mol, sgx_grids, nlc_grids, xc='wb97m-v', xc_grids=None):
        self.mol = mol
        self.xc = xc
        self.nlc = ''
        self._numint = SyntheticNumInt()
        self.with_df = SyntheticSGX(mol, sgx_grids)
        self.nlcgrids = nlc_grids
        self.grids = xc_grids if xc_grids is not None else nlc_grids

    def do_nlc(self):
        return bool(self._numint.nlc_coeff(self.xc))

    def initialize_grids(self, mol=None, dm=None):
        return self

    def get_veff(self, mol=None, dm=None, dm_last=None, vhf_last=None, hermi=1):
        """rks.get_veff (PySCF 2.14) for an RSH functional with COSX: XC and
        VV10 from _numint, J and the full-range and long-range K from
        with_df.get_jk, vxc + vj - vk / 2 tagged with ecoul and exc."""
        from ..gpu.install import tag_array
        mol = self.mol if mol is None else mol
        ni = self._numint
        _, exc, vxc = ni.nr_rks(mol, self.grids, self.xc, dm)
        if self.do_nlc():
            _, enl, vnl = ni.nr_nlc_vxc(mol, self.nlcgrids, self.xc, dm)
            exc, vxc = exc + enl, vxc + vnl
        omega, alpha, hyb = ni.rsh_coeff(self.xc)
        vj, vk = self.with_df.get_jk(dm, hermi)
        vk = vk * hyb
        if omega:
            vk = vk + self.with_df.get_jk(dm, hermi, with_j=False, omega=omega)[1] * (alpha - hyb)
        exc -= numpy.einsum('ij,ji->', dm, vk).real * 0.25
        ecoul = numpy.einsum('ij,ji->', dm, vj).real * 0.5
        return tag_array(vxc + vj - vk * 0.5, ecoul=ecoul, exc=exc, vj=vj, vk=vk)
