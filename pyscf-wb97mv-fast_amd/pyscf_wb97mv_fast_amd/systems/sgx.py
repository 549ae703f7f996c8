"""Stand-ins for the PySCF objects the hooks talk to (no PySCF offline).

``SyntheticSGX`` carries what ``sgx_jk.get_k_only`` reads from a PySCF SGX
object (``mol``, ``grids``, ``_pjs_data`` with ``_overlap_correction_matrix``
and ``sym_ovlp``, ``_build_pjs``, ``_rsh_df``).  Its PJS matrix is the
identity (Q = I) unless one is given; the production path takes PySCF's.
``SyntheticNumInt.nlc_coeff`` returns omega-B97M-V's VV10 parameters
(b = 6.0, C = 0.01) in PySCF's [((b, C), factor)] form; ``SyntheticRKS``
holds the attributes install_gpu patches.
"""
import numpy


class PJSData:
    def __init__(self, mol, grids, itol, Q, sym_ovlp):
        self.mol = mol
        self.grids = grids
        self._itol = itol
        self._overlap_correction_matrix = Q
        self.sym_ovlp = sym_ovlp


class SyntheticSGX:
    def __init__(self, mol, grids, Q=None, sym_ovlp=False):
        self.mol = mol
        self.grids = grids
        self._Q = Q
        self._sym = sym_ovlp
        self._pjs_data = None
        self._rsh_df = {}
        self.pjs = True

    def _build_pjs(self, direct_scf_tol=1e-13):
        Q = numpy.eye(self.mol.nao) if self._Q is None else self._Q
        self._pjs_data = PJSData(self.mol, self.grids, direct_scf_tol, Q, self._sym)


WB97MV_NLC = ((6.0, 0.01), 1.0)


WB97MV_RSH = (0.3, 1.0, 0.15)      # (omega, alpha, hyb): E_x = hyb K + (alpha - hyb) K_lr


class SyntheticNumInt:
    def rsh_coeff(self, xc_code):
        if 'wb97m' not in str(xc_code).lower().replace('_', '').replace('-', ''):
            return (0.0, 0.0, 0.0)
        return WB97MV_RSH

    def nlc_coeff(self, xc_code):
        if 'wb97m' not in str(xc_code).lower().replace('_', '').replace('-', ''):
            return []
        return [WB97MV_NLC]

    def nr_nlc_vxc(self, mol, grids, xc_code, dm, relativity=0, hermi=1, max_memory=2000,
                   verbose=None):
        raise NotImplementedError('no CPU NumInt offline')

    def nr_rks(self, mol, grids, xc_code, dms, relativity=0, hermi=1, max_memory=2000,
               verbose=None):
        raise NotImplementedError('no CPU NumInt offline')

    def _xc_type(self, xc_code):
        return 'MGGA' if 'wb97m' in str(xc_code).lower().replace('_', '').replace('-', '') \
            else 'GGA'

    def eval_xc(self, xc_code, rho, spin=0, relativity=0, deriv=1, omega=None, verbose=None):
        raise NotImplementedError('no libxc offline: the functional is PySCF\'s')


class SyntheticRKS:
    def __init__(self, mol, sgx_grids, nlc_grids, xc='wb97m-v'):
        self.mol = mol
        self.xc = xc
        self._numint = SyntheticNumInt()
        self.with_df = SyntheticSGX(mol, sgx_grids)
        self.nlcgrids = nlc_grids
