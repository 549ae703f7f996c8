"""Install the MI355X hot path into a PySCF RKS/COSX object.

Same removable hooks as the reference (gpu/install.py:209-336):

    mf._numint.nr_rks            -> xc.gpu_nr_rks            (semilocal XC, semilocal=True)
    mf._numint.nr_nlc_vxc        -> vv10.gpu_nr_nlc_vxc      (VV10)
    pyscf.sgx.sgx_jk.get_k_only  -> sgx_k.get_k_only_gpu     (COSX K, k=True)

The K hook dispatches only for ``mf.with_df`` and the long-range copies in
``mf.with_df._rsh_df`` (gpu/install.py:300-304); every other SGX object
keeps the original.  The operators of one SCF step (full range and the
functional's range-separated ones) are evaluated together at the first
request and the others served from that result (sgx_k.rsh_companions).  Any ``Unsupported`` input or GPU error emits a
RuntimeWarning with the traceback and calls the original PySCF function
(gpu/install.py:91-96); an out-of-memory error first retries with half the
number of G partial copies (down to one), the memory lever of this design, then
falls back (gpu/install.py:99-109).  ``restore_all()`` on the returned
HookSet removes every patch.  semilocal=False leaves the semilocal XC on
PySCF's CPU path (the reference's switch of the same name).
"""
import traceback
import warnings

from .. import lib
from ..core.hooks import HookSet
from ..errors import Unsupported, is_oom
from . import sgx_k, vv10, xc

WARN_PREFIX = 'pyscf_wb97mv_fast_amd GPU'


def _warn_fallback(where, exc):
    warnings.warn('%s %s failed (%s: %s); falling back to the CPU implementation\n%s'
                  % (WARN_PREFIX, where, type(exc).__name__, exc, traceback.format_exc()),
                  RuntimeWarning)


def run_with_oom_retry(flow, set_copies, ncopy0=8):
    """flow() with ncopy = 8, 4, 2, 1 G partial copies on out-of-memory errors
    (set_copies(n) applies n); other errors propagate at once."""
    last = None
    for attempt in range(4):
        set_copies(max(1, ncopy0 >> attempt))
        try:
            return flow()
        except Exception as exc:          # noqa: BLE001
            if not is_oom(exc):
                raise
            last = exc
    raise last


class GpuContext:
    """Everything the two flows need, built lazily once per mf (the device
    context on first use, so installing never touches the GPU)."""

    def __init__(self, mol, device=0, precision='mixed', k_tile_tol=1e-11,
                 vv10_tile_tol=0.0):
        self.mol = mol
        self.device = device
        self.precision = precision
        self.k_tile_tol = k_tile_tol
        self.vv10_tile_tol = vv10_tile_tol
        self._ctx = None
        self._kb = None
        self._nlc = None

    @property
    def ctx(self):
        if self._ctx is None:
            self._ctx = lib.Context(self.device)
        return self._ctx

    def set_copies(self, n):
        lib.check(lib.load().wb_ctx_set_copies(self.ctx.h, int(n)))

    @property
    def k_builder(self):
        if self._kb is None:
            self._kb = sgx_k.GpuKBuilder(self.mol, self.ctx, tile_tol=self.k_tile_tol,
                                         precision=self.precision)
        return self._kb

    @property
    def nlc_state(self):
        if self._nlc is None:
            self._nlc = vv10.NlcState(self.ctx, self.mol)
        return self._nlc


def gpu_available(device=0):
    try:
        lib.Context(device).close()
        return True
    except Exception:                     # noqa: BLE001
        return False


def install_gpu(mf, backend=None, precision='mixed', vv10_tile_tol=0.0, k=False,
                k_tile_tol=1e-11, device=0, sgx_module=None, semilocal=True):
    """Hook mf._numint.nr_rks (semilocal=True) and nr_nlc_vxc (and, with
    k=True, sgx_jk.get_k_only); returns a HookSet.  backend=None installs
    nothing (with a warning) when the library or the GPU is unavailable;
    backend='hip' insists."""
    ni = getattr(mf, '_numint', None)
    if ni is None:
        raise TypeError('mf has no _numint; not a DFT object?')
    if backend is None and not gpu_available(device):
        warnings.warn('MI355X GPU path not available; mf stays on the CPU', RuntimeWarning)
        return HookSet()
    ctx = GpuContext(mf.mol, device, precision, k_tile_tol, vv10_tile_tol)
    hooks = HookSet()
    hooks.k_stats = []
    hooks.ctx = ctx

    def nlc_factory(orig):
        def nr_nlc_vxc(mol, grids, xc_code, dm, relativity=0, hermi=1, max_memory=2000,
                       verbose=None):
            try:
                if mol is not ctx.mol:
                    raise Unsupported('molecule changed since install_gpu')
                return run_with_oom_retry(
                    lambda: vv10.gpu_nr_nlc_vxc(ctx.nlc_state, ni, mol, grids, xc_code, dm,
                                                relativity, hermi, max_memory, verbose,
                                                tile_tol=ctx.vv10_tile_tol,
                                                precision=ctx.precision), ctx.set_copies)
            except Exception as exc:      # noqa: BLE001
                _warn_fallback('nr_nlc_vxc', exc)
                return orig(mol, grids, xc_code, dm, relativity, hermi, max_memory, verbose)
        return nr_nlc_vxc

    def rks_factory(orig):
        def nr_rks(mol, grids, xc_code, dms, relativity=0, hermi=1, max_memory=2000,
                   verbose=None):
            try:
                if mol is not ctx.mol:
                    raise Unsupported('molecule changed since install_gpu')
                return xc.gpu_nr_rks(ctx.nlc_state, ni, mol, grids, xc_code, dms, relativity,
                                     hermi, max_memory, verbose)
            except Exception as exc:      # noqa: BLE001
                _warn_fallback('nr_rks', exc)
                return orig(mol, grids, xc_code, dms, relativity, hermi, max_memory, verbose)
        return nr_rks

    if semilocal:
        hooks.wrap(ni, 'nr_rks', rks_factory)
    hooks.wrap(ni, 'nr_nlc_vxc', nlc_factory)

    if k:
        if sgx_module is None:
            import pyscf.sgx.sgx_jk as sgx_module       # noqa: F811

        def k_factory(orig):
            def get_k_only(sgx, dm, hermi=1, direct_scf_tol=1e-13):
                df = getattr(mf, 'with_df', None)
                rsh = getattr(df, '_rsh_df', None) or {}
                if not (sgx is df or any(sgx is v for v in rsh.values() if v is not None)):
                    return orig(sgx, dm, hermi, direct_scf_tol)
                try:
                    if sgx.mol is not ctx.mol:
                        raise Unsupported('molecule changed since install_gpu')
                    comp = sgx_k.rsh_companions(mf)
                    kk = run_with_oom_retry(
                        lambda: sgx_k.get_k_only_gpu(ctx.k_builder, sgx, dm, hermi,
                                                     direct_scf_tol, companions=comp),
                        ctx.set_copies)
                except Exception as exc:  # noqa: BLE001
                    _warn_fallback('get_k_only', exc)
                    return orig(sgx, dm, hermi, direct_scf_tol)
                hooks.k_stats.append(dict(ctx.k_builder.last_stats))
                return kk
            return get_k_only

        hooks.wrap(sgx_module, 'get_k_only', k_factory)
    return hooks
