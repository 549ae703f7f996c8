"""Install the MI355X hot path into a PySCF RKS/COSX object.

Same removable hooks as the reference (gpu/install.py:209-336):

    mf._numint.nr_rks            -> xc.gpu_nr_rks            (semilocal XC, semilocal=True)
    mf._numint.nr_nlc_vxc        -> vv10.gpu_nr_nlc_vxc      (VV10)
    pyscf.sgx.sgx_jk.get_k_only  -> sgx_k.get_k_only_gpu     (COSX K, k=True)
    mf.get_veff                  -> XC + VV10 on a worker thread while this
                                    thread builds J and K   (overlap=True)
    pyscf.sgx.sgx.SGX.get_jk     -> core.sgx_patch.rsh_get_jk (rsh_fix=True)

The K hook dispatches only for ``mf.with_df`` and the long-range copies in
``mf.with_df._rsh_df`` (gpu/install.py:300-304); every other SGX object
keeps the original.  The operators of one SCF step (full range and the
functional's range-separated ones) are evaluated together at the first
request and the others served from that result (sgx_k.rsh_companions).

The K build and the XC / VV10 flows run on two device contexts (each its own
HIP stream and workspaces), so the overlapped get_veff -- K on the calling
thread, XC and VV10 on the worker, as the reference runs them on two
streams (gpu/install.py:118-135, :171-196) -- never shares a stream or a
buffer between the two threads.

Any ``Unsupported`` input or GPU error emits a RuntimeWarning with the
traceback and calls the original PySCF function (gpu/install.py:91-96).  An
out-of-memory K build first retries with half the G partial copies of the K
context (down to one) and keeps the count that worked -- the memory lever of
this design, as the reference halves its tile (gpu/install.py:99-109); the
XC and VV10 flows hold no such lever and fall back at once.
``restore_all()`` on the returned HookSet removes every patch.
"""
import threading
import time
import traceback
import warnings

import numpy

from .. import lib
from ..core import sgx_patch
from ..core.hooks import HookSet
from ..errors import Unsupported, is_oom
from . import sgx_k, vv10, xc

WARN_PREFIX = 'pyscf_wb97mv_fast_amd GPU'


def _warn_fallback(where, exc):
    warnings.warn('%s %s failed (%s: %s); falling back to the CPU implementation\n%s'
                  % (WARN_PREFIX, where, type(exc).__name__, exc, traceback.format_exc()),
                  RuntimeWarning)


def run_with_oom_retry(flow, get_copies, set_copies):
    """flow(); on an out-of-memory error halve the G copy count (read by
    get_copies(), applied by set_copies(n)) and try again, down to one copy.
    The count that succeeded stays set (the next call starts there); other
    errors propagate at once."""
    while True:
        try:
            return flow()
        except Exception as exc:          # noqa: BLE001
            if not is_oom(exc):
                raise
            n = get_copies()
            if n <= 1:
                raise
            set_copies(max(1, n // 2))


class GpuContext:
    """Everything the flows need, built lazily once per mf (the device
    contexts on first use, so installing never touches the GPU).  kctx runs
    the K builds, vctx the XC and VV10 flows."""

    def __init__(self, mol, device=0, precision='mixed', k_tile_tol=1e-11,
                 vv10_tile_tol=0.0):
        self.mol = mol
        self.device = device
        self.precision = precision
        self.k_tile_tol = k_tile_tol
        self.vv10_tile_tol = vv10_tile_tol
        self._vctx = None
        self._kb = None
        self._nlc = None

    @property
    def kctx(self):
        return self.k_builder.ctx

    @property
    def vctx(self):
        if self._vctx is None:
            self._vctx = lib.Context(self.device)
        return self._vctx

    # the K context's G copy count (the OOM lever)
    def get_copies(self):
        return self.kctx.tuning()[0]

    def set_copies(self, n):
        lib.check(lib.load().wb_ctx_set_copies(self.kctx.h, int(n)))

    @property
    def k_builder(self):
        if self._kb is None:
            # the builder makes the K context at its first device call, so
            # input checks run (and fail) before anything touches the GPU
            self._kb = sgx_k.GpuKBuilder(self.mol, None, tile_tol=self.k_tile_tol,
                                         precision=self.precision, device=self.device)
        return self._kb

    @property
    def nlc_state(self):
        if self._nlc is None:
            self._nlc = vv10.NlcState(self.vctx, self.mol)
        return self._nlc


def gpu_available(device=0):
    try:
        lib.Context(device).close()
        return True
    except Exception:                     # noqa: BLE001
        return False


def tag_array(a, **attrs):
    """PySCF's lib.tag_array (an ndarray view carrying attributes); a
    minimal equivalent when PySCF is not importable."""
    try:
        from pyscf.lib import tag_array as pyscf_tag
    except ImportError:
        pyscf_tag = None
    if pyscf_tag is not None:
        return pyscf_tag(a, **attrs)
    t = numpy.asarray(a).view(_Tagged)
    t.__dict__.update(attrs)
    return t


class _Tagged(numpy.ndarray):
    pass


def _zero_xc(*args, **kwargs):
    dm = args[3] if len(args) > 3 else kwargs.get('dm', kwargs.get('dms'))
    n = numpy.asarray(dm).shape[-1]
    return 0.0, 0.0, numpy.zeros((n, n))


def overlap_get_veff(mf, ni, timings):
    """Factory of an mf.get_veff that runs the XC and VV10 terms (the hooked
    nr_rks / nr_nlc_vxc, on the XC context) on a worker thread while the
    calling thread runs PySCF's own get_veff -- J, the COSX K builds through
    the K hook, on the K context -- with nr_rks / nr_nlc_vxc answering zero;
    the worker's exc and vxc are added afterwards (the reference's
    _overlap_factory, gpu/install.py:138-206).  rks.get_veff uses the XC
    results only in sums with J and K, so the result is the serial one with
    the terms added in another order.  ctypes releases the GIL around every
    device call, so the two threads' device work overlaps.  Each call
    appends {'jk', 'xc', 'wait'} wall seconds to ``timings``."""
    def factory(orig):
        def get_veff(mol=None, dm=None, dm_last=None, vhf_last=None, hermi=1):
            if mol is None:
                mol = mf.mol
            if dm is None:
                dm = mf.make_rdm1()
            if hermi == 2 or not (isinstance(dm, numpy.ndarray) and dm.ndim == 2):
                return orig(mol, dm, dm_last, vhf_last, hermi)
            if mf.grids.coords is None:
                mf.initialize_grids(mol, dm)
            do_nlc = mf.do_nlc()
            if do_nlc and mf.nlcgrids.coords is None:
                mf.nlcgrids.build(with_non0tab=True)
            nlc_code = None
            if do_nlc:
                nlc_code = mf.xc if ni.libxc.is_nlc(mf.xc) else mf.nlc
            nr_rks, nr_nlc_vxc = ni.nr_rks, ni.nr_nlc_vxc
            box = {}

            def xc_terms():
                t = time.perf_counter()
                try:
                    _, exc, vxc = nr_rks(mol, mf.grids, mf.xc, dm)
                    if do_nlc:
                        _, enl, vnl = nr_nlc_vxc(mol, mf.nlcgrids, nlc_code, dm)
                        exc, vxc = exc + enl, vxc + vnl
                    box['xc'] = (exc, vxc)
                except BaseException as err:          # re-raised on the caller's thread
                    box['err'] = err
                box['t'] = time.perf_counter() - t

            worker = threading.Thread(target=xc_terms, name='wb97mv-amd-xc')
            worker.start()
            ni.nr_rks, ni.nr_nlc_vxc = _zero_xc, _zero_xc
            t0 = time.perf_counter()
            try:
                veff = orig(mol, dm, dm_last, vhf_last, hermi)
            finally:
                ni.nr_rks, ni.nr_nlc_vxc = nr_rks, nr_nlc_vxc
                t1 = time.perf_counter()
                worker.join()
            timings.append({'jk': t1 - t0, 'xc': box.get('t'),
                            'wait': time.perf_counter() - t1})
            if 'err' in box:
                raise box['err']
            exc, vxc = box['xc']
            return tag_array(numpy.asarray(veff) + vxc, ecoul=veff.ecoul, exc=veff.exc + exc,
                             vj=veff.vj, vk=veff.vk)
        return get_veff
    return factory


def install_gpu(mf, backend=None, precision='mixed', vv10_tile_tol=0.0, k=False,
                k_tile_tol=1e-11, device=0, sgx_module=None, semilocal=True, overlap=True,
                rsh_fix=False, sgx_class=None):
    """Hook mf._numint.nr_rks (semilocal=True) and nr_nlc_vxc; with k=True
    also sgx_jk.get_k_only; with overlap=True mf.get_veff (XC + VV10 on a
    worker thread beside J and K); with rsh_fix=True PySCF's SGX.get_jk
    (core.sgx_patch).  Returns the HookSet.  backend=None installs nothing
    (with a warning) when the library or the GPU is unavailable;
    backend='hip' insists.  sgx_module / sgx_class name the objects to patch
    (PySCF's pyscf.sgx.sgx_jk / pyscf.sgx.sgx.SGX by default)."""
    ni = getattr(mf, '_numint', None)
    if ni is None:
        raise TypeError('mf has no _numint; not a DFT object?')
    if backend is None and not gpu_available(device):
        warnings.warn('MI355X GPU path not available; mf stays on the CPU', RuntimeWarning)
        return HookSet()
    ctx = GpuContext(mf.mol, device, precision, k_tile_tol, vv10_tile_tol)
    hooks = HookSet()
    hooks.k_stats = []
    hooks.timings = []
    hooks.ctx = ctx

    def same_mol(mol):
        if mol is not ctx.mol:
            raise Unsupported('molecule changed since install_gpu')

    def nlc_factory(orig):
        def nr_nlc_vxc(mol, grids, xc_code, dm, relativity=0, hermi=1, max_memory=2000,
                       verbose=None):
            try:
                same_mol(mol)
                return vv10.gpu_nr_nlc_vxc(ctx.nlc_state, ni, mol, grids, xc_code, dm,
                                           relativity, hermi, max_memory, verbose,
                                           tile_tol=ctx.vv10_tile_tol, precision=ctx.precision)
            except Exception as exc:      # noqa: BLE001
                _warn_fallback('nr_nlc_vxc', exc)
                return orig(mol, grids, xc_code, dm, relativity, hermi, max_memory, verbose)
        return nr_nlc_vxc

    def rks_factory(orig):
        def nr_rks(mol, grids, xc_code, dms, relativity=0, hermi=1, max_memory=2000,
                   verbose=None):
            try:
                same_mol(mol)
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
                    same_mol(sgx.mol)
                    comp = sgx_k.rsh_companions(mf)
                    kk = run_with_oom_retry(
                        lambda: sgx_k.get_k_only_gpu(ctx.k_builder, sgx, dm, hermi,
                                                     direct_scf_tol, companions=comp),
                        ctx.get_copies, ctx.set_copies)
                except Exception as exc:  # noqa: BLE001
                    _warn_fallback('get_k_only', exc)
                    return orig(sgx, dm, hermi, direct_scf_tol)
                hooks.k_stats.append(dict(ctx.k_builder.last_stats))
                return kk
            return get_k_only

        hooks.wrap(sgx_module, 'get_k_only', k_factory)
    if rsh_fix:
        sgx_patch.apply(sgx_class, hooks)
    if overlap and hasattr(mf, 'initialize_grids'):
        hooks.wrap(mf, 'get_veff', overlap_get_veff(mf, ni, hooks.timings))
    return hooks
