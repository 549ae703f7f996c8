"""ctypes binding of libwb97mv_amd.so (the gfx950 C ABI, include/wb97mv_amd.h).

The library must have been built in-tree (``__graft_entry__.build()``);
there is no CPU fallback behind this module: a missing library or a missing
GPU raises, and the hook layer (gpu/install.py) turns that into PySCF's own
CPU path with a warning, exactly as the reference does on Unsupported
(gpu/install.py:264-287).
"""
import ctypes
import os
import threading

import numpy

from .errors import GpuError, GpuOOM, Unsupported

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libwb97mv_amd.so')

WB_OK, WB_UNSUPPORTED, WB_OOM, WB_DEVICE_ERROR, WB_INVALID = range(5)
PREC = {'mixed': 0, 'fp32': 0, 'fp64': 1}

_lib = None
_lock = threading.Lock()

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double


class CycleJob(ctypes.Structure):
    """wb_cycle_job (include/wb97mv_amd.h)."""
    _fields_ = [('basis', ctypes.c_void_p), ('kgrid', ctypes.c_void_p), ('qd', ctypes.c_void_p),
                ('k', ctypes.c_void_p), ('nlcgrid', ctypes.c_void_p), ('dm', ctypes.c_void_p),
                ('nlc', ctypes.c_void_p), ('vmat', ctypes.c_void_p), ('flags', ctypes.c_int)]


JOB_SEQUENTIAL = 1


_SIGS = {
    'wb_last_error': (ctypes.c_char_p, []),
    'wb_version': (_i, []),
    'wb_ctx_create': (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    'wb_ctx_destroy': (_i, [_vp]),
    'wb_ctx_set_copies': (_i, [_vp, _i]),
    'wb_ctx_set_g64': (_i, [_vp, _i]),
    'wb_ctx_set_k_budget': (_i, [_vp, ctypes.c_longlong]),
    'wb_k_plan': (_i, [_vp, _vp, _vp, _i, _i, _vp]),
    'wb_sync': (_i, [_vp]),
    'wb_last_ms': (ctypes.c_float, [_vp]),
    'wb_ctx_set_profile': (_i, [_vp, _i]),
    'wb_profile_read': (_i, [_vp, _vp, _i, _i]),
    'wb_k_work': (_i, [_vp, _vp]),
    'wb_basis_create': (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _d, ctypes.POINTER(_vp)]),
    'wb_basis_info': (_i, [_vp, _vp]),
    'wb_basis_destroy': (_i, [_vp]),
    'wb_grid_create': (_i, [_vp, _i, _vp, _vp, ctypes.POINTER(_vp)]),
    'wb_grid_info': (_i, [_vp, _vp]),
    'wb_grid_destroy': (_i, [_vp]),
    'wb_k_build': (_i, [_vp, _vp, _vp, _vp, _i, _vp, _d, _i, _vp]),
    'wb_k_build_dev': (_i, [_vp, _vp, _vp, _vp, _i, _vp, _d, _i, _vp]),
    'wb_k_build_noalloc': (_i, [_vp, _vp, _vp, _vp, _i, _vp, _d, _i, _vp]),
    'wb_k_stats': (_i, [_vp, _vp, _vp]),
    'wb_nlc_vxc': (_i, [_vp, _vp, _vp, _vp, _d, _d, _d, _i, _vp, _vp]),
    'wb_nlc_vxc_dev': (_i, [_vp, _vp, _vp, _vp, _d, _d, _d, _i, _vp, _vp]),
    'wb_vv10_sum': (_i, [_vp, _i, _vp, _vp, _vp, _vp, _d, _i, _vp, _vp, _vp]),
    'wb_ao_eval': (_i, [_vp, _vp, _i, _vp, _i, _vp]),
    'wb_boys_eval': (_i, [_vp, _i, _vp, _i, _i, _vp]),
    'wb_grid_partition': (_i, [_vp, _i, _vp, _i, _vp, _vp, _vp]),
    'wb_grid_overlap': (_i, [_vp, _vp, _vp, _vp]),
    'wb_pjs_projector': (_i, [_vp, _vp, _vp, _vp, _vp]),
    'wb_grid_build': (_i, [_vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _d,
                           ctypes.c_longlong, _vp, _vp, ctypes.POINTER(ctypes.c_longlong)]),
    'wb_cycle_batch': (_i, [_vp, _i, ctypes.POINTER(CycleJob), _i, _vp, _d, _d, _d, _d, _i, _i,
                            _vp]),
    'wb_cycle_queue': (_i, [_vp, _i, ctypes.POINTER(CycleJob), _i, _vp, _d, _d, _d, _d, _i,
                            _vp]),
    'wb_ctx_mem': (_i, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    'wb_cycle': (_i, [_vp, ctypes.POINTER(CycleJob), _i, _vp, _d, _d, _d, _d, _i]),
    'wb_job_bytes': (_i, [_vp, ctypes.POINTER(CycleJob), _i, _i, ctypes.POINTER(ctypes.c_longlong),
                          ctypes.POINTER(ctypes.c_longlong)]),
    'wb_ctx_trim': (_i, [_vp]),
    'wb_ctx_autotune': (_i, [_vp, _vp, _vp, _vp, _i, _vp, _d, _i, _i, _vp]),
    'wb_ctx_get_tuning': (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_uint),
                               ctypes.POINTER(_i)]),
    'wb_ctx_set_tuning': (_i, [_vp, _i, ctypes.c_uint, _i]),
    'wb_ctx_get_class_waves': (_i, [_vp, _vp]),
    'wb_ctx_set_class_waves': (_i, [_vp, _vp]),
    'wb_test_gemm': (_i, [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp]),
    'wb_test_pair_bounds': (_i, [_i, _vp, _vp, _vp, _vp, _vp, _d, _vp, _vp, _vp]),
    'wb_xc_rho': (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    'wb_xc_vmat': (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    'wb_cycle_graph_create': (_i, [_vp, _vp, _vp, _vp, _i, _vp, _d, _vp, _vp, _d, _d, _d, _i,
                                   _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    'wb_cycle_dev': (_i, [_vp, _vp, _vp, _vp, _i, _vp, _d, _vp, _vp, _d, _d, _d, _i, _vp, _vp,
                          _vp]),
    'wb_vv10_stats': (_i, [_vp, _vp]),
    'wb_basis_bytes': (_i, [_vp, _i, _vp]),
    'wb_grid_bytes': (_i, [_vp, _vp]),
    'wb_cycle_graph_launch': (_i, [_vp]),
    'wb_cycle_graph_destroy': (_i, [_vp]),
}
EXPORTS = tuple(_SIGS)


def load(path=LIB_PATH):
    """The loaded library (ctypes releases the GIL around every call)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise Unsupported('libwb97mv_amd.so is not built (%s); run '
                                  '__graft_entry__.build()' % path)
            try:
                # PyTorch ships its own HIP runtime; loaded after this
                # library it finds no device, so it goes first when present
                import torch  # noqa: F401
            except ImportError:
                pass
            lib = ctypes.CDLL(path)
            for name, (res, args) in _SIGS.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(status):
    if status == WB_OK:
        return
    msg = load().wb_last_error().decode(errors='replace')
    if status == WB_UNSUPPORTED:
        raise Unsupported(msg)
    if status == WB_OOM:
        raise GpuOOM(msg)
    raise GpuError('wb97mv_amd status %d: %s' % (status, msg))


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def f64(a):
    return numpy.ascontiguousarray(a, dtype=numpy.float64)


class Context:
    """One device context (a private HIP stream) of the library."""

    def __init__(self, device=0, ncopy=None):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.wb_ctx_create(int(device), None, ctypes.byref(h)))
        self.h = h
        self.device = int(device)
        if ncopy is not None:
            check(lib.wb_ctx_set_copies(self.h, int(ncopy)))

    def close(self):
        if getattr(self, 'h', None):
            load().wb_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:      # noqa: BLE001 -- interpreter shutdown
            pass

    def sync(self):
        check(load().wb_sync(self.h))

    @property
    def last_ms(self):
        return float(load().wb_last_ms(self.h))

    @property
    def mem_bytes(self):
        """Device bytes held by the context's workspaces (batch lanes included)."""
        v = ctypes.c_longlong()
        check(load().wb_ctx_mem(self.h, ctypes.byref(v)))
        return int(v.value)

    def trim(self):
        """Release the workspaces (they regrow on the next call)."""
        check(load().wb_ctx_trim(self.h))

    def vv10_stats(self):
        """(tile pairs tested, skipped) by the pruning of the last VV10 sum
        on this context (both 0 without pruning)."""
        out = numpy.zeros(2, dtype=numpy.int64)
        check(load().wb_vv10_stats(self.h, ptr(out)))
        return int(out[0]), int(out[1])

    def set_k_budget(self, nbytes):
        """Workspace budget of this context's K builds (0: no limit); a
        build over budget runs over chunks of the grid, bit for bit the same."""
        check(load().wb_ctx_set_k_budget(self.h, int(nbytes)))

    def k_plan(self, basis, grid, nop=2, precision='mixed'):
        """(workspace bytes, chunks, sub-blocks per chunk) of a K build."""
        out = numpy.zeros(3, dtype=numpy.int64)
        check(load().wb_k_plan(self.h, basis.h, grid.h, int(nop), PREC[precision], ptr(out)))
        return tuple(int(v) for v in out)

    def tuning(self):
        """(G copies, fused-class mask, GEMM tiles) the context runs with."""
        n, m, t = ctypes.c_int(), ctypes.c_uint(), ctypes.c_int()
        check(load().wb_ctx_get_tuning(self.h, ctypes.byref(n), ctypes.byref(m),
                                       ctypes.byref(t)))
        return n.value, m.value, t.value


class DeviceBasis:
    """Basis handle; shells as arrays from basis.pack_mol."""

    def __init__(self, ctx, pack, pair_cut=1e-14):
        lib = load()
        self.ctx = ctx
        self.pack = pack
        l = numpy.ascontiguousarray(pack.l, dtype=numpy.int32)
        nprim = numpy.ascontiguousarray(pack.nprim, dtype=numpy.int32)
        cen, ex, co = f64(pack.center), f64(pack.exps), f64(pack.coefs)
        h = ctypes.c_void_p()
        check(lib.wb_basis_create(ctx.h, len(l), ptr(l), ptr(nprim), ptr(cen), ptr(ex), ptr(co),
                                  float(pair_cut), ctypes.byref(h)))
        self.h = h
        info = numpy.zeros(4, dtype=numpy.int64)
        check(lib.wb_basis_info(self.h, ptr(info)))
        self.nao, self.nshell, self.npairs, self.nprims = (int(v) for v in info)

    def device_bytes(self, nop=0):
        """Device bytes held, plus the records a K build of nop operators
        would still add."""
        v = ctypes.c_longlong()
        check(load().wb_basis_bytes(self.h, int(nop), ctypes.byref(v)))
        return int(v.value)

    def __del__(self):
        try:
            if getattr(self, 'h', None):
                load().wb_basis_destroy(self.h)
        except Exception:      # noqa: BLE001
            pass


class DeviceGrid:
    """Grid handle (coords bohr (n, 3), weights (n,)), resident on the device."""

    def __init__(self, ctx, coords, weights):
        lib = load()
        self.ctx = ctx
        coords, weights = f64(coords).reshape(-1, 3), f64(weights).ravel()
        if len(coords) != len(weights):
            raise ValueError('coords / weights size mismatch')
        h = ctypes.c_void_p()
        check(lib.wb_grid_create(ctx.h, len(weights), ptr(coords), ptr(weights),
                                 ctypes.byref(h)))
        self.h = h
        info = numpy.zeros(3, dtype=numpy.int64)
        check(lib.wb_grid_info(self.h, ptr(info)))
        self.npts, self.ppad, self.nsub = (int(v) for v in info)

    def device_bytes(self):
        v = ctypes.c_longlong()
        check(load().wb_grid_bytes(self.h, ctypes.byref(v)))
        return int(v.value)

    def __del__(self):
        try:
            if getattr(self, 'h', None):
                load().wb_grid_destroy(self.h)
        except Exception:      # noqa: BLE001
            pass


class CycleGraph:
    """One SCF cycle recorded as a HIP graph (wb_cycle_graph_create): the K
    build (kgrid: d_qd -> d_k, len(omegas) operators) and/or the VV10 term
    (nlcgrid: d_dm -> d_out2, d_vmat).  The d_* arguments are device
    addresses (ints, e.g. ``tensor.data_ptr()``) fixed at recording; write
    each cycle's inputs into the same buffers and call :meth:`launch`.  The
    basis and grids must stay alive as long as the graph."""

    def __init__(self, ctx, basis, kgrid=None, d_qd=0, omegas=(), k_tol=1e-11, nlcgrid=None,
                 d_dm=0, nlc=(6.0, 0.01), vv_tol=0.0, precision='mixed', d_k=0, d_out2=0,
                 d_vmat=0):
        lib = load()
        self.ctx, self.basis, self.kgrid, self.nlcgrid = ctx, basis, kgrid, nlcgrid
        om = f64(omegas) if len(omegas) else numpy.zeros(1)
        h = ctypes.c_void_p()
        check(lib.wb_cycle_graph_create(
            ctx.h, basis.h, kgrid.h if kgrid is not None else None, d_qd or None,
            len(omegas), ptr(om), float(k_tol), nlcgrid.h if nlcgrid is not None else None,
            d_dm or None, float(nlc[0]), float(nlc[1]), float(vv_tol), PREC[precision],
            d_k or None, d_out2 or None, d_vmat or None, ctypes.byref(h)))
        self.h = h

    def launch(self):
        """Enqueue one replay on the context's stream (ctx.sync() waits)."""
        check(load().wb_cycle_graph_launch(self.h))

    def close(self):
        if getattr(self, 'h', None):
            load().wb_cycle_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:      # noqa: BLE001
            pass
