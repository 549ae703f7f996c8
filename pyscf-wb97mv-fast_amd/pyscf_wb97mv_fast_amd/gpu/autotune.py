"""Autotuner of the K build's run-time choices (wb_ctx_autotune).

The library times, on the caller's own molecule and grid, the number of G
partial copies the exchange kernels write into (fewer copies: less memory
and a cheaper fold; more: more waves per sub-block), per (la, lb) class
with a fused two-operator kernel fused vs two passes, the FP64 GEMMs' MFMA
tiles per wave side, and the waves per sub-block of each (la, lb) class's
exchange kernel, and keeps the fastest on the context.  ``Tuner`` caches the choice per size class
(AOs and grid points rounded up to powers of two, operator count,
precision), so molecules of similar size reuse one tuning run; the cache
can be saved as JSON and loaded by a later process.  The reference runs
fixed launch shapes and has no tuner.
"""
import ctypes
import json

import numpy

from .. import lib


def autotune(ctx, basis, grid, qd, omegas=(0.0, 0.3), tile_tol=1e-11, precision='mixed',
             reps=3):
    """Tune ``ctx`` on one K build; returns the report as a dict."""
    om = lib.f64(omegas)
    qd = lib.f64(qd)
    if qd.shape != (basis.nao, basis.nao):
        raise ValueError('qd must be (nao, nao)')
    rep = numpy.zeros(6)
    lib.check(lib.load().wb_ctx_autotune(ctx.h, basis.h, grid.h, lib.ptr(qd), len(om),
                                         lib.ptr(om), float(tile_tol), lib.PREC[precision],
                                         int(reps), lib.ptr(rep)))
    return {'ncopy': int(rep[0]), 'fuse_mask': int(rep[1]), 'best_ms': float(rep[2]),
            'default_ms': float(rep[3]), 'configs': int(rep[4]), 'gemm_tiles': int(rep[5]),
            'class_waves': get_class_waves(ctx)}


def get_class_waves(ctx):
    """Waves per sub-block of each (la, lb) class (0: the copy count)."""
    w = numpy.zeros(10, dtype=numpy.int32)
    lib.check(lib.load().wb_ctx_get_class_waves(ctx.h, lib.ptr(w)))
    return [int(v) for v in w]


def set_class_waves(ctx, waves=None):
    if waves is None:
        lib.check(lib.load().wb_ctx_set_class_waves(ctx.h, None))
        return
    w = numpy.ascontiguousarray(waves, dtype=numpy.int32)
    if w.shape != (10,):
        raise ValueError('one value per (la, lb) class: 10')
    lib.check(lib.load().wb_ctx_set_class_waves(ctx.h, lib.ptr(w)))


def get_tuning(ctx):
    """(copies, fuse mask, GEMM wave tiles) of the context."""
    n, m, t = ctypes.c_int(), ctypes.c_uint(), ctypes.c_int()
    lib.check(lib.load().wb_ctx_get_tuning(ctx.h, ctypes.byref(n), ctypes.byref(m),
                                           ctypes.byref(t)))
    return int(n.value), int(m.value), int(t.value)


def set_tuning(ctx, ncopy, fuse_mask, gemm_tiles=0):
    lib.check(lib.load().wb_ctx_set_tuning(ctx.h, int(ncopy), ctypes.c_uint(int(fuse_mask)),
                                           int(gemm_tiles)))


def _pow2(n):
    return 1 << max(0, int(n) - 1).bit_length()


class Tuner:
    """Per-size-class cache of autotune results."""

    def __init__(self, reps=3):
        self.reps = reps
        self.table = {}

    @staticmethod
    def key(nao, npts, nop, precision):
        return '%d/%d/%d/%s' % (_pow2(nao), _pow2(npts), nop, precision)

    def apply(self, ctx, basis, grid, qd, omegas=(0.0, 0.3), tile_tol=1e-11,
              precision='mixed'):
        """Set ctx's tuning for this size class, tuning on this molecule the
        first time the class is seen; returns (key, entry)."""
        k = self.key(basis.nao, grid.npts, len(omegas), precision)
        e = self.table.get(k)
        if e is None:
            e = autotune(ctx, basis, grid, qd, omegas, tile_tol, precision, self.reps)
            self.table[k] = e
        else:
            set_tuning(ctx, e['ncopy'], e['fuse_mask'], e.get('gemm_tiles', 0))
            set_class_waves(ctx, e.get('class_waves'))
        return k, e

    def save(self, path):
        with open(path, 'w') as f:
            json.dump(self.table, f, indent=1, sort_keys=True)

    def load(self, path):
        with open(path) as f:
            self.table.update(json.load(f))
        return self
