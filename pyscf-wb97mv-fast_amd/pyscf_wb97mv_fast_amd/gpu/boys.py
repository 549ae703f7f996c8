"""Boys function F_0..F_nmax on the device (the routines the exchange
kernels inline), for tests and callers that want the values themselves.

Mirrors boys_eval(cp, T, nmax) of the reference (gpu/boys.py:194-205):
returns an (nmax + 1, len(T)) array; nmax outside 0..6 raises ValueError.
precision='mixed' evaluates in FP32 (T rounded to FP32 first), 'fp64' in
FP64.
"""
import numpy

from .. import lib

NMAX = 6


def boys_eval(ctx, T, nmax, precision='mixed'):
    if not 0 <= int(nmax) <= NMAX:
        raise ValueError('nmax must be 0..%d' % NMAX)
    T = lib.f64(numpy.ravel(T))
    out = numpy.empty((int(nmax) + 1, T.size))
    lib.check(lib.load().wb_boys_eval(ctx.h, T.size, lib.ptr(T), int(nmax),
                                      lib.PREC[precision], lib.ptr(out)))
    return out
