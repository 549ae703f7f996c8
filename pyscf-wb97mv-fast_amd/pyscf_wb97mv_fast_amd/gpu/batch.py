"""Cycles of many molecules in one call (wb_cycle_batch, wb_cycle_queue).

Each job is one molecule's SCF-cycle work on the device: the raw K for the
given operators on its SGX grid (what get_k_only computes per operator,
gpu/sgx_k.py:663-751) and/or its VV10 term (nr_nlc_vxc, gpu/vv10.py:268-342).
``run_batch``: the library drains the jobs largest first from one shared
queue with ``lanes`` host threads, each driving its own HIP stream and
workspaces, so small molecules' launches overlap on the device.
``run_queue``: the exchange kernels of all the jobs' K builds run as one
persistent launch per shell-pair class that drains a device work queue of
(molecule, sub-block, chunk) items.  Either way a job's outputs equal the
single-molecule calls' for it alone, bit for bit.  The reference has no
batch path; it runs one molecule per SCF (staging/schedule.py:99-452).
"""
import ctypes
from dataclasses import dataclass, field

import numpy

from .. import lib


@dataclass
class CycleJob:
    basis: 'lib.DeviceBasis'
    kgrid: 'lib.DeviceGrid' = None
    qd: numpy.ndarray = None            # Q dm (nao, nao)
    nlcgrid: 'lib.DeviceGrid' = None
    dm: numpy.ndarray = None            # dm (nao, nao) for VV10
    # outputs
    k: numpy.ndarray = field(default=None, repr=False)       # (nop, nao, nao)
    nlc: numpy.ndarray = field(default=None, repr=False)     # (nelec, exc)
    vmat: numpy.ndarray = field(default=None, repr=False)    # (nao, nao)
    ms: float = None                    # completion time from the call's start


def _pack(jobs, omegas):
    om = lib.f64(omegas) if len(omegas) else numpy.zeros(1)
    arr = (lib.CycleJob * max(1, len(jobs)))()
    for i, j in enumerate(jobs):
        nao = j.basis.nao
        c = arr[i]
        c.basis = j.basis.h.value
        if j.kgrid is not None:
            j.qd = lib.f64(j.qd)
            if j.qd.shape != (nao, nao):
                raise ValueError('job %d: qd must be (nao, nao)' % i)
            j.k = numpy.empty((len(omegas), nao, nao))
            c.kgrid, c.qd, c.k = j.kgrid.h.value, j.qd.ctypes.data, j.k.ctypes.data
        if j.nlcgrid is not None:
            j.dm = lib.f64(j.dm)
            if j.dm.shape != (nao, nao):
                raise ValueError('job %d: dm must be (nao, nao)' % i)
            j.nlc = numpy.empty(2)
            j.vmat = numpy.empty((nao, nao))
            c.nlcgrid, c.dm = j.nlcgrid.h.value, j.dm.ctypes.data
            c.nlc, c.vmat = j.nlc.ctypes.data, j.vmat.ctypes.data
    return om, arr


def run_batch(ctx, jobs, omegas=(0.0, 0.3), k_tol=1e-11, nlc=(6.0, 0.01), vv_tol=0.0,
              precision='mixed', lanes=4):
    """Run every job on ``lanes`` host threads; fills job.k / job.nlc /
    job.vmat / job.ms and returns jobs."""
    om, arr = _pack(jobs, omegas)
    ms = numpy.zeros(max(1, len(jobs)))
    lib.check(lib.load().wb_cycle_batch(ctx.h, len(jobs), arr, len(omegas), lib.ptr(om),
                                        float(k_tol), float(nlc[0]), float(nlc[1]),
                                        float(vv_tol), lib.PREC[precision], int(lanes),
                                        lib.ptr(ms)))
    for j, t in zip(jobs, ms):
        j.ms = float(t)
    return jobs


def run_queue(ctx, jobs, omegas=(0.0, 0.3), k_tol=1e-11, nlc=(6.0, 0.01), vv_tol=0.0,
              precision='mixed'):
    """Run every job with one exchange-kernel launch per shell-pair class
    over all of them; fills job.k / job.nlc / job.vmat / job.ms and returns
    jobs."""
    om, arr = _pack(jobs, omegas)
    ms = numpy.zeros(max(1, len(jobs)))
    lib.check(lib.load().wb_cycle_queue(ctx.h, len(jobs), arr, len(omegas), lib.ptr(om),
                                        float(k_tol), float(nlc[0]), float(nlc[1]),
                                        float(vv_tol), lib.PREC[precision], lib.ptr(ms)))
    for j, t in zip(jobs, ms):
        j.ms = float(t)
    return jobs
