"""Continuous-batching request loop for SCF-cycle jobs on one GPU.

A request is one molecule's job: its basis and grids (host arrays), an FP64
density, and a number of cycles; each cycle runs the device K build for the
given operators and the VV10 term, then an optional host ``update(dm, k,
nlc, vmat) -> dm`` (the SCF step the caller owns; PySCF's in the reference,
staging/schedule.py:99-452).  Requests arrive at any time; ``slots`` worker
threads, each with its own library context (HIP stream and workspaces),
take them in arrival order as soon as a slot is free and the job's
estimated device footprint fits the HBM budget beside what the other slots
hold.  Requests that arrive while others run are admitted into free slots,
their device work overlaps with the running jobs', and every job runs
exactly the single-molecule code on a context of its own, so its results do
not depend on the mix.  Each job's basis, grids and density stay on the
device for all its cycles and are released when it ends; an idle slot
releases its workspaces when the next request does not fit, so the head of
the queue always runs once the busy slots finish.

Latency is time from submission to the job's last cycle (p50/p99 over the
finished jobs), beside throughput in jobs/s.  The reference has no serving
path.
"""
import collections
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy

from . import lib
from .errors import GpuOOM


@dataclass
class Request:
    pack: object                        # basis.pack_mol(mol)
    kgrid: tuple = None                 # (coords (n, 3), weights (n,)) of the SGX grid
    nlcgrid: tuple = None               # (coords, weights) of the VV10 grid
    dm: numpy.ndarray = None            # FP64 density (nao, nao)
    q: numpy.ndarray = None             # PJS projector (K uses Q dm); None = identity
    cycles: int = 1
    update: object = None               # host SCF step: update(dm, k, nlc, vmat) -> dm
    tag: object = None
    # filled by the server
    t_submit: float = field(default=None, repr=False)
    t_start: float = field(default=None, repr=False)
    t_end: float = field(default=None, repr=False)
    est_bytes: int = field(default=0, repr=False)


def footprint(nao, kpts=0, nlcpts=0, nop=2, ncopy=8, precision='mixed', npairs=0):
    """Estimate of the device bytes one job needs (workspaces + basis and grid
    data): X, F, the G copies and the task lists (at most one word pair per
    quarter of the shell pairs per 64-point sub-block) for the K build, AO
    values, density intermediates and the pair-sum buffers for VV10."""
    pk = -(-kpts // 64) * 64
    pv = -(-nlcpts // 64) * 64
    gb = 8 if precision == 'fp64' else 4
    k = 0
    if kpts:
        k = nao * pk * (8 + 8) + ncopy * nop * (nao + 8) * pk * gb \
            + (pk // 64) * npairs * 2 + 96 * nao * nao * 8 + 8 * nao * nao * (1 + nop)
    v = 0
    if nlcpts:
        v = nao * pv * 8 * 4 + pv * 8 * 32 + 8 * nao * nao * 3
    return int(1.10 * (k + v)) + (64 << 20)


def _percentile(xs, p):
    return float(numpy.percentile(numpy.asarray(xs), p)) if len(xs) else None


class CycleServer:
    def __init__(self, device=0, slots=4, hbm_budget=200e9, omegas=(0.0, 0.3), k_tol=1e-11,
                 nlc=(6.0, 0.01), vv_tol=0.0, precision='mixed', ncopy=8, runner=None,
                 contexts=None):
        """runner / contexts: for tests of the scheduling logic (a callable
        runner(ctx, request) -> result and one context object per slot)."""
        self.omegas = tuple(omegas)
        self.k_tol, self.nlc, self.vv_tol = float(k_tol), tuple(nlc), float(vv_tol)
        self.precision, self.ncopy = precision, int(ncopy)
        self.budget = int(hbm_budget)
        self._run = runner or self._run_device
        self._ctx = contexts if contexts is not None else \
            [lib.Context(device, ncopy=ncopy) for _ in range(slots)]
        n = len(self._ctx)
        self._held = [0] * n
        self._q = collections.deque()
        self._cv = threading.Condition()
        self._closed = False
        self.active = 0
        self.max_active = 0
        self.done = []
        self._threads = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                         for i in range(n)]
        for t in self._threads:
            t.start()

    # -- client side -------------------------------------------------------
    def submit(self, req):
        if req.est_bytes <= 0:
            nsh = len(req.pack.l)
            req.est_bytes = footprint(int(req.pack.nao), len(req.kgrid[1]) if req.kgrid else 0,
                                      len(req.nlcgrid[1]) if req.nlcgrid else 0,
                                      len(self.omegas), self.ncopy, self.precision,
                                      nsh * (nsh + 1) // 2)
        fut = Future()
        req.t_submit = time.perf_counter()
        with self._cv:
            if self._closed:
                raise RuntimeError('server is closed')
            self._q.append((req, fut))
            self._cv.notify_all()
        return fut

    def close(self):
        """Finish every submitted request, then stop the slots."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    def stats(self):
        lat = [r.t_end - r.t_submit for r in self.done]
        if not self.done:
            return {'jobs': 0}
        span = max(r.t_end for r in self.done) - min(r.t_submit for r in self.done)
        return {'jobs': len(self.done), 'jobs_per_s': len(self.done) / span if span > 0 else None,
                'p50_s': _percentile(lat, 50), 'p99_s': _percentile(lat, 99),
                'max_active': self.max_active}

    # -- slots ---------------------------------------------------------------
    def _worker(self, i):
        ctx = self._ctx[i]
        while True:
            with self._cv:
                trim = False
                while True:
                    if not self._q:
                        if self._closed:
                            return
                        self._cv.wait()
                        continue
                    req, fut = self._q[0]
                    if req.est_bytes > self.budget:
                        self._q.popleft()
                        fut.set_exception(GpuOOM('job needs about %d bytes, budget %d'
                                                 % (req.est_bytes, self.budget)))
                        continue
                    others = sum(h for j, h in enumerate(self._held) if j != i)
                    need = max(self._held[i], req.est_bytes)
                    if others + need <= self.budget:
                        self._q.popleft()
                        self._held[i] = need
                        self.active += 1
                        self.max_active = max(self.max_active, self.active)
                        break
                    if self._held[i] > 0:
                        # idle and holding memory that may be what blocks the
                        # head of the queue (here or on another slot): release
                        # it; once every idle slot has, the head fits
                        trim = True
                        break
                    self._cv.wait()
            if trim:
                ctx.trim()
                with self._cv:
                    self._held[i] = 0
                    self._cv.notify_all()
                continue
            req.t_start = time.perf_counter()
            try:
                res = self._run(ctx, req)
            except BaseException as e:   # noqa: BLE001 -- delivered to the caller
                res, err = None, e
            else:
                err = None
            req.t_end = time.perf_counter()
            with self._cv:
                self.active -= 1
                mem = getattr(ctx, 'mem_bytes', None)
                self._held[i] = int(mem) if mem is not None else self._held[i]
                if err is None:
                    self.done.append(req)
                self._cv.notify_all()
            if err is None:
                fut.set_result(res)
            else:
                fut.set_exception(err)

    def _run_device(self, ctx, req):
        L = lib.load()
        basis = lib.DeviceBasis(ctx, req.pack)
        nao = basis.nao
        kg = lib.DeviceGrid(ctx, *req.kgrid) if req.kgrid is not None else None
        vg = lib.DeviceGrid(ctx, *req.nlcgrid) if req.nlcgrid is not None else None
        om = lib.f64(self.omegas)
        dm = lib.f64(req.dm)
        out = {}
        for _ in range(max(1, int(req.cycles))):
            if kg is not None:
                qd = lib.f64(dm if req.q is None else req.q @ dm)
                k = numpy.empty((len(om), nao, nao))
                lib.check(L.wb_k_build(ctx.h, basis.h, kg.h, lib.ptr(qd), len(om), lib.ptr(om),
                                       self.k_tol, lib.PREC[self.precision], lib.ptr(k)))
                out['k'] = k
            if vg is not None:
                o2, vmat = numpy.empty(2), numpy.empty((nao, nao))
                lib.check(L.wb_nlc_vxc(ctx.h, basis.h, vg.h, lib.ptr(dm), self.nlc[0],
                                       self.nlc[1], self.vv_tol, lib.PREC[self.precision],
                                       lib.ptr(o2), lib.ptr(vmat)))
                out['nlc'], out['vmat'] = o2, vmat
            if req.update is not None:
                dm = lib.f64(req.update(dm, out.get('k'), out.get('nlc'), out.get('vmat')))
        out['dm'] = dm
        return out
