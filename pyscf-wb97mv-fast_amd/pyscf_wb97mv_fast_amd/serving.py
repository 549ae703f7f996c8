"""Continuous-batching request loop for SCF-cycle jobs on one GPU.

A request is one molecule's job: its basis and grids (host arrays), an FP64
density, and a number of cycles; each cycle runs the device K build for the
given operators and the VV10 term, then an optional host ``update(dm, k,
nlc, vmat) -> dm`` (the SCF step the caller owns; PySCF's in the reference,
staging/schedule.py:99-452).  Requests arrive at any time; ``slots`` worker
threads, each with its own library context (HIP stream and workspaces),
take them in arrival order as soon as a slot is free; a taken job starts
when the device bytes its slot will hold afterwards (exact, wb_job_bytes)
plus what the other slots hold fit the HBM budget, and no job that arrived
earlier is still waiting for memory (so jobs start in arrival order).  Requests that arrive
while others run are admitted into free slots, their device work overlaps
with the running jobs', and every job runs exactly the single-molecule code
on a context of its own, so its results do not depend on the mix.  Each
job's basis, grids and density stay on the device for all its cycles and
are released when it ends.  Slots keep their grow-only workspaces between
jobs; a waiting job makes its own slot and the idle slots release theirs,
so it starts once the busy slots finish, and a job that needs more than the
whole budget is refused with GpuOOM.

Latency is time from submission to the job's last cycle (p50/p99 over the
finished jobs), beside throughput in jobs/s.  The reference has no serving
path.
"""
import collections
import ctypes
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
    est_bytes: int = field(default=0, repr=False)     # device bytes the job needs from empty
    k_budget: int = field(default=0, repr=False)      # K workspace budget set to fit (0: none)
    _seq: int = field(default=0, repr=False)          # arrival number
    _shared: bool = field(default=False, repr=False)  # started while other jobs ran


def _percentile(xs, p):
    return float(numpy.percentile(numpy.asarray(xs), p)) if len(xs) else None


class CycleServer:
    def __init__(self, device=0, slots=4, hbm_budget=200e9, omegas=(0.0, 0.3), k_tol=1e-11,
                 nlc=(6.0, 0.01), vv_tol=0.0, precision='mixed', ncopy=8, runner=None,
                 contexts=None):
        """runner / contexts: for tests of the scheduling logic (a callable
        runner(ctx, request) -> result, one context object per slot, and
        request.est_bytes as the job's memory need)."""
        self.omegas = tuple(omegas)
        self.k_tol, self.nlc, self.vv_tol = float(k_tol), tuple(nlc), float(vv_tol)
        self.precision, self.ncopy = precision, int(ncopy)
        self.budget = int(hbm_budget)
        self._runner = runner
        self._own = contexts is None
        self._ctx = contexts if contexts is not None else \
            [lib.Context(device, ncopy=ncopy) for _ in range(slots)]
        n = len(self._ctx)
        self._held = [0] * n
        self._q = collections.deque()
        self._cv = threading.Condition()
        self._closed = False
        self._mem_waiters = 0
        self._seq = 0
        self._waiting = set()           # arrival numbers of taken jobs waiting for memory
        self.active = 0
        self.max_active = 0
        self.done = []
        self._threads = [threading.Thread(target=self._worker, args=(i,), daemon=True)
                         for i in range(n)]
        for t in self._threads:
            t.start()

    # -- client side -------------------------------------------------------
    def submit(self, req):
        fut = Future()
        req.t_submit = time.perf_counter()
        with self._cv:
            if self._closed:
                raise RuntimeError('server is closed')
            self._seq += 1
            req._seq = self._seq
            self._q.append((req, fut))
            self._cv.notify_all()
        return fut

    def close(self):
        """Finish every submitted request, stop the slots and release the
        device memory of the contexts the server made."""
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()
        if self._own:
            for c in self._ctx:
                c.close()

    def stats(self):
        lat = [r.t_end - r.t_submit for r in self.done]
        if not self.done:
            return {'jobs': 0}
        span = max(r.t_end for r in self.done) - min(r.t_submit for r in self.done)
        return {'jobs': len(self.done), 'jobs_per_s': len(self.done) / span if span > 0 else None,
                'p50_s': _percentile(lat, 50), 'p99_s': _percentile(lat, 99),
                'max_active': self.max_active}

    # -- slots ---------------------------------------------------------------
    def _release(self, i):
        """Trim slot i's workspaces (called with the lock held, slot idle)."""
        trim = getattr(self._ctx[i], 'trim', None)
        if trim is not None:
            trim()
        self._held[i] = 0
        self._cv.notify_all()

    def _worker(self, i):
        ctx = self._ctx[i]
        while True:
            with self._cv:
                while not self._q:
                    if self._closed:
                        return
                    if self._mem_waiters and self._held[i] > 0:
                        self._release(i)            # idle memory another slot waits for
                        continue
                    self._cv.wait()
                req, fut = self._q.popleft()
            try:
                # a job that starts while others run shares the device: its two
                # flows run one after the other (fixed for all its cycles)
                req._shared = self.active > 0
                state = self._prepare(ctx, req)
                after, alone = self._need(ctx, req, state)
            except BaseException as e:   # noqa: BLE001 -- delivered to the caller
                fut.set_exception(e)
                continue
            with self._cv:
                ok = True
                waiting = False
                while True:
                    others = sum(h for j, h in enumerate(self._held) if j != i)
                    older = any(q < req._seq for q in self._waiting)
                    if others + after <= self.budget and not older:
                        break
                    if alone > self.budget:
                        ok = False
                        break
                    if self._held[i] > 0:
                        self._release(i)
                        after = alone
                        continue
                    if not waiting:
                        waiting = True
                        self._mem_waiters += 1
                        self._waiting.add(req._seq)
                        self._cv.notify_all()       # idle slots may release memory
                    self._cv.wait()
                if waiting:
                    self._mem_waiters -= 1
                    self._waiting.discard(req._seq)
                    self._cv.notify_all()
                if ok:
                    self._held[i] = after
                    self.active += 1
                    self.max_active = max(self.max_active, self.active)
            if not ok:
                fut.set_exception(GpuOOM('job needs %d bytes of device memory, budget %d'
                                         % (alone, self.budget)))
                continue
            req.t_start = time.perf_counter()
            try:
                res = self._run(ctx, req, state)
            except BaseException as e:   # noqa: BLE001 -- delivered to the caller
                res, err = None, e
            else:
                err = None
            del state
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

    # -- device work -----------------------------------------------------------
    def _prepare(self, ctx, req):
        """The job's device state: basis, grids (resident for all its cycles)."""
        if self._runner is not None:
            return None
        basis = lib.DeviceBasis(ctx, req.pack)
        kg = lib.DeviceGrid(ctx, *req.kgrid) if req.kgrid is not None else None
        vg = lib.DeviceGrid(ctx, *req.nlcgrid) if req.nlcgrid is not None else None
        return basis, kg, vg

    def _need(self, ctx, req, state):
        """(bytes the slot holds after the job, bytes the job needs from empty).
        A job whose K build alone would not fit the budget gets a K workspace
        budget on its slot (wb_ctx_set_k_budget): its K build then runs over
        chunks of the grid, with the same result, in as little as the budget
        leaves; the slot's budget is cleared for the next job."""
        if self._runner is not None:
            held = int(getattr(ctx, 'mem_bytes', 0) or 0)
            return max(held, req.est_bytes), req.est_bytes
        basis, kg, vg = state
        ctx.set_k_budget(0)
        after, alone = self._job_bytes(ctx, req, basis, kg, vg)
        if alone > self.budget and kg is not None:
            arena = ctx.k_plan(basis, kg, len(self.omegas), self.precision)[0]
            kb = arena - (alone - self.budget)
            while kb > 0:
                ctx.set_k_budget(kb)
                try:
                    after, alone = self._job_bytes(ctx, req, basis, kg, vg)
                except GpuOOM:                   # below one split's workspace
                    break
                if alone <= self.budget:
                    req.k_budget = kb
                    break
                kb -= max(alone - self.budget, arena // 64)
            else:
                ctx.set_k_budget(0)
        req.est_bytes = int(alone)
        return int(after), int(alone)

    def _job_bytes(self, ctx, req, basis, kg, vg):
        job = lib.CycleJob()
        job.basis = basis.h.value
        job.flags = lib.JOB_SEQUENTIAL if req._shared else 0
        job.kgrid = kg.h.value if kg is not None else None
        job.nlcgrid = vg.h.value if vg is not None else None
        after, alone = ctypes.c_longlong(), ctypes.c_longlong()
        lib.check(lib.load().wb_job_bytes(ctx.h, ctypes.byref(job), len(self.omegas),
                                          lib.PREC[self.precision], ctypes.byref(after),
                                          ctypes.byref(alone)))
        return int(after.value), int(alone.value)

    def _run(self, ctx, req, state):
        if self._runner is not None:
            return self._runner(ctx, req)
        return self._run_device(ctx, req, state)

    def _run_device(self, ctx, req, state=None):
        L = lib.load()
        basis, kg, vg = state if state is not None else self._prepare(ctx, req)
        nao = basis.nao
        om = lib.f64(self.omegas)
        dm = lib.f64(req.dm)
        out = {}
        for _ in range(max(1, int(req.cycles))):
            # K build and VV10 concurrent on the device (wb_cycle) when the job
            # started alone, one after the other when it shares the device
            job = lib.CycleJob()
            job.basis = basis.h.value
            job.flags = lib.JOB_SEQUENTIAL if req._shared else 0
            if kg is not None:
                qd = lib.f64(dm if req.q is None else req.q @ dm)
                k = numpy.empty((len(om), nao, nao))
                job.kgrid, job.qd, job.k = kg.h.value, qd.ctypes.data, k.ctypes.data
            if vg is not None:
                o2, vmat = numpy.empty(2), numpy.empty((nao, nao))
                job.nlcgrid, job.dm = vg.h.value, dm.ctypes.data
                job.nlc, job.vmat = o2.ctypes.data, vmat.ctypes.data
            lib.check(L.wb_cycle(ctx.h, ctypes.byref(job), len(om), lib.ptr(om), self.k_tol,
                                 self.nlc[0], self.nlc[1], self.vv_tol, lib.PREC[self.precision]))
            if kg is not None:
                out['k'] = k
            if vg is not None:
                out['nlc'], out['vmat'] = o2, vmat
            if req.update is not None:
                dm = lib.f64(req.update(dm, out.get('k'), out.get('nlc'), out.get('vmat')))
        out['dm'] = dm
        return out
