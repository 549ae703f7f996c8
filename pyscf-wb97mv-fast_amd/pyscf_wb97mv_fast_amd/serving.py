"""Continuous-batching request loop for SCF jobs on one GPU.

A request is one molecule's job: its basis and grids (host arrays) and an
FP64 density.  It either runs a fixed number of cycles -- each cycle the
device K build for the given operators and the VV10 term, then an optional
host ``update(dm, k, nlc, vmat) -> dm`` -- or, given an SCF step ``scf(dm,
k, nlc, vmat) -> (dm_next, energy, gradient_norm)`` (PySCF's Fock build and
diagonalization in real use; the reference's staged SCF,
staging/schedule.py:99-452), runs to convergence through ``stages``: grids
and precisions in order, coarse to fine, ending with an FP64 tail, the next
stage taken when the energy change and the orbital gradient fall below the
stage's thresholds or after its maximum number of cycles (the switching
rules of staging/schedule.py:289-350, :368-394), with the step's history
reset at each switch; the job is converged when, in the last stage, the
energy changes by less than conv_tol and the gradient norm is below
conv_tol_grad.

Requests arrive at any time; ``slots`` worker threads, each with its own
library context (HIP stream and workspaces), take them in arrival order as
soon as a slot is free; a taken job starts when the device bytes its slot
will hold afterwards (exact, wb_job_bytes, the most demanding stage) plus
what the other slots hold fit the HBM budget, and no job that arrived
earlier is still waiting for memory (so jobs start in arrival order).  A job
whose K build alone would not fit gets a K workspace budget on its slot (the
build runs over grid chunks, same result).  Requests that arrive while
others run are admitted into free slots, their device work overlaps with the
running jobs', and every job runs exactly the single-molecule code on a
context of its own, so its results do not depend on the mix.  Each job's
basis and every stage's grids stay on the device for the whole job and are
released when it ends.  Slots keep their grow-only workspaces between jobs;
a waiting job makes its own slot and the idle slots release theirs, so it
starts once the busy slots finish, and a job that needs more than the whole
budget is refused with GpuOOM.

Latency is time from submission to the job's last cycle -- for SCF jobs the
time to the converged energy -- (p50/p99 over the finished jobs), beside
throughput in jobs/s.  The reference has no serving path.

The host steps (``scf`` / ``update``: Fock matrices, DIIS, diagonalization)
of the slots run concurrently on the host's cores.  A BLAS library that
gives each of them every core oversubscribes the host (measured on a
192-AO water cluster: 0.15-0.23 s per SCF step in the serving process
against 2-9 ms of device cycle); ``host_threads`` (default: the process's
thread budget, ``OMP_NUM_THREADS`` or the usable cores, divided by the
slots) caps the BLAS thread pools while the server runs, through
threadpoolctl when it is installed, and restores them at close().  BLAS
results can round differently with another thread count, so a job's
results equal bit for bit across servers with the same host_threads, and
within one server alone and in any mix.
"""
import collections
import ctypes
import os
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy

from . import lib
from .errors import GpuOOM


@dataclass
class Stage:
    """One stage of a staged SCF job: its grids, the precision of the
    integral and VV10 arithmetic, and when to move on (|dE| < tau_e and
    gradient norm < tau_g, or max_cycles cycles in the stage)."""
    kgrid: tuple = None                 # (coords, weights) of the SGX grid
    nlcgrid: tuple = None               # (coords, weights) of the VV10 grid
    precision: str = 'mixed'
    tau_e: float = 1e-6
    tau_g: float = 1e-3
    max_cycles: int = 8


def staged(coarse, fine, fp64_tail=True, tau=(1e-5, 1e-7), tau_g=(1e-2, 1e-4), max_cycles=8):
    """The coarse -> fine schedule with an FP64 tail: coarse grids (mixed),
    fine grids (mixed), fine grids in FP64 to convergence.  coarse and fine
    are (kgrid, nlcgrid) pairs of (coords, weights)."""
    out = [Stage(coarse[0], coarse[1], 'mixed', tau[0], tau_g[0], max_cycles),
           Stage(fine[0], fine[1], 'mixed', tau[1], tau_g[1], max_cycles)]
    if fp64_tail:
        out.append(Stage(fine[0], fine[1], 'fp64', 0.0, 0.0, 10 ** 6))
    return out


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
    # run to convergence: scf(dm, k, nlc, vmat) -> (dm, energy, gradient norm)
    # through stages (kgrid / nlcgrid above are then unused)
    scf: object = None
    stages: list = None
    conv_tol: float = 1e-8
    conv_tol_grad: float = 1e-5
    max_cycles: int = 100
    # filled by the server
    t_submit: float = field(default=None, repr=False)
    t_start: float = field(default=None, repr=False)
    t_end: float = field(default=None, repr=False)
    est_bytes: int = field(default=0, repr=False)     # device bytes the job needs from empty
    state_bytes: int = field(default=0, repr=False)   # of which its basis, records and grids
    k_budget: int = field(default=0, repr=False)      # K workspace budget set to fit (0: none)
    _seq: int = field(default=0, repr=False)          # arrival number
    _shared: bool = field(default=False, repr=False)  # started while other jobs ran


def host_thread_budget():
    """Threads this process may keep busy: OMP_NUM_THREADS when set, else the
    cores it may run on."""
    try:
        n = int(os.environ.get('OMP_NUM_THREADS', '0'))
    except ValueError:
        n = 0
    if n > 0:
        return n
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def _limit_blas(n):
    """Cap every loaded BLAS thread pool at n threads (process-wide); the
    returned object restores them.  None without threadpoolctl."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        return None
    return threadpool_limits(limits=int(n), user_api='blas')


def _percentile(xs, p):
    return float(numpy.percentile(numpy.asarray(xs), p)) if len(xs) else None


class CycleServer:
    def __init__(self, device=0, slots=4, hbm_budget=200e9, omegas=(0.0, 0.3), k_tol=1e-11,
                 nlc=(6.0, 0.01), vv_tol=0.0, precision='mixed', ncopy=8, runner=None,
                 contexts=None, host_threads=None):
        """runner / contexts: for tests of the scheduling logic (a callable
        runner(ctx, request) -> result, one context object per slot, and
        request.est_bytes as the job's memory need).  host_threads: BLAS
        threads per slot's host step (None: the thread budget / slots; 0:
        leave the BLAS pools as they are)."""
        self.omegas = tuple(omegas)
        self.k_tol, self.nlc, self.vv_tol = float(k_tol), tuple(nlc), float(vv_tol)
        self.precision, self.ncopy = precision, int(ncopy)
        self.budget = int(hbm_budget)
        self._runner = runner
        self._own = contexts is None
        self._ctx = contexts if contexts is not None else \
            [lib.Context(device, ncopy=ncopy) for _ in range(slots)]
        n = len(self._ctx)
        if host_threads is None:
            host_threads = max(1, host_thread_budget() // max(1, n))
        self.host_threads = int(host_threads)
        self._blas_limit = _limit_blas(self.host_threads) if self.host_threads > 0 else None
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
        if self._blas_limit is not None:
            self._blas_limit.restore_original_limits()
            self._blas_limit = None

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
    def _stages(self, req):
        if req.scf is not None:
            if not req.stages:
                raise ValueError('an SCF job needs its stages')
            return list(req.stages)
        return [Stage(req.kgrid, req.nlcgrid, self.precision)]

    def _prepare(self, ctx, req):
        """The job's device state: its basis and every stage's grids (one
        device grid per distinct grid), resident for the whole job."""
        if self._runner is not None:
            return None
        basis = lib.DeviceBasis(ctx, req.pack)
        made = {}

        def grid(g):
            if g is None:
                return None
            key = (id(g[0]), id(g[1]))
            if key not in made:
                made[key] = lib.DeviceGrid(ctx, *g)
            return made[key]
        return basis, [(grid(st.kgrid), grid(st.nlcgrid), st.precision)
                       for st in self._stages(req)]

    def _need(self, ctx, req, state):
        """(bytes the slot holds after the job, bytes the job needs from empty),
        over its most demanding stage.  A job whose K build alone would not fit
        the budget gets a K workspace budget on its slot (wb_ctx_set_k_budget):
        its K builds then run over chunks of the grid, with the same result, in
        as little as the budget leaves; the slot's budget is cleared for the
        next job."""
        if self._runner is not None:
            held = int(getattr(ctx, 'mem_bytes', 0) or 0)
            return max(held, req.est_bytes), req.est_bytes
        basis, stages = state

        def worst():
            after, alone, big = 0, -1, None
            for kg, vg, prec in stages:
                a, b = self._job_bytes(ctx, req, basis, kg, vg, prec)
                after = max(after, a)
                if b > alone:
                    alone, big = b, (kg, prec)
            return after, alone, big
        # the job's own device state (its basis with the operator records its
        # K builds add, every stage's grid) is held while it runs, on top of
        # the slot's workspaces
        grids_seen = {}
        for kg, vg, _ in stages:
            for g in (kg, vg):
                if g is not None:
                    grids_seen[id(g)] = g
        state_bytes = basis.device_bytes(len(self.omegas)) + sum(
            g.device_bytes() for g in grids_seen.values())
        room = self.budget - state_bytes         # what the workspaces may take
        ctx.set_k_budget(0)
        after, alone, big = worst()
        if alone > room and big is not None and big[0] is not None:
            arena = ctx.k_plan(basis, big[0], len(self.omegas), big[1])[0]
            kb = arena - (alone - room)
            while kb > 0:
                ctx.set_k_budget(kb)
                try:
                    after, alone, _ = worst()
                except GpuOOM:                   # below one split's workspace
                    break
                if alone <= room:
                    req.k_budget = kb
                    break
                kb -= max(alone - room, arena // 64)
            else:
                ctx.set_k_budget(0)
        req.state_bytes = int(state_bytes)
        req.est_bytes = int(alone + state_bytes)
        return int(after + state_bytes), int(alone + state_bytes)

    def _job_bytes(self, ctx, req, basis, kg, vg, precision):
        job = lib.CycleJob()
        job.basis = basis.h.value
        job.flags = lib.JOB_SEQUENTIAL if req._shared else 0
        job.kgrid = kg.h.value if kg is not None else None
        job.nlcgrid = vg.h.value if vg is not None else None
        after, alone = ctypes.c_longlong(), ctypes.c_longlong()
        lib.check(lib.load().wb_job_bytes(ctx.h, ctypes.byref(job), len(self.omegas),
                                          lib.PREC[precision], ctypes.byref(after),
                                          ctypes.byref(alone)))
        return int(after.value), int(alone.value)

    def _run(self, ctx, req, state):
        if self._runner is not None:
            return self._runner(ctx, req)
        return self._run_device(ctx, req, state)

    def _cycle(self, ctx, req, basis, kg, vg, precision, dm):
        """One cycle's device work (K build and VV10 concurrent on the device,
        wb_cycle, when the job started alone; one after the other when it
        shares the device): (k, nlc, vmat)."""
        nao = basis.nao
        om = lib.f64(self.omegas)
        job = lib.CycleJob()
        job.basis = basis.h.value
        job.flags = lib.JOB_SEQUENTIAL if req._shared else 0
        k = nlc = vmat = None
        if kg is not None:
            qd = lib.f64(dm if req.q is None else req.q @ dm)
            k = numpy.empty((len(om), nao, nao))
            job.kgrid, job.qd, job.k = kg.h.value, qd.ctypes.data, k.ctypes.data
        if vg is not None:
            nlc, vmat = numpy.empty(2), numpy.empty((nao, nao))
            job.nlcgrid, job.dm = vg.h.value, dm.ctypes.data
            job.nlc, job.vmat = nlc.ctypes.data, vmat.ctypes.data
        lib.check(lib.load().wb_cycle(ctx.h, ctypes.byref(job), len(om), lib.ptr(om),
                                      self.k_tol, self.nlc[0], self.nlc[1], self.vv_tol,
                                      lib.PREC[precision]))
        return k, nlc, vmat

    def _run_device(self, ctx, req, state=None):
        basis, stages = state if state is not None else self._prepare(ctx, req)
        dm = lib.f64(req.dm)
        if req.scf is None:
            kg, vg, prec = stages[0]
            out = {}
            for _ in range(max(1, int(req.cycles))):
                k, nlc, vmat = self._cycle(ctx, req, basis, kg, vg, prec, dm)
                if k is not None:
                    out['k'] = k
                if nlc is not None:
                    out['nlc'], out['vmat'] = nlc, vmat
                if req.update is not None:
                    dm = lib.f64(req.update(dm, out.get('k'), out.get('nlc'), out.get('vmat')))
            out['dm'] = dm
            return out
        # staged SCF to convergence
        specs = self._stages(req)
        si, in_stage, e_last = 0, 0, None
        history = []
        converged = False
        e = None
        for cycle in range(int(req.max_cycles)):
            kg, vg, prec = stages[si]
            k, nlc, vmat = self._cycle(ctx, req, basis, kg, vg, prec, dm)
            dm_next, e, gnorm = req.scf(dm, k, nlc, vmat)
            de = float('inf') if e_last is None else abs(e - e_last)
            history.append({'cycle': cycle, 'stage': si, 'precision': prec, 'e': float(e),
                            'de': de, 'gnorm': float(gnorm), 't': time.perf_counter()})
            e_last = e
            in_stage += 1
            last = si == len(stages) - 1
            if last and de < req.conv_tol and gnorm < req.conv_tol_grad:
                converged = True
                break
            st = specs[si]
            if not last and ((de < st.tau_e and gnorm < st.tau_g) or in_stage >= st.max_cycles):
                si, in_stage, e_last = si + 1, 0, None       # energies of another grid differ
                reset = getattr(req.scf, 'reset', None)
                if reset is not None:
                    reset()
            dm = lib.f64(dm_next)
        return {'e_tot': float(e), 'dm': dm, 'converged': converged, 'cycles': len(history),
                'history': history, 'k': k, 'nlc': nlc, 'vmat': vmat}
