"""Scheduling logic of the continuous-batching loop (serving.py) on the CPU:
slots fill as requests arrive, admission by estimated HBM footprint against
the budget (an idle slot releases its workspaces when the next request would
not fit), over-budget and failing jobs reach their callers without stopping
the loop.  The device runner is replaced by a host stand-in with the same
interface; the device path is covered by tests/test_gpu_serving.py."""
import threading
import time

import pytest

from pyscf_wb97mv_fast_amd.errors import GpuOOM
from pyscf_wb97mv_fast_amd.serving import CycleServer, Request


class FakeCtx:
    def __init__(self, hold=0):
        self.hold = hold
        self.mem_bytes = 0
        self.trims = 0

    def trim(self):
        self.trims += 1
        self.mem_bytes = 0


def _runner(seconds=0.05):
    def run(ctx, req):
        if req.tag == 'boom':
            raise ValueError('job failed')
        time.sleep(seconds)
        ctx.mem_bytes = max(ctx.mem_bytes, ctx.hold or req.est_bytes)
        return req.tag
    return run


def _req(tag, est):
    return Request(pack=None, tag=tag, est_bytes=est)


def test_slots_fill_and_results_arrive():
    srv = CycleServer(hbm_budget=10_000, runner=_runner(0.2), contexts=[FakeCtx() for _ in range(3)])
    futs = [srv.submit(_req(i, 100)) for i in range(6)]
    assert [f.result(timeout=10) for f in futs] == list(range(6))
    srv.close()
    st = srv.stats()
    assert st['jobs'] == 6 and st['max_active'] == 3
    assert st['p99_s'] >= st['p50_s'] > 0


def test_idle_slot_releases_memory_a_waiting_job_needs():
    # slot 0 ends up holding 1000 of a 1500 budget and goes idle; a job of
    # 800 taken by slot 1 starts only after slot 0 releases its memory
    ctxs = [FakeCtx(hold=1000), FakeCtx(hold=800)]
    srv = CycleServer(hbm_budget=1500, runner=_runner(0.01), contexts=ctxs)
    srv.submit(_req('a', 1000)).result(timeout=10)
    time.sleep(0.05)
    futs = [srv.submit(_req(t, 800)) for t in 'bcd']
    assert [f.result(timeout=10) for f in futs] == list('bcd')
    srv.close()
    assert srv.stats()['jobs'] == 4


def test_budget_serializes_jobs_in_arrival_order():
    order = []
    lock = threading.Lock()

    def run(ctx, req):
        with lock:
            order.append(req.tag)
        time.sleep(0.02)
        ctx.mem_bytes = req.est_bytes
        return req.tag

    # each job needs 100 of a 150 budget: never two at once, even on 3 slots
    srv = CycleServer(hbm_budget=150, runner=run, contexts=[FakeCtx() for _ in range(3)])
    futs = [srv.submit(_req(i, 100)) for i in range(5)]
    for f in futs:
        f.result(timeout=10)
    srv.close()
    assert srv.max_active == 1
    assert order == list(range(5))


def test_over_budget_and_failing_jobs_reach_the_caller():
    srv = CycleServer(hbm_budget=1000, runner=_runner(0.01), contexts=[FakeCtx() for _ in range(2)])
    big = srv.submit(_req('big', 5000))
    boom = srv.submit(_req('boom', 10))
    ok = srv.submit(_req('ok', 10))
    with pytest.raises(GpuOOM):
        big.result(timeout=10)
    with pytest.raises(ValueError):
        boom.result(timeout=10)
    assert ok.result(timeout=10) == 'ok'
    srv.close()
    assert srv.stats()['jobs'] == 1
    with pytest.raises(RuntimeError):
        srv.submit(_req('late', 10))


def test_idle_slots_trim_to_admit_the_next_request():
    # two slots each end up holding 1000 of a 2000 budget; the next job (1200)
    # fits only after both idle slots release their workspaces
    ctxs = [FakeCtx(hold=1000), FakeCtx(hold=1000)]
    srv = CycleServer(hbm_budget=2000, runner=_runner(0.2), contexts=ctxs)
    a, b = srv.submit(_req('a', 1000)), srv.submit(_req('b', 1000))
    a.result(timeout=10), b.result(timeout=10)
    assert srv.max_active == 2
    assert srv.submit(_req('c', 1200)).result(timeout=10) == 'c'
    srv.close()
    assert sum(c.trims for c in ctxs) >= 1
    assert srv.stats()['jobs'] == 3
