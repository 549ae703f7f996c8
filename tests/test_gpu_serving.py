"""Many molecules on one GPU: the batch entry (wb_cycle_batch, gpu/batch.py)
and the continuous-batching loop (serving.py) give every molecule exactly
the results of its single-molecule calls alone (bit for bit), whatever the
mix, the lane count or the arrival order; the workspace accounting used
for admission by HBM footprint reads and releases the device memory."""
import numpy
import pytest

from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.gpu import batch
from pyscf_wb97mv_fast_amd.serving import CycleServer, Request
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu
OM = numpy.array([0.0, 0.3])


@pytest.fixture(scope='module')
def ctx():
    return lib.Context(0)


@pytest.fixture(scope='module')
def systems():
    out = []
    for n, seed in ((2, 0), (5, 1), (3, 2), (4, 3)):
        mol = mole.Mole(geometry.water_cluster(n, seed=seed), 'def2-svp')
        g = grids.Grids(mol, level=0).build()
        out.append((pack_mol(mol), g.coords, g.weights, guess.huckel_density(mol)))
    return out


def _alone(pk, coords, weights, dm):
    c = lib.Context(0)
    b = lib.DeviceBasis(c, pk)
    g = lib.DeviceGrid(c, coords, weights)
    k = numpy.empty((2, b.nao, b.nao))
    lib.check(lib.load().wb_k_build(c.h, b.h, g.h, lib.ptr(dm), 2, lib.ptr(OM), 1e-11,
                                    lib.PREC['mixed'], lib.ptr(k)))
    o2, v = numpy.empty(2), numpy.empty((b.nao, b.nao))
    lib.check(lib.load().wb_nlc_vxc(c.h, b.h, g.h, lib.ptr(dm), 6.0, 0.01, 0.0,
                                    lib.PREC['mixed'], lib.ptr(o2), lib.ptr(v)))
    return k, o2, v


@pytest.fixture(scope='module')
def alone(systems):
    return [_alone(*s) for s in systems]


def test_batch_equals_single_calls(ctx, systems, alone):
    bases = [lib.DeviceBasis(ctx, s[0]) for s in systems]
    gr = [lib.DeviceGrid(ctx, s[1], s[2]) for s in systems]
    jobs = [batch.CycleJob(b, g, s[3], g, s[3]) for b, g, s in zip(bases, gr, systems)]
    # the same basis in two jobs (K only, VV10 only): shared operator constants
    jobs.append(batch.CycleJob(bases[1], kgrid=gr[1], qd=systems[1][3]))
    jobs.append(batch.CycleJob(bases[1], nlcgrid=gr[1], dm=systems[1][3]))
    for lanes in (1, 3):
        print('lanes', lanes, flush=True)
        batch.run_batch(ctx, jobs, OM, lanes=lanes)
        for j, (k, o2, v) in zip(jobs, alone):
            assert numpy.array_equal(j.k, k)
            assert numpy.array_equal(j.nlc, o2) and numpy.array_equal(j.vmat, v)
        assert numpy.array_equal(jobs[4].k, alone[1][0])
        assert jobs[4].nlc is None
        assert numpy.array_equal(jobs[5].vmat, alone[1][2]) and jobs[5].k is None
        assert all(j.ms > 0 for j in jobs)


def test_batch_out_of_memory_job_is_retried_alone(ctx, systems, alone, monkeypatch):
    bases = [lib.DeviceBasis(ctx, s[0]) for s in systems]
    gr = [lib.DeviceGrid(ctx, s[1], s[2]) for s in systems]
    jobs = [batch.CycleJob(b, g, s[3], g, s[3]) for b, g, s in zip(bases, gr, systems)]
    monkeypatch.setenv('WB_BATCH_FAIL_ONCE', '2')
    batch.run_batch(ctx, jobs, OM, lanes=3)
    for j, (k, o2, v) in zip(jobs, alone):
        assert numpy.array_equal(j.k, k) and numpy.array_equal(j.vmat, v)


def test_batch_errors_name_the_job(ctx, systems):
    b = lib.DeviceBasis(ctx, systems[0][0])
    g = lib.DeviceGrid(ctx, systems[0][1], systems[0][2])
    jobs = [batch.CycleJob(b, g, systems[0][3])]
    with pytest.raises(Exception, match='nop'):
        batch.run_batch(ctx, jobs, omegas=[0.0] * 9)
    with pytest.raises(Exception, match='lanes'):
        batch.run_batch(ctx, jobs, OM, lanes=0)


def test_server_mix_equals_alone(systems):
    def update(dm, k, nlc, vmat):
        # a deterministic host step standing in for the caller's SCF update
        return dm - 1e-3 * (k[0] + k[0].T) + 1e-4 * vmat

    reqs = [Request(pack=s[0], kgrid=(s[1], s[2]), nlcgrid=(s[1], s[2]), dm=s[3], cycles=2,
                    update=update, tag=i) for i, s in enumerate(systems)]
    srv = CycleServer(slots=3, hbm_budget=200e9)
    futs = [srv.submit(r) for r in reqs]
    got = [f.result(timeout=120) for f in futs]
    srv.close()
    st = srv.stats()
    print(st, flush=True)
    assert st['jobs'] == len(reqs)
    ref_srv = CycleServer(slots=1, hbm_budget=200e9)
    for r, g in zip(reqs, got):
        ref = ref_srv._run_device(lib.Context(0), r)
        for key in ('k', 'nlc', 'vmat', 'dm'):
            assert numpy.array_equal(g[key], ref[key]), (r.tag, key)
    ref_srv.close()


def test_workspace_accounting(systems):
    c = lib.Context(0)
    assert c.mem_bytes == 0
    pk, coords, weights, dm = systems[1]
    srv = CycleServer(slots=1, contexts=[c])
    r = Request(pack=pk, kgrid=(coords, weights), nlcgrid=(coords, weights), dm=dm)
    out = srv.submit(r).result(timeout=120)
    srv.close()
    held = c.mem_bytes
    print('held', held, 'job bytes', r.est_bytes, 'of which its state', r.state_bytes, flush=True)
    # the job's workspace bytes (wb_job_bytes) are exactly what it allocated;
    # its basis (with the operator records its K builds add) and grids come
    # on top, at least what they hold after a K build of both operators
    assert held == r.est_bytes - r.state_bytes > 0
    b = lib.DeviceBasis(c, pk)
    g = lib.DeviceGrid(c, coords, weights)
    need0 = b.device_bytes(2)
    assert b.device_bytes(0) < need0
    kk = numpy.empty((2, b.nao, b.nao))
    lib.check(lib.load().wb_k_build(c.h, b.h, g.h, lib.ptr(lib.f64(dm)), 2,
                                    lib.ptr(numpy.array([0.0, 0.3])), 1e-11, lib.PREC['mixed'],
                                    lib.ptr(kk)))
    assert b.device_bytes(2) == b.device_bytes(0) <= need0
    assert b.device_bytes(0) + g.device_bytes() <= r.state_bytes
    del b, g
    c.trim()
    assert c.mem_bytes == 0
    k, o2, v = _alone(pk, coords, weights, dm)
    assert numpy.array_equal(out['k'], k) and numpy.array_equal(out['vmat'], v)
