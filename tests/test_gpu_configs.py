"""The many-molecule configurations at their stated sizes:

* 64 water clusters, n ~ U{4..64} (seeded), def2-SVP, level-1 grids,
  extended-Hueckel densities, all 128 K builds (1/r and erf) and the VV10
  terms in one wb_cycle_batch call: every batched result equals the
  single-molecule call bit for bit (checked on eight of them, smallest to
  largest), and every molecule's mixed-form energies are within the per-cycle
  budget of the FP64 batch (|dE_x| <= 1e-6, |dE_nl| <= 1e-7 Ha);
* a serving mix of water clusters (8 and 80 molecules, def2-SVP) and
  drug-like chains (10 and 60 heavy atoms in def2-TZVP, 30 in def2-SVP) with
  seeded Poisson arrivals through the continuous-batching server under a
  200 GB device budget: every job is admitted within the budget (the
  60-heavy-atom def2-TZVP job's FP64 K build runs tiled to fit), each job's
  results in the mix equal its results alone, bit for bit, and the mixed
  form's energies are within the budget of the FP64 form's.
"""
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
    c = lib.Context(0)
    yield c
    c.close()


def _system(mol, pctx):
    g = grids.Grids(mol, level=1, device=pctx).build()
    return pack_mol(mol), g.coords, g.weights, guess.huckel_density(mol)


def _energies(dm, k, nlc):
    ex = -0.25 * float(numpy.einsum('ij,ji->', dm, 0.15 * k[0] + 0.85 * k[1]))
    return ex, float(nlc[1])


@pytest.mark.timeout(600)
def test_batch_of_64_clusters(ctx):
    rng = numpy.random.default_rng(11)
    sizes = rng.integers(4, 65, 64)
    systems = [_system(mole.Mole(geometry.water_cluster(int(n), seed=11 + i), 'def2-svp'), ctx)
               for i, n in enumerate(sizes)]
    print('64 clusters, n %d..%d, %d grid points' % (sizes.min(), sizes.max(),
                                                   sum(len(s[2]) for s in systems)), flush=True)
    bases = [lib.DeviceBasis(ctx, s[0]) for s in systems]
    grs = [lib.DeviceGrid(ctx, s[1], s[2]) for s in systems]

    def jobs():
        return [batch.CycleJob(b, g, s[3], g, s[3]) for b, g, s in zip(bases, grs, systems)]
    mixed = batch.run_batch(ctx, jobs(), OM, precision='mixed', lanes=1)
    print('mixed batch done', flush=True)
    fp64 = batch.run_batch(ctx, jobs(), OM, precision='fp64', lanes=1)
    print('fp64 batch done', flush=True)
    worst = [0.0, 0.0]
    for s, jm, jd in zip(systems, mixed, fp64):
        em, ed = _energies(s[3], jm.k, jm.nlc), _energies(s[3], jd.k, jd.nlc)
        worst = [max(worst[0], abs(em[0] - ed[0])), max(worst[1], abs(em[1] - ed[1]))]
    print('max |dE_x| %.2e, max |dE_nl| %.2e' % tuple(worst), flush=True)
    assert worst[0] <= 1e-6 and worst[1] <= 1e-7
    order = numpy.argsort(sizes)
    for i in order[::9]:
        b, g, s = bases[i], grs[i], systems[i]
        k = numpy.empty((2, b.nao, b.nao))
        lib.check(lib.load().wb_k_build(ctx.h, b.h, g.h, lib.ptr(s[3]), 2, lib.ptr(OM), 1e-11,
                                        lib.PREC['mixed'], lib.ptr(k)))
        o2, v = numpy.empty(2), numpy.empty((b.nao, b.nao))
        lib.check(lib.load().wb_nlc_vxc(ctx.h, b.h, g.h, lib.ptr(s[3]), 6.0, 0.01, 0.0,
                                        lib.PREC['mixed'], lib.ptr(o2), lib.ptr(v)))
        assert numpy.array_equal(mixed[i].k, k) and numpy.array_equal(mixed[i].nlc, o2)
        assert numpy.array_equal(mixed[i].vmat, v)
    ctx.trim()


def _update(dm, k, nlc, vmat):
    # a deterministic host step standing in for the caller's SCF update
    return dm - 1e-3 * (k[0] + k[0].T) + 1e-4 * vmat


@pytest.mark.timeout(600)
def test_serving_mix_with_tzvp_drug_chains(ctx):
    mols = [mole.Mole(geometry.water_cluster(8, seed=3), 'def2-svp'),
            mole.Mole(geometry.drug_like(10, seed=2), 'def2-tzvp'),
            mole.Mole(geometry.water_cluster(80, seed=4), 'def2-svp'),
            mole.Mole(geometry.drug_like(30, seed=5), 'def2-svp'),
            mole.Mole(geometry.drug_like(60, seed=8), 'def2-tzvp')]
    systems = [_system(m, ctx) for m in mols]
    print('serving mix: nao %s' % [m.nao for m in mols], flush=True)
    gaps = numpy.random.default_rng(5).exponential(0.5, len(systems))

    def reqs():
        return [Request(pack=s[0], kgrid=(s[1], s[2]), nlcgrid=(s[1], s[2]), dm=s[3], cycles=2,
                        update=_update, tag=i) for i, s in enumerate(systems)]
    out = {}
    for precision in ('mixed', 'fp64'):
        srv = CycleServer(slots=3, hbm_budget=200e9, precision=precision)
        futs = []
        import time
        for gap, r in zip(gaps, reqs()):
            time.sleep(float(gap))
            futs.append((r, srv.submit(r)))
        out[precision] = [(r, f.result(timeout=500)) for r, f in futs]
        srv.close()
        st = srv.stats()
        print(precision, {k: st[k] for k in ('jobs', 'jobs_per_s', 'p50_s', 'p99_s',
                                             'max_active')}, flush=True)
        for r, _ in futs:
            print('  job %d: %.1f GB on the device, K workspace budget %s' % (
                r.tag, r.est_bytes / 1e9, '%.1f GB' % (r.k_budget / 1e9) if r.k_budget else
                'none'), flush=True)
            assert r.est_bytes <= 200e9
        assert st['jobs'] == len(systems)
        if precision == 'mixed':                  # FP64 jobs of 186-200 GB run one at a time
            assert st['max_active'] >= 2
    ref_srv = CycleServer(slots=1, hbm_budget=200e9, precision='mixed')
    alone_ctx = lib.Context(0)
    try:
        for r, got in out['mixed']:
            ref = ref_srv._run_device(alone_ctx, r)
            for key in ('k', 'nlc', 'vmat', 'dm'):
                assert numpy.array_equal(got[key], ref[key]), (r.tag, key)
            alone_ctx.trim()
    finally:
        ref_srv.close()
        alone_ctx.close()
    for (r, m), (_, d) in zip(out['mixed'], out['fp64']):
        em, ed = _energies(m['dm'], m['k'], m['nlc']), _energies(m['dm'], d['k'], d['nlc'])
        print('job %d: dE_x %.2e dE_nl %.2e' % (r.tag, em[0] - ed[0], em[1] - ed[1]))
        assert abs(em[0] - ed[0]) <= 1e-6 and abs(em[1] - ed[1]) <= 1e-7
