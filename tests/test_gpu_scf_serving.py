"""SCF jobs to convergence through the continuous-batching server on the
device: water clusters and a drug-like chain arrive over time, each runs
the staged schedule (level-0 grids in the mixed form, level-1 grids in the
mixed form, level-1 grids in FP64 until |dE| < 1e-9 Ha and |g| < 1e-6) with
the model SCF step (systems/scf.py) and converges; its final energy equals,
within 1e-8 Ha, that of the same SCF run entirely in FP64 on the level-1
grids; a job's results in the mix equal its results alone bit for bit; and
the server reports p50/p99 time to the converged energy."""
import time

import numpy
import pytest

from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.serving import CycleServer, Request, Stage, staged
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole, scf

pytestmark = pytest.mark.gpu
HOST_THREADS = 4


@pytest.fixture(scope='module')
def jobs():
    pctx = lib.Context(0)
    out = []
    for mol in (mole.Mole(geometry.water_cluster(4, seed=1), 'def2-svp'),
                mole.Mole(geometry.drug_like(10, seed=2), 'def2-svp'),
                mole.Mole(geometry.water_cluster(8, seed=0), 'def2-svp')):
        g0 = grids.Grids(mol, level=0, device=pctx).build()
        g1 = grids.Grids(mol, level=1, device=pctx).build()
        coarse, fine = (g0.coords, g0.weights), (g1.coords, g1.weights)
        out.append((mol, coarse, fine, guess.huckel_density(mol)))
    pctx.close()
    return out


def _req(mol, coarse, fine, dm, fp64_only=False, tag=None):
    st = ([Stage(fine, fine, 'fp64', 0.0, 0.0, 10 ** 6)] if fp64_only
          else staged((coarse, coarse), (fine, fine)))
    return Request(pack=pack_mol(mol), dm=dm, scf=scf.ModelSCF(mol), stages=st, conv_tol=1e-9,
                   conv_tol_grad=1e-6, max_cycles=80, tag=tag)


@pytest.mark.timeout(600)
def test_staged_scf_jobs_converge_in_the_mix(jobs):
    # the host steps' BLAS rounds differently with another thread count:
    # both servers give each slot the same number of threads
    srv = CycleServer(slots=2, hbm_budget=200e9, host_threads=HOST_THREADS)
    futs = []
    for i, j in enumerate(jobs):
        futs.append(srv.submit(_req(*j, tag=i)))
        time.sleep(0.05)
    res = [f.result(timeout=500) for f in futs]
    srv.close()
    st = srv.stats()
    print('staged: %d jobs, %.2f jobs/s, time to converged energy p50 %.2f s, p99 %.2f s, '
          'at most %d at once' % (st['jobs'], st['jobs_per_s'], st['p50_s'], st['p99_s'],
                                  st['max_active']), flush=True)
    assert st['jobs'] == len(jobs) and st['max_active'] >= 2
    ref_srv = CycleServer(slots=1, hbm_budget=200e9, host_threads=HOST_THREADS)
    try:
        for i, (j, r) in enumerate(zip(jobs, res)):
            stages = [h['stage'] for h in r['history']]
            print('job %d: %d cycles (%s), E %.10f' % (i, r['cycles'], [stages.count(s) for s in
                                                                      (0, 1, 2)], r['e_tot']),
                  flush=True)
            assert r['converged'] and stages[-1] == 2 and stages == sorted(stages)
            ref = ref_srv.submit(_req(*j, fp64_only=True)).result(timeout=500)
            print('   all-FP64 level 1: %d cycles, E %.10f, dE %.2e' % (
                ref['cycles'], ref['e_tot'], r['e_tot'] - ref['e_tot']), flush=True)
            assert ref['converged'] and abs(r['e_tot'] - ref['e_tot']) < 1e-8
        alone = ref_srv.submit(_req(*jobs[1])).result(timeout=500)
        assert alone['e_tot'] == res[1]['e_tot'] and numpy.array_equal(alone['dm'], res[1]['dm'])
        assert alone['cycles'] == res[1]['cycles']
    finally:
        ref_srv.close()
