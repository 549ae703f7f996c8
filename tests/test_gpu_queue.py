"""The exchange kernels of many molecules' K builds in one persistent launch
per shell-pair class (wb_cycle_queue, gpu/batch.run_queue): every molecule's
K equals its own wb_k_build bit for bit, and its VV10 terms equal
wb_nlc_vxc's, in both precision forms, for a mix of def2-SVP water clusters,
a def2-TZVP drug-like chain (every (la, lb) class up to (f, f)) and a
2064-AO cluster whose G is cleared by the first class launch (past the rows
the written-shell masks cover); under a workspace budget the molecules run
in several groups and the largest alone, tiled, with the same results."""
import types

import numpy
import pytest

from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.gpu import batch
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu
OM = numpy.array([0.0, 0.3])


@pytest.fixture(scope='module')
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


def _big_cluster():
    # 64 points near the atoms: the kernel path of a 2064-AO basis, quickly
    mol = mole.Mole(geometry.water_cluster(86, seed=4), 'def2-svp')
    rng = numpy.random.default_rng(0)
    xyz = mol.atom_coords()
    pts = xyz[rng.choice(len(xyz), 64)] + rng.normal(scale=0.8, size=(64, 3))
    a = rng.normal(scale=0.01, size=(mol.nao, mol.nao))
    return pack_mol(mol), pts, rng.uniform(0.01, 0.1, 64), a + a.T


@pytest.fixture(scope='module')
def systems(ctx):
    out = []
    for n, seed in ((2, 0), (9, 1), (3, 2), (5, 3), (14, 5)):
        mol = mole.Mole(geometry.water_cluster(n, seed=seed), 'def2-svp')
        g = grids.Grids(mol, level=0).build()
        out.append((pack_mol(mol), g.coords, g.weights, guess.huckel_density(mol)))
    mol = mole.Mole(geometry.drug_like(10, seed=2), 'def2-tzvp')
    g = grids.Grids(mol, level=0).build()
    out.append((pack_mol(mol), g.coords, g.weights, guess.huckel_density(mol)))
    out.append(_big_cluster())
    print('queue systems: grid points %s' % [len(s[1]) for s in out], flush=True)
    return out


def _handles(ctx, systems):
    bases = [lib.DeviceBasis(ctx, s[0]) for s in systems]
    grs = [lib.DeviceGrid(ctx, s[1], s[2]) for s in systems]
    return bases, grs


def _alone(ctx, b, g, dm, precision, vv10=True):
    k = numpy.empty((2, b.nao, b.nao))
    lib.check(lib.load().wb_k_build(ctx.h, b.h, g.h, lib.ptr(dm), 2, lib.ptr(OM), 1e-11,
                                    lib.PREC[precision], lib.ptr(k)))
    if not vv10:
        return k, None, None
    o2, v = numpy.empty(2), numpy.empty((b.nao, b.nao))
    lib.check(lib.load().wb_nlc_vxc(ctx.h, b.h, g.h, lib.ptr(dm), 6.0, 0.01, 0.0,
                                    lib.PREC[precision], lib.ptr(o2), lib.ptr(v)))
    return k, o2, v


def _jobs(bases, grs, systems):
    jobs = []
    for i, (b, g, s) in enumerate(zip(bases, grs, systems)):
        vv = i != len(systems) - 1            # the 64-point cluster: K only
        jobs.append(batch.CycleJob(b, g, s[3], g if vv else None, s[3] if vv else None))
    # the same basis twice (K only, VV10 only): shared operator records
    jobs.append(batch.CycleJob(bases[1], kgrid=grs[1], qd=systems[1][3]))
    jobs.append(batch.CycleJob(bases[1], nlcgrid=grs[1], dm=systems[1][3]))
    return jobs


def _check(ctx, jobs, precision):
    for i, j in enumerate(jobs):
        dm = j.qd if j.kgrid is not None else j.dm
        grid = j.kgrid if j.kgrid is not None else j.nlcgrid
        k, o2, v = _alone(ctx, j.basis, grid, dm, precision, vv10=j.nlcgrid is not None)
        if j.kgrid is not None:
            assert numpy.array_equal(j.k, k), (i, precision, 'k')
        if j.nlcgrid is not None:
            assert numpy.array_equal(j.nlc, o2) and numpy.array_equal(j.vmat, v), (i, precision)
        assert j.ms is not None and j.ms > 0


@pytest.mark.parametrize('precision', ['mixed', 'fp64'])
def test_queue_equals_single_calls(ctx, systems, precision):
    bases, grs = _handles(ctx, systems)
    jobs = batch.run_queue(ctx, _jobs(bases, grs, systems), OM, precision=precision)
    print('%s queue: completion ms %s' % (precision, ['%.1f' % j.ms for j in jobs]), flush=True)
    _check(ctx, jobs, precision)


def test_queue_groups_under_a_budget(systems):
    c = lib.Context(0)
    try:
        bases, grs = _handles(c, systems)
        plans = [c.k_plan(b, g, 2, 'mixed')[0] for b, g in zip(bases, grs)]
        big = max(plans)
        # groups of about two molecules; the largest runs alone, tiled
        c.set_k_budget(int(0.8 * big))
        assert c.k_plan(bases[int(numpy.argmax(plans))], grs[int(numpy.argmax(plans))])[1] > 1
        jobs = batch.run_queue(c, _jobs(bases, grs, systems), OM, precision='mixed')
        c.set_k_budget(0)
        _check(c, jobs, 'mixed')
    finally:
        c.close()


def test_queue_edges(ctx, systems):
    assert batch.run_queue(ctx, [], OM) == []
    bases, grs = _handles(ctx, systems[:1])
    s = systems[0]
    with pytest.raises(lib.GpuError, match='job 0'):
        bad = (lib.CycleJob * 1)()
        bad[0].basis = bases[0].h.value
        bad[0].kgrid = grs[0].h.value
        lib.check(lib.load().wb_cycle_queue(ctx.h, 1, bad, 2, lib.ptr(OM), 1e-11, 6.0, 0.01, 0.0,
                                            lib.PREC['mixed'], None))
    with pytest.raises(lib.Unsupported):
        batch.run_queue(ctx, [batch.CycleJob(bases[0], grs[0], s[3])], (0.0, -0.3))
    # VV10 only: no K group at all
    jobs = batch.run_queue(ctx, [batch.CycleJob(bases[0], nlcgrid=grs[0], dm=s[3])], OM)
    _, o2, v = _alone(ctx, bases[0], grs[0], s[3], 'mixed')
    assert numpy.array_equal(jobs[0].nlc, o2) and numpy.array_equal(jobs[0].vmat, v)
