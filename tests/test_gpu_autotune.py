"""The K-build autotuner (wb_ctx_autotune, gpu/autotune.py): it keeps the
fastest configuration it timed (never slower than the default it also
timed), the tuned K stays within the mixed form's error of the FP64 K, and
the per-size-class cache reapplies a tuning without timing again.  The
waves per sub-block of each (la, lb) class change only the FP64 summation
order of the G copies."""
import numpy
import pytest

from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.errors import GpuError
from pyscf_wb97mv_fast_amd.gpu import autotune
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu
OM = numpy.array([0.0, 0.3])


@pytest.fixture(scope='module')
def case():
    mol = mole.Mole(geometry.water_cluster(6, seed=4), 'def2-svp')
    g = grids.Grids(mol, level=0).build()
    return pack_mol(mol), g, guess.huckel_density(mol)


def _k(ctx, b, gr, dm, prec='mixed'):
    k = numpy.empty((2, b.nao, b.nao))
    lib.check(lib.load().wb_k_build(ctx.h, b.h, gr.h, lib.ptr(dm), 2, lib.ptr(OM), 1e-11,
                                    lib.PREC[prec], lib.ptr(k)))
    return k


def test_autotune_keeps_fastest_and_accuracy(case):
    pk, g, dm = case
    ctx = lib.Context(0)
    b = lib.DeviceBasis(ctx, pk)
    gr = lib.DeviceGrid(ctx, g.coords, g.weights)
    k_def = _k(ctx, b, gr, dm)
    k64 = _k(ctx, b, gr, dm, 'fp64')
    rep = autotune.autotune(ctx, b, gr, dm, OM, reps=2)
    print(rep, flush=True)
    assert rep['configs'] >= 5
    assert rep['best_ms'] <= rep['default_ms']
    assert autotune.get_tuning(ctx) == (rep['ncopy'], rep['fuse_mask'], rep['gemm_tiles'])
    assert rep['configs'] >= 7 and rep['gemm_tiles'] in (0, 3, 4)
    k_tuned = _k(ctx, b, gr, dm)
    err_def = numpy.abs(k_def - k64).max()
    err_tuned = numpy.abs(k_tuned - k64).max()
    print('max|dK| default %.3e tuned %.3e' % (err_def, err_tuned), flush=True)
    assert err_tuned <= 2 * err_def + 1e-10
    assert len(rep['class_waves']) == 10 and rep['class_waves'] == autotune.get_class_waves(ctx)
    # restoring the default configuration restores the default K bit for bit
    autotune.set_tuning(ctx, 8, 0xffffffff, 0)
    autotune.set_class_waves(ctx, None)
    assert numpy.array_equal(_k(ctx, b, gr, dm), k_def)
    # either GEMM tile shape gives K within rounding of the default
    for tiles in (3, 4):
        autotune.set_tuning(ctx, 8, 0xffffffff, tiles)
        kt = _k(ctx, b, gr, dm)
        print('tiles %d: max|dK| %.3e' % (tiles, numpy.abs(kt - k_def).max()), flush=True)
        assert numpy.abs(kt - k_def).max() <= 1e-9 * numpy.abs(k_def).max()
    with pytest.raises(GpuError):
        autotune.set_tuning(ctx, 0, 0)
    with pytest.raises(GpuError):
        autotune.set_tuning(ctx, 8, 0, 5)


def test_tuner_cache_reuses_the_size_class(case, tmp_path):
    pk, g, dm = case
    ctx = lib.Context(0)
    b = lib.DeviceBasis(ctx, pk)
    gr = lib.DeviceGrid(ctx, g.coords, g.weights)
    t = autotune.Tuner(reps=1)
    k1, e1 = t.apply(ctx, b, gr, dm, OM)
    autotune.set_tuning(ctx, 3, 0)
    k2, e2 = t.apply(ctx, b, gr, dm, OM)             # cached: no timing, tuning restored
    assert k1 == k2 and e1 is e2
    assert autotune.get_tuning(ctx) == (e1['ncopy'], e1['fuse_mask'], e1['gemm_tiles'])
    path = tmp_path / 'tuning.json'
    t.save(str(path))
    t2 = autotune.Tuner().load(str(path))
    assert t2.table == t.table


def test_class_waves_change_only_the_summation_order(case):
    pk, g, dm = case
    ctx = lib.Context(0)
    b = lib.DeviceBasis(ctx, pk)
    gr = lib.DeviceGrid(ctx, g.coords, g.weights)
    ref = _k(ctx, b, gr, dm, 'fp64')
    for waves in ([1] * 10, [0, 2, 4, 1, 8, 3, 1, 2, 1, 64]):
        autotune.set_class_waves(ctx, waves)
        assert autotune.get_class_waves(ctx) == waves
        k = _k(ctx, b, gr, dm, 'fp64')
        print(waves, 'max|dK| %.2e' % numpy.abs(k - ref).max(), flush=True)
        assert numpy.abs(k - ref).max() <= 1e-12 * numpy.abs(ref).max()
    autotune.set_class_waves(ctx, None)
    assert autotune.get_class_waves(ctx) == [0] * 10
    assert numpy.array_equal(_k(ctx, b, gr, dm, 'fp64'), ref)
    with pytest.raises(GpuError):
        autotune.set_class_waves(ctx, [65] + [0] * 9)
