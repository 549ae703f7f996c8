"""The largest water cluster of the benchmark sets, (H2O)64 def2-SVP on its
full level-1 grid (1536 AOs, ~5e5 points): the mixed form against the FP64
form of the same kernels, through size-independent properties -- the
per-cycle energy errors within the 1e-6 Ha budget, the electron count, the
VV10 potential matrix symmetric -- and the batch lanes' memory gating at this size
(two jobs of this size on two lanes give the single-call results)."""
import numpy
import pytest

from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.gpu import batch
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu
OM = numpy.array([0.0, 0.3])


@pytest.fixture(scope='module')
def big():
    ctx = lib.Context(0)
    mol = mole.Mole(geometry.water_cluster(64, seed=0), 'def2-svp')
    g = grids.Grids(mol, level=1, device=ctx).build()
    dm = guess.huckel_density(mol)
    b = lib.DeviceBasis(ctx, pack_mol(mol))
    gr = lib.DeviceGrid(ctx, g.coords, g.weights)
    print('(H2O)64: nao %d, %d points' % (b.nao, g.size), flush=True)
    return ctx, b, gr, dm


def _cycle(ctx, b, gr, dm, prec):
    k = numpy.empty((2, b.nao, b.nao))
    lib.check(lib.load().wb_k_build(ctx.h, b.h, gr.h, lib.ptr(dm), 2, lib.ptr(OM), 1e-11,
                                    lib.PREC[prec], lib.ptr(k)))
    o2, v = numpy.empty(2), numpy.empty((b.nao, b.nao))
    lib.check(lib.load().wb_nlc_vxc(ctx.h, b.h, gr.h, lib.ptr(dm), 6.0, 0.01, 0.0,
                                    lib.PREC[prec], lib.ptr(o2), lib.ptr(v)))
    return k, o2, v


def test_water64_mixed_within_budget_of_fp64(big):
    ctx, b, gr, dm = big
    k32, e32, v32 = _cycle(ctx, b, gr, dm, 'mixed')
    k64, e64, v64 = _cycle(ctx, b, gr, dm, 'fp64')
    dK = k32 - k64
    dEx = -0.25 * float(numpy.einsum('ij,ji->', dm, 0.15 * dK[0] + 0.85 * dK[1]))
    print('dE_x %.2e dE_nl %.2e max|dK| %.2e' % (dEx, e32[1] - e64[1], abs(dK).max()), flush=True)
    assert abs(dEx) <= 1e-6
    assert abs(e32[1] - e64[1]) <= 1e-7
    assert abs(e32[0] - e64[0]) <= 1e-9 * abs(e64[0])
    assert numpy.array_equal(v32, v32.T)
    # the symmetrized K (what get_k_only returns) carries the same energy
    ks = 0.5 * (k32 + k32.transpose(0, 2, 1))
    assert abs(numpy.einsum('ij,oji->o', dm, ks - k32)).max() <= 1e-9


def test_water64_two_lanes_equal_single_calls(big):
    ctx, b, gr, dm = big
    ref = _cycle(ctx, b, gr, dm, 'mixed')
    dm2 = dm * 0.999
    ref2 = _cycle(ctx, b, gr, dm2, 'mixed')
    ctx.trim()
    jobs = [batch.CycleJob(b, gr, dm, gr, dm), batch.CycleJob(b, gr, dm2, gr, dm2)]
    batch.run_batch(ctx, jobs, OM, lanes=2)
    for j, r in zip(jobs, (ref, ref2)):
        assert numpy.array_equal(j.k, r[0]) and numpy.array_equal(j.nlc, r[1])
        assert numpy.array_equal(j.vmat, r[2])
    ctx.trim()
