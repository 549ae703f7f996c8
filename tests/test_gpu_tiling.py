"""K builds under a workspace budget (wb_ctx_set_k_budget): the grid runs in
chunks of whole K-GEMM point splits and K is accumulated chunk by chunk in
split order, so a tiled build equals the one-chunk build bit for bit, in
both precisions, for both operators, with screening and primitive tails, and
for a def2-TZVP molecule with f shells.  A budget below one split's
workspace is an out-of-memory error (the hooks' retry signal)."""
import numpy
import pytest

from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.errors import GpuOOM
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu
OM = numpy.array([0.0, 0.3])


@pytest.fixture(scope='module')
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


def _k(ctx, b, g, dm, prec):
    k = numpy.empty((2, b.nao, b.nao))
    lib.check(lib.load().wb_k_build(ctx.h, b.h, g.h, lib.ptr(dm), 2, lib.ptr(OM), 1e-11,
                                    lib.PREC[prec], lib.ptr(k)))
    return k


@pytest.mark.parametrize('system', ['water8', 'drug10_tzvp'])
def test_tiled_k_equals_untiled_bit_for_bit(ctx, system):
    if system == 'water8':
        mol = mole.Mole(geometry.water_cluster(8, seed=0), 'def2-svp')
    else:
        mol = mole.Mole(geometry.drug_like(10, seed=2), 'def2-tzvp')
    g = grids.Grids(mol, level=1, device=ctx).build()
    dm = guess.huckel_density(mol)
    b = lib.DeviceBasis(ctx, pack_mol(mol))
    gr = lib.DeviceGrid(ctx, g.coords, g.weights)
    try:
        for prec in ('mixed', 'fp64'):
            ctx.set_k_budget(0)
            full, n1, _ = ctx.k_plan(b, gr, 2, prec)
            assert n1 == 1
            ref = _k(ctx, b, gr, dm, prec)
            for frac in (2.5, 7.0):
                ctx.set_k_budget(int(full / frac))
                total, nchunk, csub = ctx.k_plan(b, gr, 2, prec)
                print('%s %s: budget 1/%.1f -> %d chunks of %d sub-blocks, %d of %d bytes'
                      % (system, prec, frac, nchunk, csub, total, full), flush=True)
                assert nchunk > 1 and total <= full / frac
                ctx.trim()
                got = _k(ctx, b, gr, dm, prec)
                assert ctx.mem_bytes <= full / frac + 8 * 2 * b.nao ** 2 * 3 + 4096
                assert numpy.array_equal(got, ref), (prec, frac, numpy.abs(got - ref).max())
        ctx.set_k_budget(1 << 20)                     # below one split's workspace
        with pytest.raises(GpuOOM):
            _k(ctx, b, gr, dm, 'mixed')
    finally:
        ctx.set_k_budget(0)
        ctx.trim()
