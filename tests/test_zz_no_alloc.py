"""The allocation-free K build (wb_k_build_noalloc) through the C ABI against
the allocating entry it mirrors (wb_k_build_dev) on the same device inputs:
bit for bit, both precisions, two calls in a row with the second on a
smaller molecule (the shared scratch keeps the first call's contents), and
no growth of the context's workspaces.  (H2O)27 on its level-1 grid needs
more than the 256 MiB scratch in one chunk, so it runs tiled.  Sizes beyond
the scratch and operator records not yet on the device are clear errors."""
import numpy
import pytest

from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.errors import GpuError, Unsupported
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu
OM = numpy.array([0.0, 0.3])


def _system(ctx, n, level):
    mol = mole.Mole(geometry.water_cluster(n, seed=0), 'def2-svp')
    g = grids.Grids(mol, level=level, device=ctx).build()
    return (lib.DeviceBasis(ctx, pack_mol(mol)), lib.DeviceGrid(ctx, g.coords, g.weights),
            guess.huckel_density(mol))


def _call(fn, ctx, b, g, d_qd, d_k, prec):
    lib.check(fn(ctx.h, b.h, g.h, d_qd.data_ptr(), 2, lib.ptr(OM), 1e-11, lib.PREC[prec],
                 d_k.data_ptr()))
    ctx.sync()


def test_noalloc_k_build_equals_the_allocating_entry():
    import torch
    L = lib.load()
    ctx = lib.Context(0)
    dev = torch.device('cuda:0')
    try:
        for n, level in ((27, 1), (3, 0)):               # the second call is smaller
            b, g, dm = _system(ctx, n, level)
            plan = ctx.k_plan(b, g, 2, 'mixed')
            print('(H2O)%d: nao %d, %d points, one-chunk workspace %d bytes' % (n, b.nao, g.npts,
                                                                             plan[0]), flush=True)
            d_qd = torch.from_numpy(numpy.ascontiguousarray(dm)).to(dev)
            for prec in ('mixed', 'fp64'):
                ref = torch.zeros((2, b.nao, b.nao), dtype=torch.float64, device=dev)
                got = torch.full_like(ref, numpy.nan)
                torch.cuda.synchronize()
                _call(L.wb_k_build_dev, ctx, b, g, d_qd, ref, prec)
                held = ctx.mem_bytes
                _call(L.wb_k_build_noalloc, ctx, b, g, d_qd, got, prec)
                assert ctx.mem_bytes == held                    # nothing allocated
                same = torch.equal(got, ref)
                print('  %s: noalloc == dev: %s, max|K| %.3e' % (prec, same, ref.abs().max().item()),
                      flush=True)
                assert same
    finally:
        ctx.close()


def test_noalloc_rejects_what_it_cannot_do():
    import torch
    L = lib.load()
    ctx = lib.Context(0)
    dev = torch.device('cuda:0')
    try:
        b, g, dm = _system(ctx, 3, 0)
        d_qd = torch.from_numpy(numpy.ascontiguousarray(dm)).to(dev)
        out = torch.zeros((2, b.nao, b.nao), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        with pytest.raises(GpuError, match='not on the device yet'):
            _call(L.wb_k_build_noalloc, ctx, b, g, d_qd, out, 'mixed')   # no records yet
        b27, g27, dm27 = _system(ctx, 27, 1)
        q27 = torch.from_numpy(numpy.ascontiguousarray(dm27)).to(dev)
        o27 = torch.zeros((2, b27.nao, b27.nao), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        _call(L.wb_k_build_dev, ctx, b27, g27, q27, o27, 'fp64')
        lib.check(L.wb_ctx_set_copies(ctx.h, 64))        # 64 FP64 G copies of one split: too big
        with pytest.raises(Unsupported, match='256 MiB'):
            _call(L.wb_k_build_noalloc, ctx, b27, g27, q27, o27, 'fp64')
    finally:
        ctx.close()
