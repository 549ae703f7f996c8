"""Per-cycle HIP graph (wb_cycle_graph_*, K build and VV10 on two forked
branches): every replay equals the eager calls (wb_k_build_dev +
wb_nlc_vxc_dev, in sequence) on the same density bit for bit, with eager work
and synchronizations between the replays, in both precisions, and the
graph's workspaces survive the owner context's buffers regrowing.  The eager
concurrent cycle (wb_cycle_dev, VV10 on a second stream) equals the
sequential calls bit for bit too."""
import numpy
import pytest

pytestmark = pytest.mark.gpu

OMEGAS = numpy.array([0.0, 0.3])
NLC = (6.0, 0.01)


def _system(nwater=3, npts=6000, level=0):
    from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole
    mol = mole.Mole(geometry.water_cluster(nwater, seed=2), 'def2-svp')
    g = grids.Grids(mol, level=level).build()
    rng = numpy.random.default_rng(5)
    sel = numpy.sort(rng.choice(g.size, min(g.size, npts), replace=False))
    dm = guess.huckel_density(mol)
    return mol, g.coords[sel], g.weights[sel], dm


def _densities(dm, n=3):
    rng = numpy.random.default_rng(11)
    out = [dm]
    for _ in range(n - 1):
        p = rng.normal(scale=0.02, size=dm.shape)
        out.append(dm + p + p.T)
    return out


@pytest.mark.parametrize('precision', ['mixed', 'fp64'])
def test_cycle_graph_replays_equal_eager(precision):
    import torch
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.basis import pack_mol
    mol, coords, weights, dm0 = _system()
    nao = mol.nao
    ctx = lib.Context(0)
    L = lib.load()
    basis = lib.DeviceBasis(ctx, pack_mol(mol))
    grid = lib.DeviceGrid(ctx, coords, weights)
    dev = torch.device('cuda:0')
    f64 = dict(dtype=torch.float64, device=dev)
    g_qd, g_dm = torch.zeros((nao, nao), **f64), torch.zeros((nao, nao), **f64)
    g_k, g_e, g_v = torch.zeros((2, nao, nao), **f64), torch.zeros(2, **f64), torch.zeros((nao, nao), **f64)
    e_k, e_e, e_v = torch.zeros_like(g_k), torch.zeros_like(g_e), torch.zeros_like(g_v)
    prec = lib.PREC[precision]
    dms = _densities(dm0)
    g_qd.copy_(torch.from_numpy(dms[0])); g_dm.copy_(torch.from_numpy(dms[0]))
    torch.cuda.synchronize()
    graph = lib.CycleGraph(ctx, basis, grid, g_qd.data_ptr(), OMEGAS, 1e-11, grid, g_dm.data_ptr(),
                           NLC, 0.0, precision, g_k.data_ptr(), g_e.data_ptr(), g_v.data_ptr())
    try:
        for rep, dm in enumerate(dms + dms[::-1]):
            d = torch.from_numpy(numpy.ascontiguousarray(dm)).to(dev)
            torch.cuda.synchronize()
            # eager reference (the owner context's own workspaces)
            lib.check(L.wb_k_build_dev(ctx.h, basis.h, grid.h, d.data_ptr(), 2, lib.ptr(OMEGAS),
                                       1e-11, prec, e_k.data_ptr()))
            lib.check(L.wb_nlc_vxc_dev(ctx.h, basis.h, grid.h, d.data_ptr(), NLC[0], NLC[1], 0.0,
                                       prec, e_e.data_ptr(), e_v.data_ptr()))
            ctx.sync()
            g_qd.copy_(d); g_dm.copy_(d)
            torch.cuda.synchronize()
            graph.launch()
            ctx.sync()
            assert torch.equal(g_k, e_k), (precision, rep, (g_k - e_k).abs().max().item())
            assert torch.equal(g_e, e_e), (precision, rep)
            assert torch.equal(g_v, e_v), (precision, rep, (g_v - e_v).abs().max().item())
            assert float(g_k.abs().max()) > 0.0 and float(g_e[0]) > 0.0
            if rep == 1:
                # regrow the owner context's buffers: the graph keeps its own
                lib.check(L.wb_ctx_set_copies(ctx.h, 12))
                lib.check(L.wb_k_build_dev(ctx.h, basis.h, grid.h, d.data_ptr(), 2,
                                           lib.ptr(OMEGAS), 1e-11, prec, e_k.data_ptr()))
                ctx.sync()
                lib.check(L.wb_ctx_set_copies(ctx.h, 8))
    finally:
        graph.close()


def test_cycle_graph_k_only_and_argument_checks():
    import ctypes
    import torch
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.basis import pack_mol
    mol, coords, weights, dm0 = _system()
    nao = mol.nao
    ctx = lib.Context(0)
    L = lib.load()
    basis = lib.DeviceBasis(ctx, pack_mol(mol))
    grid = lib.DeviceGrid(ctx, coords, weights)
    dev = torch.device('cuda:0')
    d = torch.from_numpy(dm0).to(dev)
    g_k = torch.zeros((1, nao, nao), dtype=torch.float64, device=dev)
    e_k = torch.zeros_like(g_k)
    torch.cuda.synchronize()
    graph = lib.CycleGraph(ctx, basis, grid, d.data_ptr(), [0.3], 1e-11, d_k=g_k.data_ptr())
    try:
        for _ in range(3):
            lib.check(L.wb_k_build_dev(ctx.h, basis.h, grid.h, d.data_ptr(), 1,
                                       lib.ptr(numpy.array([0.3])), 1e-11, 0, e_k.data_ptr()))
            ctx.sync()
            graph.launch()
            ctx.sync()
            assert torch.equal(g_k, e_k)
    finally:
        graph.close()
    h = ctypes.c_void_p()
    assert L.wb_cycle_graph_create(ctx.h, basis.h, None, None, 0, None, 0.0, None, None, 6.0,
                                   0.01, 0.0, 0, None, None, None, ctypes.byref(h)) == lib.WB_INVALID


@pytest.mark.parametrize('precision', ['mixed', 'fp64'])
@pytest.mark.parametrize('size', ['small', 'large'])
def test_concurrent_cycle_equals_sequential(precision, size):
    """large: 60 000 points, 118 VV10 point groups x 16 parts = 1888 work
    units, more than the concurrent cycle's capped VV10 grid holds blocks
    (4 per CU mixed, 3 per 4 CUs FP64), so its blocks take several units."""
    import torch
    from pyscf_wb97mv_fast_amd import lib
    from pyscf_wb97mv_fast_amd.basis import pack_mol
    mol, coords, weights, dm0 = _system(*((3, 6000, 0) if size == 'small' else (8, 60000, 1)))
    assert size == 'small' or len(coords) == 60000
    nao = mol.nao
    ctx = lib.Context(0)
    L = lib.load()
    basis = lib.DeviceBasis(ctx, pack_mol(mol))
    grid = lib.DeviceGrid(ctx, coords, weights)
    dev = torch.device('cuda:0')
    f64 = dict(dtype=torch.float64, device=dev)
    c_k, c_e, c_v = torch.zeros((2, nao, nao), **f64), torch.zeros(2, **f64), torch.zeros((nao, nao), **f64)
    s_k, s_e, s_v = torch.zeros_like(c_k), torch.zeros_like(c_e), torch.zeros_like(c_v)
    prec = lib.PREC[precision]
    dens = _densities(dm0) * 2 if size == 'small' else [dm0, dm0]
    half = len(dens) // 2
    for rep, dm in enumerate(dens):
        d = torch.from_numpy(numpy.ascontiguousarray(dm)).to(dev)
        torch.cuda.synchronize()
        lib.check(L.wb_k_build_dev(ctx.h, basis.h, grid.h, d.data_ptr(), 2, lib.ptr(OMEGAS),
                                   1e-11, prec, s_k.data_ptr()))
        lib.check(L.wb_nlc_vxc_dev(ctx.h, basis.h, grid.h, d.data_ptr(), NLC[0], NLC[1], 0.0,
                                   prec, s_e.data_ptr(), s_v.data_ptr()))
        # phase timing on for the second half: the flows then run in sequence
        lib.check(L.wb_ctx_set_profile(ctx.h, 1 if rep >= half else 0))
        lib.check(L.wb_cycle_dev(ctx.h, basis.h, grid.h, d.data_ptr(), 2, lib.ptr(OMEGAS), 1e-11,
                                 grid.h, d.data_ptr(), NLC[0], NLC[1], 0.0, prec,
                                 c_k.data_ptr(), c_e.data_ptr(), c_v.data_ptr()))
        ctx.sync()
        assert torch.equal(c_k, s_k), (precision, rep)
        assert torch.equal(c_e, s_e) and torch.equal(c_v, s_v), (precision, rep)
    lib.check(L.wb_ctx_set_profile(ctx.h, 0))
    # K only / VV10 only through the same entry
    lib.check(L.wb_cycle_dev(ctx.h, basis.h, None, None, 0, None, 1e-11, grid.h, d.data_ptr(),
                             NLC[0], NLC[1], 0.0, prec, None, c_e.data_ptr(), c_v.data_ptr()))
    ctx.sync()
    assert torch.equal(c_e, s_e)
