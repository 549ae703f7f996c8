"""Device grid construction: the Stratmann-Scuseria-Frisch partition
(wb_grid_partition, gpu/grids.py) against a brute-force cell weight over all
atoms (tests/independent.py) and against the host partition of
systems/grids.py (pinned in test_grids.py); and the host-threaded shell-pair
build, whose result must not depend on the thread count."""
import os

import numpy
import pytest

from independent import ssf_brute
from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.errors import GpuError
from pyscf_wb97mv_fast_amd.gpu import grids as dgrids
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    return lib.Context(0)


def test_partition_vs_brute_force(ctx):
    for seed in range(3):
        rng = numpy.random.default_rng(seed)
        natm = 7
        atoms = rng.uniform(-3, 3, (natm, 3))
        owner = rng.integers(0, natm, 500)
        coords = atoms[owner] + rng.normal(size=(500, 3)) * rng.uniform(0.1, 3, (500, 1))
        got = dgrids.ssf_partition(ctx, coords, atoms, owner)
        ref = ssf_brute(coords, atoms, owner)
        assert numpy.allclose(got, ref, rtol=0, atol=1e-13), seed


def test_partition_vs_host_water8(ctx):
    mol = mole.Mole(geometry.water_cluster(8, seed=0), 'def2-svp')
    coords, w, owner = grids.Grids(mol, level=1).raw_points()
    atoms = mol.atom_coords()
    print('points', len(coords), 'atoms', len(atoms), flush=True)
    got = dgrids.ssf_partition(ctx, coords, atoms, owner)
    ref = grids._ssf_partition(coords, atoms, owner)
    assert numpy.abs(got - ref).max() < 1e-13


def test_device_grid_matches_host_grid(ctx):
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    gh = grids.Grids(mol, level=3).build()
    gd = grids.Grids(mol, level=3, device=ctx).build()
    assert gd.size == gh.size
    assert numpy.array_equal(gd.coords, gh.coords)
    # weight = quadrature weight x partition; the partitions agree to 1e-13
    coords, wq, _ = grids.Grids(mol, level=3).raw_points()
    wq = wq[numpy.isin(coords.view('V24').ravel(), gh.coords.view('V24').ravel())]
    assert len(wq) == gh.size
    assert (numpy.abs(gd.weights - gh.weights) <= 1e-13 * wq).all()


@pytest.mark.parametrize('case', ['water8_l1', 'drug10_l3'])
def test_device_built_grid_equals_host_points_with_device_partition(ctx, case):
    """wb_grid_build: points, partition and cut on the device equal the host's
    points and cut with the device partition (wb_grid_partition), bit for
    bit, in the host's order."""
    if case == 'water8_l1':
        mol, level = mole.Mole(geometry.water_cluster(8, seed=0), 'def2-svp'), 1
    else:
        mol, level = mole.Mole(geometry.drug_like(10, seed=2), 'def2-tzvp'), 3
    g = grids.Grids(mol, level=level)
    coords, wq, owner = g.raw_points()
    w = wq * dgrids.ssf_partition(ctx, coords, mol.atom_coords(), owner)
    keep = w > g.w_cut
    gd = grids.Grids(mol, level=level, device=ctx).build()
    print(case, 'raw', len(coords), 'kept', gd.size, flush=True)
    assert gd.size == int(keep.sum()) and gd.size < len(coords)
    assert numpy.array_equal(gd.coords, coords[keep])
    assert numpy.array_equal(gd.weights, w[keep])


def test_device_grid_build_edges(ctx):
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    t = grids.Grids(mol, level=0).shell_tables()
    atoms = mol.atom_coords()
    c, w = dgrids.build_grid(ctx, atoms, *t, w_cut=1e300)          # everything cut
    assert c.shape == (0, 3) and w.shape == (0,)
    c, w = dgrids.build_grid(ctx, atoms, t[0][:0], t[1][:0], t[2][:0], t[3][:0], t[4])
    assert c.shape == (0, 3)
    bad = (t[0].copy(),) + t[1:]
    bad[0][0] = len(atoms)
    with pytest.raises(GpuError):
        dgrids.build_grid(ctx, atoms, *bad)


def test_partition_edge_cases(ctx):
    atoms = numpy.array([[0.0, 0, 0], [1.5, 0, 0]])
    assert dgrids.ssf_partition(ctx, numpy.zeros((0, 3)), atoms, []).shape == (0,)
    # one atom: every point belongs to it entirely
    one = dgrids.ssf_partition(ctx, numpy.random.default_rng(0).normal(size=(50, 3)),
                               atoms[:1], numpy.zeros(50, dtype=int))
    assert (one == 1.0).all()
    with pytest.raises(GpuError):
        dgrids.ssf_partition(ctx, numpy.zeros((2, 3)), atoms, [0, 2])
    with pytest.raises(GpuError):
        dgrids.ssf_partition(ctx, numpy.ones((1, 3)), numpy.zeros((2, 3)), [0])


def _k(ctx, basis, grid, dm):
    om = numpy.array([0.0, 0.3])
    out = numpy.empty((2, basis.nao, basis.nao))
    lib.check(lib.load().wb_k_build(ctx.h, basis.h, grid.h, lib.ptr(lib.f64(dm)), 2, lib.ptr(om),
                                    1e-11, lib.PREC['mixed'], lib.ptr(out)))
    return out


def test_threaded_shell_pairs_are_thread_count_independent(ctx):
    mol = mole.Mole(geometry.water_cluster(8, seed=1), 'def2-svp')
    g = grids.Grids(mol, level=0).build()
    pk = pack_mol(mol)
    dm = guess.huckel_density(mol)
    grid = lib.DeviceGrid(ctx, g.coords, g.weights)
    res = []
    old = os.environ.get('WB_HOST_THREADS')
    try:
        for nt in ('1', '3', '16'):
            os.environ['WB_HOST_THREADS'] = nt
            b = lib.DeviceBasis(ctx, pk)
            res.append((b.npairs, b.nprims, _k(ctx, b, grid, dm)))
    finally:
        if old is None:
            os.environ.pop('WB_HOST_THREADS', None)
        else:
            os.environ['WB_HOST_THREADS'] = old
    for r in res[1:]:
        assert r[:2] == res[0][:2]
        assert numpy.array_equal(r[2], res[0][2])


def test_grid_overlap_vs_quadrature_and_analytic(ctx):
    import oracle
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    b = lib.DeviceBasis(ctx, pack_mol(mol))
    # against the oracle's AO values on the same points (dense quadrature)
    g = grids.Grids(mol, level=1).build()
    gr = lib.DeviceGrid(ctx, g.coords, g.weights)
    S = dgrids.grid_overlap(ctx, b, gr)
    X = oracle.ao_eval(oracle.OracleBasis.from_mol(mol), g.coords)[0]
    ref = (X * g.weights[:, None]).T @ X
    assert numpy.abs(S - ref).max() <= 1e-12 * numpy.abs(ref).max()
    assert numpy.array_equal(S, S.T) or numpy.abs(S - S.T).max() <= 1e-14
    # a fine grid reproduces the analytic overlap
    gf = grids.Grids(mol, level=5, device=ctx).build()
    Sf = dgrids.grid_overlap(ctx, b, lib.DeviceGrid(ctx, gf.coords, gf.weights))
    assert numpy.abs(Sf - guess.overlap_matrix(mol)).max() < 1e-6


def test_pjs_projector_on_the_device_matches_the_host_solve():
    """Q = S_g^-1 S (wb_pjs_projector, rocSOLVER LU) against numpy's solve on
    the oracle's quadrature overlap, and against the identity in the limit it
    has: S_g Q = S."""
    import oracle
    from pyscf_wb97mv_fast_amd.gpu import grids as dgrids
    ctx = lib.Context(0)
    mol = mole.Mole(geometry.water_cluster(3, seed=2), 'def2-svp')
    g = grids.Grids(mol, level=1, device=ctx).build()
    b = lib.DeviceBasis(ctx, pack_mol(mol))
    gr = lib.DeviceGrid(ctx, g.coords, g.weights)
    S = guess.overlap_matrix(mol)
    Q = dgrids.pjs_projector(ctx, b, gr, S)
    X = oracle.ao_eval(oracle.OracleBasis.from_mol(mol), g.coords)[0]
    Sg = (X * g.weights[:, None]).T @ X
    ref = numpy.linalg.solve(Sg, S)
    print('max|dQ| %.2e, max|Q - 1| %.2e' % (numpy.abs(Q - ref).max(), numpy.abs(Q - numpy.eye(len(Q))).max()))
    assert numpy.allclose(Q, ref, rtol=0, atol=1e-9 * numpy.abs(ref).max())
    assert numpy.allclose(Sg @ Q, S, atol=1e-9)
    ctx.close()


def test_pjs_projector_brings_the_k_hook_closer_to_the_analytic_exchange():
    """The overlap-fitted projector built on the device, used through the K
    hook's PJS contract on the s-only molecule of the closed-form pin: the
    K with Q is closer to the exact exchange than the K with Q = I."""
    from pyscf_wb97mv_fast_amd.gpu import sgx_k
    from pyscf_wb97mv_fast_amd.systems import sgx
    from test_oracle_pins import s_only_system, ss_eri
    ctx = lib.Context(0)
    mol = s_only_system()
    rng = numpy.random.default_rng(1)
    D = rng.normal(size=(mol.nao, mol.nao))
    D = 0.5 * (D + D.T)
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='fp64')
    g = grids.Grids(mol, level=2).build()
    for omega in (0.0, 0.3):
        exact = numpy.einsum('mlns,ls->mn', ss_eri(mol, omega), D)
        with mol.with_range_coulomb(omega):
            k_i = sgx_k.get_k_only_gpu(kb, sgx.SyntheticSGX(mol, g), D)
            k_q = sgx_k.get_k_only_gpu(kb, sgx.SyntheticSGX(mol, g, fit_ovlp=ctx), D)
        e_i, e_q = numpy.abs(k_i - exact).max(), numpy.abs(k_q - exact).max()
        print('omega %.1f: max|dK| Q=I %.2e, fitted Q %.2e' % (omega, e_i, e_q))
        assert e_q < e_i
    ctx.close()
