"""GPU parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerances: the FP64 form is checked at 1e-10 relative; the mixed form (FP32
integrals / pair sums) follows the reference's gates (tests/gpu/
test_sgx_k.py:22-44: max|dK| < 1e-6 max|K|; tests/gpu/test_xc_vv10.py:
274-292: F/U/W relative < 1e-6)."""
import numpy
import pytest

import oracle
from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.gpu import sgx_k, vv10
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole, sgx

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


def _system(basis='def2-svp', level=0, npts=2500, seed=0, atoms=None):
    mol = mole.Mole(atoms or geometry.WATER_DIMER, basis)
    g = grids.Grids(mol, level=level).build()
    if npts and g.size > npts:
        rng = numpy.random.default_rng(seed)
        sel = numpy.sort(rng.choice(g.size, npts, replace=False))
        g.coords, g.weights = g.coords[sel], g.weights[sel]
    return mol, g, guess.huckel_density(mol)


@pytest.mark.parametrize('kind', [0, 1, 2, 3])
@pytest.mark.parametrize('shape', [(16, 64, 4), (48, 1536, 48), (100, 70, 333), (648, 640, 648),
                                   (130, 2002, 46), (97, 129, 4098), (200, 648, 20480)])
def test_fp64_mfma_gemm(ctx, kind, shape):
    M, N, K = shape
    rng = numpy.random.default_rng(M + N + K)
    A = rng.normal(size=(M, K))
    B = rng.normal(size=(K, N)) if kind % 2 == 0 else rng.normal(size=(N, K))
    w = None if kind % 2 == 0 else rng.uniform(0.5, 1.5, K)
    C = numpy.empty((M, N))
    wp = lib.ptr(w) if w is not None else None
    lib.check(lib.load().wb_test_gemm(ctx.h, kind, M, N, K, lib.ptr(A), lib.ptr(B), wp, lib.ptr(C)))
    ref = A @ B if kind % 2 == 0 else (A * w) @ B.T
    assert numpy.allclose(C, ref, rtol=1e-12, atol=1e-12 * numpy.abs(ref).max())


def test_ao_eval_matches_oracle(ctx):
    mol, g, _ = _system('def2-tzvp', npts=700)
    b = lib.DeviceBasis(ctx, __import__('pyscf_wb97mv_fast_amd.basis', fromlist=['x']).pack_mol(mol))
    out = numpy.empty((4, g.size, b.nao))
    lib.check(lib.load().wb_ao_eval(ctx.h, b.h, g.size, lib.ptr(lib.f64(g.coords)), 1,
                                    lib.ptr(out)))
    ref = oracle.ao_eval(oracle.OracleBasis.from_mol(mol), g.coords, deriv=1)
    assert numpy.allclose(out, ref, rtol=0, atol=1e-12 * numpy.abs(ref).max())


@pytest.mark.parametrize('basis', ['def2-svp', 'def2-tzvp'])
@pytest.mark.parametrize('omega', [0.0, 0.3])
def test_k_fp64_matches_oracle(ctx, basis, omega):
    mol, g, dm = _system(basis, npts=1500)
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='fp64')
    K = kb.k_raw(g, dm, [omega])[0]
    ref = oracle.k_build(oracle.OracleBasis.from_mol(mol), g.coords, g.weights, dm, omega,
                         pair_cut=1e-14)
    print('%s omega=%.1f fp64 max|dK| %.3e of %.3e' % (basis, omega, numpy.abs(K - ref).max(),
                                                      numpy.abs(ref).max()))
    assert numpy.allclose(K, ref, rtol=0, atol=1e-10 * numpy.abs(ref).max())


@pytest.mark.parametrize('basis', ['def2-svp', 'def2-tzvp'])
@pytest.mark.parametrize('omega', [0.0, 0.3])
def test_k_mixed_matches_oracle(ctx, basis, omega):
    mol, g, dm = _system(basis, npts=1500)
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='mixed')
    K = kb.k_raw(g, dm, [omega])[0]
    ref = oracle.k_build(oracle.OracleBasis.from_mol(mol), g.coords, g.weights, dm, omega,
                         pair_cut=1e-14)
    dE = 0.25 * numpy.einsum('ij,ji->', dm, K - ref)
    print('%s omega=%.1f mixed max|dK| %.3e of %.3e, dE_K %.3e' % (
        basis, omega, numpy.abs(K - ref).max(), numpy.abs(ref).max(), dE))
    assert numpy.abs(K - ref).max() < 1e-6 * numpy.abs(ref).max()
    # energy gate on def2-SVP as the reference's (3e-7, tests/gpu/test_sgx_k.py:
    # 38-43); for def2-TZVP the reference only reports it (:47-66)
    if basis == 'def2-svp':
        assert abs(dE) < 3e-7


@pytest.mark.parametrize('atoms,odd_nao', [(None, False),
                                           ([('F', (0.0, 0.0, 0.0)), ('H', (0.0, 0.0, 0.92)),
                                             ('O', (2.8, 0.0, 0.0)), ('H', (3.3, 0.8, 0.0)),
                                             ('H', (3.3, -0.8, 0.0))], True)])
def test_k_and_vv10_on_ragged_shapes(ctx, atoms, odd_nao):
    """The FP64 GEMMs run over per-tile AO lists (128-point tiles) and per
    point-split lists; an odd number of 64-point sub-blocks leaves a last
    tile of 64 points, and an odd AO count (HF + H2O, 43 AOs) takes the dense
    GEMMs instead.  K (both operators) and the VV10 flow in FP64 against the
    oracle either way."""
    mol, g, dm = _system(npts=1600, atoms=atoms)          # 25 sub-blocks
    assert (g.size + 63) // 64 % 2 == 1
    assert (mol.ao_loc_nr()[-1] % 2 == 1) == odd_nao
    ob = oracle.OracleBasis.from_mol(mol)
    K = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='fp64').k_raw(g, dm, [0.0, 0.3])
    for op, omega in enumerate((0.0, 0.3)):
        ref = oracle.k_build(ob, g.coords, g.weights, dm, omega, pair_cut=1e-14)
        assert numpy.allclose(K[op], ref, rtol=0, atol=1e-10 * numpy.abs(ref).max()), omega
    n_g, e_g, v_g = vv10.gpu_nr_nlc_vxc(vv10.NlcState(ctx, mol), sgx.SyntheticNumInt(), mol, g,
                                        'wb97m-v', dm, precision='fp64')
    n_o, e_o, v_o = oracle.nr_nlc_vxc(ob, g.coords, g.weights, dm)
    assert abs(n_g - n_o) < 1e-10 and abs(e_g - e_o) < 1e-10
    assert numpy.abs(v_g - v_o).max() < 1e-9


def test_k_two_operators_one_call_and_repeatable(ctx):
    mol, g, dm = _system(npts=1200)
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='mixed')
    both = kb.k_raw(g, dm, [0.0, 0.3])
    again = kb.k_raw(g, dm, [0.0, 0.3])
    assert numpy.array_equal(both, again)                     # bit-reproducible
    assert numpy.array_equal(both[1], kb.k_raw(g, dm, [0.3])[0])


def test_k_block_padding_and_copies(ctx):
    """(H2O)5 in def2-SVP has 35 s, 20 p and 5 d shells: the last B block of
    the ss/sp/ds (8), pp (8) and dd (2) classes is partly empty.  K in FP64
    matches the oracle for 1, 3 and 8 G copies, and the copy count changes
    only the summation order."""
    mol, g, dm = _system(atoms=geometry.water_cluster(5, seed=3), level=0, npts=1800)
    ref = oracle.k_build(oracle.OracleBasis.from_mol(mol), g.coords, g.weights, dm, 0.0,
                         pair_cut=1e-14)
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='fp64')
    Ks = []
    try:
        for nc in (1, 3, 8):
            lib.check(lib.load().wb_ctx_set_copies(ctx.h, nc))
            Ks.append(kb.k_raw(g, dm, [0.0])[0])
    finally:
        lib.check(lib.load().wb_ctx_set_copies(ctx.h, 8))
    for K in Ks:
        assert numpy.allclose(K, ref, rtol=0, atol=1e-10 * numpy.abs(ref).max())
        assert numpy.allclose(K, Ks[0], rtol=0, atol=1e-13 * numpy.abs(ref).max())


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_k_drug_like_tzvp(ctx, precision):
    """A small seeded drug-like chain (C/N/O backbone with F, Cl, S-CH3
    substituents) in def2-TZVP: every class up to (f, f), both operators."""
    mol, g, dm = _system('def2-tzvp', atoms=geometry.drug_like(10, seed=2), level=0, npts=600)
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision=precision)
    Kb = kb.k_raw(g, dm, [0.0, 0.3])
    ob = oracle.OracleBasis.from_mol(mol)
    for K, omega in zip(Kb, (0.0, 0.3)):
        ref = oracle.k_build(ob, g.coords, g.weights, dm, omega, pair_cut=1e-14)
        err = numpy.abs(K - ref).max() / numpy.abs(ref).max()
        print('drug-like tzvp %s omega=%.1f nao %d max|dK|/max|K| %.2e' % (precision, omega,
                                                                      mol.nao, err))
        assert err < (1e-10 if precision == 'fp64' else 1e-6)


def test_k_large_basis_clears_g(ctx):
    """(H2O)86 in def2-SVP has 2064 AOs: past the 2048 rows that the
    written-shell masks of the G copies cover, the first class launched
    clears G instead.  K for both operators on one 64-point sub-block
    (random symmetric density, no screening) matches the oracle."""
    import types
    mol = mole.Mole(geometry.water_cluster(86, seed=4), 'def2-svp')
    assert mol.nao + 7 > 2048
    rng = numpy.random.default_rng(0)
    xyz = mol.atom_coords()
    pts = xyz[rng.choice(len(xyz), 64)] + rng.normal(scale=0.8, size=(64, 3))
    g = types.SimpleNamespace(coords=pts, weights=rng.uniform(0.01, 0.1, 64))
    a = rng.normal(scale=0.01, size=(mol.nao, mol.nao))
    dm = a + a.T
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='fp64')
    Kb = kb.k_raw(g, dm, [0.0, 0.3])
    ob = oracle.OracleBasis.from_mol(mol)
    for K, omega in zip(Kb, (0.0, 0.3)):
        ref = oracle.k_build(ob, g.coords, g.weights, dm, omega, pair_cut=1e-14)
        assert numpy.allclose(K, ref, rtol=0, atol=1e-10 * numpy.abs(ref).max()), omega


def test_k_screening(ctx):
    mol, g, dm = _system(atoms=geometry.water_cluster(8, seed=1), level=0, npts=6000)
    kb0 = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='fp64')
    kb1 = sgx_k.GpuKBuilder(mol, ctx, tile_tol=1e-11, precision='fp64')
    K0 = kb0.k_raw(g, dm, [0.0])[0]
    K1 = kb1.k_raw(g, dm, [0.0])[0]
    st = kb1.last_stats
    print('screening kept %d of %d, max|dK| %.2e' % (st['pairs_kept'], st['pairs_total'],
                                                      numpy.abs(K1 - K0).max()))
    assert numpy.abs(K1 - K0).max() < 1e-8
    assert st['pairs_kept'] < st['pairs_total']


def test_get_k_only_contract(ctx):
    mol, g, dm = _system(npts=1000)
    rng = numpy.random.default_rng(3)
    Q = numpy.eye(mol.nao) + 0.01 * rng.normal(size=(mol.nao, mol.nao))
    s = sgx.SyntheticSGX(mol, g, Q=Q, sym_ovlp=True)
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision='fp64')
    with mol.with_range_coulomb(0.3):
        K = sgx_k.get_k_only_gpu(kb, s, dm)
    ref = oracle.get_k_only(oracle.OracleBasis.from_mol(mol), g.coords, g.weights, dm, Q=Q,
                            sym_ovlp=True, omega=0.3)
    assert numpy.allclose(K, ref, atol=1e-10 * numpy.abs(ref).max())
    assert numpy.array_equal(K, K.T)


def _real_vv10_points(n=1100):
    """VV10 inputs at real grid points of the water dimer (the reference's
    device test uses real points too, tests/gpu/test_xc_vv10.py:244-260)."""
    mol, g, dm = _system(npts=6000, seed=9)
    ob = oracle.OracleBasis.from_mol(mol)
    ao = oracle.ao_eval(ob, g.coords, deriv=1)
    c0 = ao[0] @ dm
    rho0 = numpy.einsum('pm,pm->p', c0, ao[0])
    grad = 2 * numpy.einsum('pm,gpm->gp', c0, ao[1:4])
    keep = numpy.flatnonzero(rho0 >= 1e-8)[:n]
    q, W0, kappa = oracle.vv10_pointwise(rho0[keep], grad[:, keep], g.weights[keep],
                                         (6.0, 0.01))[:3]
    return g.coords[keep], q, W0, kappa


@pytest.mark.parametrize('precision,rtol', [('fp64', 1e-11), ('mixed', 1e-6)])
def test_vv10_sum_matches_oracle(ctx, precision, rtol):
    """F, U, W relative error < 1e-6 in FP32 on 1100 real points: several
    256-point tiles plus a ragged tail (tests/gpu/test_xc_vv10.py:274-292)."""
    coords, q, W0, kappa = _real_vv10_points()
    assert len(coords) == 1100
    k = vv10.Vv10Kernel(ctx, precision=precision)
    got = k.kernel_sum(coords, q, W0, kappa)
    ref = oracle.vv10_sum(coords, q, W0, kappa)
    for a, b, name in zip(got, ref, 'FUW'):
        rel = numpy.abs(a - b) / numpy.abs(b)
        assert rel.max() < rtol, (name, rel.max())


def test_vv10_sum_fp64_random_points(ctx):
    rng = numpy.random.default_rng(4)
    n = 900
    coords = rng.normal(size=(n, 3)) * 4
    q, W0, kappa = rng.uniform(1e-4, 0.1, n), rng.uniform(0.05, 2.0, n), rng.uniform(0.05, 1, n)
    got = vv10.Vv10Kernel(ctx, precision='fp64').kernel_sum(coords, q, W0, kappa)
    for a, b in zip(got, oracle.vv10_sum(coords, q, W0, kappa)):
        assert numpy.allclose(a, b, rtol=1e-11, atol=0)


def test_vv10_pruning_two_far_clusters(ctx):
    """inner/outer tile indexing with two clusters 1000 bohr apart
    (tests/gpu/test_xc_vv10.py:320-343)"""
    rng = numpy.random.default_rng(5)
    n = 600
    coords = rng.normal(size=(n, 3))
    coords[n // 2:] += 1000.0
    q, W0, kappa = rng.uniform(0.01, 0.1, n), rng.uniform(0.5, 1.0, n), rng.uniform(0.3, 0.6, n)
    ref = oracle.vv10_sum(coords, q, W0, kappa)
    got = vv10.Vv10Kernel(ctx, tile_tol=1e-14, precision='fp64').kernel_sum(coords, q, W0, kappa)
    for a, b in zip(got, ref):
        assert numpy.allclose(a, b, rtol=1e-9)


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_vv10_pruning_counts_and_skips_far_tiles(ctx, precision):
    """Two clusters 1000 bohr apart: every (wave, tile) pair across the gap
    is skipped (the device's own count, wb_vv10_stats), none within a
    cluster, and F, U, W stay those of the exact sum (FP64 to 1e-9; the
    mixed form to 1e-4 here, since its FP32 coordinates are relative to the
    centroid of both clusters, 500 bohr from every point: an ulp of 3e-5
    bohr on distances of about 1 bohr); without pruning nothing is tested."""
    rng = numpy.random.default_rng(6)
    n = 1024                                  # two 512-point groups, 8 tiles each
    coords = rng.normal(size=(n, 3))
    coords[n // 2:] += 1000.0
    q, W0, kappa = rng.uniform(0.01, 0.1, n), rng.uniform(0.5, 1.0, n), rng.uniform(0.3, 0.6, n)
    ref = oracle.vv10_sum(coords, q, W0, kappa)
    k = vv10.Vv10Kernel(ctx, tile_tol=1e-12, precision=precision)
    got = k.kernel_sum(coords, q, W0, kappa)
    tested, skipped = ctx.vv10_stats()
    # points are kept in order, so each 512-point group (4 waves, 8 tiles) is
    # one cluster; every unordered pair of groups is visited once, so group
    # 0's waves test its own 8 tiles and group 1's 8, group 1's waves its own
    assert tested == 4 * (8 + 8) + 4 * 8, tested
    assert skipped == 4 * 8, skipped
    for a, b in zip(got, ref):
        assert numpy.allclose(a, b, rtol=1e-9 if precision == 'fp64' else 1e-4)
    vv10.Vv10Kernel(ctx, tile_tol=0.0, precision=precision).kernel_sum(coords, q, W0, kappa)
    assert ctx.vv10_stats() == (0, 0)


def test_vv10_pruning_on_a_real_density(ctx):
    """Mixed form, two (H2O)8 clusters 60 bohr apart, level-0 grid: with
    tile_tol 1e-10 and 1e-8 the device skips tile pairs (more at the larger
    tolerance), and E_nl moves by less than the bound the tolerance implies
    (each skipped pair's contribution to any F_i, U_i or W_i is below tol,
    so |dE_nl| <= 0.75 sum_i q_i * (tiles skipped per point) * tol).  The
    bound takes each tile's smallest W0 and kappa and its whole charge, so
    on real densities it prunes only pairs far from both tiles' dense
    points."""
    base = geometry.water_cluster(8, seed=2)
    shift = 60.0 * mole.BOHR
    atoms = base + [(s_, (x + shift, y, z)) for s_, (x, y, z) in base]
    mol = mole.Mole(atoms, 'def2-svp')
    g = grids.Grids(mol, level=0).build()
    dm = guess.huckel_density(mol)
    st = vv10.NlcState(ctx, mol)
    ni = sgx.SyntheticNumInt()
    n0, e0, v0 = vv10.gpu_nr_nlc_vxc(st, ni, mol, g, 'wb97m-v', dm, precision='mixed')
    ntiles = (g.size + 63) // 64
    last = 0
    for tol in (1e-10, 1e-8):
        n1, e1, v1 = vv10.gpu_nr_nlc_vxc(st, ni, mol, g, 'wb97m-v', dm, precision='mixed',
                                         tile_tol=tol)
        tested, skipped = ctx.vv10_stats()
        print('tol %.0e: tile pairs tested %d, skipped %d (%.1f %%); dE_nl %.2e, max|dV| %.2e'
              % (tol, tested, skipped, 100.0 * skipped / max(tested, 1), e1 - e0,
                 numpy.abs(v1 - v0).max()))
        assert skipped > last
        last = skipped
        assert abs(e1 - e0) <= 0.75 * n0 * ntiles * tol
        assert abs(n1 - n0) == 0.0


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_nlc_vxc_matches_oracle(ctx, precision):
    mol, g, dm = _system(npts=3000)
    st = vv10.NlcState(ctx, mol)
    n_g, e_g, v_g = vv10.gpu_nr_nlc_vxc(st, sgx.SyntheticNumInt(), mol, g, 'wb97m-v', dm,
                                        precision=precision)
    n_o, e_o, v_o = oracle.nr_nlc_vxc(oracle.OracleBasis.from_mol(mol), g.coords, g.weights, dm)
    print('%s nelec %.12f vs %.12f, exc %.3e, max|dV| %.2e' % (precision, n_g, n_o, e_g - e_o,
                                                              numpy.abs(v_g - v_o).max()))
    assert abs(n_g - n_o) < 1e-10
    tol = 1e-10 if precision == 'fp64' else 1e-7
    assert abs(e_g - e_o) < tol
    assert numpy.abs(v_g - v_o).max() < (1e-9 if precision == 'fp64' else 1e-6)
