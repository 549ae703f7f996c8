"""Oracle parity on the benchmark workloads themselves, with screening on.

For (H2O)8 and (H2O)27 in def2-SVP and the 40-heavy-atom drug-like chain in
def2-TZVP (f shells), each on its level-1 grid with the cell partition on
the device and an extended-Hueckel density (the molecules bench.py times):

* the production K build (tile_tol 1e-11, both operators in one call, the
  fused classes, primitive tails skipped) in FP64 and in the mixed form
  against the unscreened FP64 oracle on a subsample of the real grid.  The
  device grid keeps every point of the real grid, so the sub-blocks and
  their screening are the production ones, and zero weight on the points
  outside the sample (K is linear in the weights; a zero-weight point adds
  nothing and its sub-block's bound shrinks to the sampled points);
* the whole-grid cycle in both forms: the per-cycle energy errors of the
  mixed form against the FP64 form (dE_x = -1/4 tr(D (0.15 dK_1/r + 0.85
  dK_erf)), dE_K per operator, dE_nl) within the 1e-6 Ha budget and the
  reference's K gates (tests/gpu/test_sgx_k.py:69-89: |dE_K| < 3e-6 for
  1/r, < 2e-6 for erf on water27), and on (H2O)27 at least half of the
  (sub-block, pair) tasks skipped (tests/gpu/test_sgx_k.py:226-243);
* on (H2O)27, the VV10 kernel sums of the real density (device rho,
  pointwise terms by the oracle) for all points against the oracle's
  literal sums for 1024 outer points over the full inner set
  (tests/gpu/test_xc_vv10.py:274-292: relative < 1e-6 in FP32).
"""
import numpy
import pytest

import oracle
from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.gpu import vv10, xc
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu
OM = numpy.array([0.0, 0.3])
TOL = 1e-11
NLC = (6.0, 0.01)

WORKLOADS = {
    'water8': (lambda: mole.Mole(geometry.water_cluster(8, seed=0), 'def2-svp'), 512),
    'water27': (lambda: mole.Mole(geometry.water_cluster(27, seed=0), 'def2-svp'), 512),
    'drug40_tzvp': (lambda: mole.Mole(geometry.drug_like(40, seed=0), 'def2-tzvp'), 128),
}


@pytest.fixture(scope='module')
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


_CACHE = {}


def workload(ctx, name):
    if name not in _CACHE:
        mk, nsample = WORKLOADS[name]
        mol = mk()
        g = grids.Grids(mol, level=1, device=ctx).build()
        dm = guess.huckel_density(mol)
        b = lib.DeviceBasis(ctx, pack_mol(mol))
        full = lib.DeviceGrid(ctx, g.coords, g.weights)
        rng = numpy.random.default_rng(7)
        sel = numpy.sort(rng.choice(g.size, nsample, replace=False))
        wmask = numpy.zeros_like(g.weights)
        wmask[sel] = g.weights[sel]
        masked = lib.DeviceGrid(ctx, g.coords, wmask)
        print('%s: nao %d, %d points, %d sampled' % (name, b.nao, g.size, nsample), flush=True)
        _CACHE[name] = dict(mol=mol, g=g, dm=dm, b=b, full=full, masked=masked, sel=sel)
    return _CACHE[name]


def k_build(ctx, b, grid, dm, prec, tol=TOL):
    k = numpy.empty((2, b.nao, b.nao))
    lib.check(lib.load().wb_k_build(ctx.h, b.h, grid.h, lib.ptr(dm), 2, lib.ptr(OM), tol,
                                    lib.PREC[prec], lib.ptr(k)))
    st = numpy.zeros(2, dtype=numpy.int64)
    lib.check(lib.load().wb_k_stats(ctx.h, grid.h, lib.ptr(st)))
    return k, st


def nlc(ctx, b, grid, dm, prec):
    o2, v = numpy.empty(2), numpy.empty((b.nao, b.nao))
    lib.check(lib.load().wb_nlc_vxc(ctx.h, b.h, grid.h, lib.ptr(dm), NLC[0], NLC[1], 0.0,
                                    lib.PREC[prec], lib.ptr(o2), lib.ptr(v)))
    return o2, v


@pytest.mark.parametrize('name', sorted(WORKLOADS))
def test_screened_k_on_the_workload_grid_matches_the_oracle(ctx, name):
    w = workload(ctx, name)
    mol, g, dm, sel = w['mol'], w['g'], w['dm'], w['sel']
    ob = oracle.OracleBasis.from_mol(mol)
    refs = [oracle.k_build(ob, g.coords[sel], g.weights[sel], dm, om, pair_cut=1e-14)
            for om in OM]
    for prec in ('fp64', 'mixed'):
        k, st = k_build(ctx, w['b'], w['masked'], dm, prec)
        for op, ref in enumerate(refs):
            dk = numpy.abs(k[op] - ref).max()
            de = 0.25 * numpy.einsum('ij,ji->', dm, k[op] - ref)
            print('%s %s omega=%.1f: max|dK| %.2e of max|K| %.2e, dE_K %.2e (kept %d of %d)'
                  % (name, prec, OM[op], dk, numpy.abs(ref).max(), de, st[1], st[0]), flush=True)
            if prec == 'fp64':
                assert dk < 1e-8 and abs(de) < 1e-8           # screening + tails only
            else:
                assert dk < 1e-6 * numpy.abs(ref).max()
                assert abs(de) < 3e-7


def test_unscreened_k_on_water27_matches_the_oracle_to_fp64(ctx):
    w = workload(ctx, 'water27')
    ob = oracle.OracleBasis.from_mol(w['mol'])
    sel = w['sel']
    k, st = k_build(ctx, w['b'], w['masked'], w['dm'], 'fp64', tol=0.0)
    for op, om in enumerate(OM):
        ref = oracle.k_build(ob, w['g'].coords[sel], w['g'].weights[sel], w['dm'], om,
                             pair_cut=1e-14)
        assert numpy.allclose(k[op], ref, rtol=0, atol=1e-10 * numpy.abs(ref).max())


@pytest.mark.parametrize('name', sorted(WORKLOADS))
def test_full_cycle_mixed_within_budget_of_fp64(ctx, name):
    w = workload(ctx, name)
    b, full, dm = w['b'], w['full'], w['dm']
    k32, st = k_build(ctx, b, full, dm, 'mixed')
    k64, _ = k_build(ctx, b, full, dm, 'fp64')
    n32, v32 = nlc(ctx, b, full, dm, 'mixed')
    n64, v64 = nlc(ctx, b, full, dm, 'fp64')
    dK = k32 - k64
    deK = [0.25 * float(numpy.einsum('ij,ji->', dm, dK[op])) for op in range(2)]
    dEx = -0.25 * float(numpy.einsum('ij,ji->', dm, 0.15 * dK[0] + 0.85 * dK[1]))
    dEnl = n32[1] - n64[1]
    kept = st[1] / st[0]
    print('%s: dE_x %.2e, dE_K(1/r) %.2e, dE_K(erf) %.2e, dE_nl %.2e, max|dK| %.2e, '
          'kept %.3f of tasks, nelec %.8f' % (name, dEx, deK[0], deK[1], dEnl, abs(dK).max(),
                                               kept, n64[0]), flush=True)
    assert abs(dEx) <= 1e-6 and abs(dEnl) <= 1e-7
    assert abs(deK[0]) < 3e-6 and abs(deK[1]) < 2e-6
    assert abs(n32[0] - n64[0]) <= 1e-9 * abs(n64[0])
    assert abs(n64[0] - w['mol'].nelectron) < 1e-3 * w['mol'].nelectron   # the grid integrates
    if name == 'water27':
        assert kept <= 0.5


@pytest.mark.parametrize('precision,rtol', [('fp64', 1e-11), ('mixed', 1e-6)])
def test_vv10_sums_of_the_water27_density_match_the_oracle(ctx, precision, rtol):
    w = workload(ctx, 'water27')
    st = vv10.NlcState(ctx, w['mol'])
    rho = xc.xc_rho(st, w['g'], w['dm'], False)
    keep = numpy.flatnonzero(rho[0] >= 1e-8)
    q, W0, kappa = oracle.vv10_pointwise(rho[0, keep], rho[1:4, keep], w['g'].weights[keep],
                                         NLC)[:3]
    coords = w['g'].coords[keep]
    got = vv10.Vv10Kernel(ctx, precision=precision).kernel_sum(coords, q, W0, kappa)
    outer = numpy.sort(numpy.random.default_rng(2).choice(len(keep), 1024, replace=False))
    ref = oracle.vv10_sum_outer(coords, q, W0, kappa, outer)
    for a, r, name in zip(got, ref, 'FUW'):
        rel = numpy.abs(a[outer] - r) / numpy.abs(r)
        print('%s %s: %d points, max rel %.2e' % (precision, name, len(keep), rel.max()))
        assert rel.max() < rtol, (name, rel.max())
