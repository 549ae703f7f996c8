"""The screening bounds the device uses, checked against exact integrals.

The task screen keeps (sub-block, shell pair) when beta min(1, sqrt(pi) /
(2 sqrt(q) d)) max_g w |F| >= tol, which is safe only if
|A_ab(g)| <= beta min(1, sqrt(pi) / (2 sqrt(q) d)), d = |g - centre| - delta,
for every point g and every element of the pair's block (the reference
checks its host bound the same way, tests/gpu/test_shellpairs.py:36-56).
The primitive-tail skip drops a kept task's trailing primitive pairs whose
c_m / d_m bounds sum to at most tol / 2, which is safe only if each
primitive pair's block obeys |A^(m)(g)| <= c_m / |g - P_m|.  Both are
checked here on the records as the device holds them (FP32, built by the
library's host code, wb_test_pair_bounds), against oracle.esp_dense on
near and far points, for every class up to (f, f) and both operators (the
erf operator is bounded by 1/r's)."""
import ctypes

import numpy
import pytest

import oracle
from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.systems import geometry, mole

SQPI2 = 0.8862269254527580


def _records(mol, pair_cut=1e-14):
    try:
        L = lib.load()
    except Exception as exc:           # noqa: BLE001
        pytest.skip('library not built: %s' % exc)
    p = pack_mol(mol)
    l = numpy.ascontiguousarray(p.l, dtype=numpy.int32)
    npr = numpy.ascontiguousarray(p.nprim, dtype=numpy.int32)
    cen, ex, co = lib.f64(p.center), lib.f64(p.exps), lib.f64(p.coefs)
    cnt = numpy.zeros(2, dtype=numpy.int64)
    args = (len(l), lib.ptr(l), lib.ptr(npr), lib.ptr(cen), lib.ptr(ex), lib.ptr(co),
            float(pair_cut), lib.ptr(cnt))
    lib.check(L.wb_test_pair_bounds(*args, None, None))
    pairs = numpy.zeros((cnt[0], 10))
    prims = numpy.zeros((cnt[1], 6))
    lib.check(L.wb_test_pair_bounds(*args, lib.ptr(pairs), lib.ptr(prims)))
    return p, pairs, prims


def _points(mol, rng, n_near=40, n_far=20):
    xyz = mol.atom_coords()
    near = xyz[rng.integers(0, len(xyz), n_near)] + rng.normal(scale=1.0, size=(n_near, 3))
    d = rng.normal(size=(n_far, 3))
    d /= numpy.linalg.norm(d, axis=1)[:, None]
    far = xyz.mean(0) + d * rng.uniform(6.0, 25.0, (n_far, 1))
    return numpy.vstack([near, far, xyz[:3]])


SYSTEMS = {
    'water_dimer_tzvp': lambda: mole.Mole(geometry.WATER_DIMER, 'def2-tzvp'),
    'drug_like_svp': lambda: mole.Mole(geometry.drug_like(10, seed=2), 'def2-svp'),
}


@pytest.mark.parametrize('name', sorted(SYSTEMS))
def test_pair_bound_covers_every_block(name):
    mol = SYSTEMS[name]()
    p, pairs, _ = _records(mol)
    ob = oracle.OracleBasis.from_mol(mol)
    assert ob.nao == p.nao
    pts = _points(mol, numpy.random.default_rng(3))
    A = numpy.abs(oracle.esp_dense(ob, pts, 0.0))
    classes = set()
    worst = 0.0
    for rec in pairs:
        sa, sb = int(rec[0]), int(rec[1])
        la, lb = ob.l[sa], ob.l[sb]
        classes.add((max(la, lb), min(la, lb)))
        blk = A[:, ob.ao0[sa]:ob.ao0[sa] + 2 * la + 1, ob.ao0[sb]:ob.ao0[sb] + 2 * lb + 1]
        exact = blk.reshape(len(pts), -1).max(axis=1)
        d = numpy.linalg.norm(pts - rec[2:5], axis=1) - rec[5]
        f0 = numpy.where(d > 0, numpy.minimum(1.0, SQPI2 / (numpy.sqrt(rec[7]) * numpy.maximum(d, 1e-300))), 1.0)
        bound = rec[6] * f0
        assert numpy.all(exact <= bound), (sa, sb, (exact / bound).max())
        worst = max(worst, (exact / bound).max())
    print('%s: %d pairs, classes %s, tightest exact/bound %.3f' % (name, len(pairs),
                                                                  sorted(classes), worst))
    if name == 'water_dimer_tzvp':
        assert (3, 3) in classes and (3, 0) in classes
    assert worst > 0.05            # the bound is not vacuous


@pytest.mark.parametrize('name', sorted(SYSTEMS))
def test_primitive_tail_bound_covers_each_primitive_pair(name):
    mol = SYSTEMS[name]()
    p, pairs, prims = _records(mol)
    rng = numpy.random.default_rng(4)
    pts = _points(mol, rng, 30, 20)
    shell_prim0 = numpy.concatenate([[0], numpy.cumsum(p.nprim)[:-1]])
    sel = rng.choice(len(pairs), min(len(pairs), 600), replace=False)
    worst, checked = 0.0, 0
    for rec in pairs[sel]:
        sa, sb = int(rec[0]), int(rec[1])
        for m in range(int(rec[8]), int(rec[8] + rec[9])):
            P, c, ia, ib = prims[m, :3], prims[m, 3], int(prims[m, 4]), int(prims[m, 5])
            ka, kb = shell_prim0[sa] + ia, shell_prim0[sb] + ib
            one = oracle.OracleBasis([p.l[sa], p.l[sb]], [1, 1],
                                     numpy.array([p.center[sa], p.center[sb]]),
                                     numpy.array([p.exps[ka], p.exps[kb]]),
                                     numpy.array([p.coefs[ka], p.coefs[kb]]))
            na = 2 * p.l[sa] + 1
            blk = numpy.abs(oracle.esp_dense(one, pts, 0.0)[:, :na, na:])
            exact = blk.reshape(len(pts), -1).max(axis=1)
            d = numpy.linalg.norm(pts - P, axis=1)
            ok = d > 1e-3
            assert numpy.all(exact[ok] * d[ok] <= c), (sa, sb, m, (exact[ok] * d[ok]).max() / c)
            worst = max(worst, (exact[ok] * d[ok]).max() / c)
            checked += 1
    print('%s: %d primitive pairs, tightest exact*d/c %.3f' % (name, checked, worst))
    assert checked > 500 and worst > 0.05
