"""Pins the FP64 oracle (oracle/wb_oracle.c) against independent checks.

The reference's own parity tests compare against live PySCF (absent here);
these checks replace them with what the mathematics fixes:
  * Boys vs scipy's incomplete-gamma closed form (tests/gpu/test_boys.py:9-17)
  * ESP blocks vs an independent Obara-Saika recursion (1/r and erf),
    vs the analytic s-s formula and vs brute-force quadrature
  * symmetry, translation invariance, the 1/d far field
    (tests/gpu/test_shellpairs.py:59-71), erf -> 1/r as omega -> inf
  * AO normalization (quadrature S = analytic S) and finite-difference
    gradients
  * VV10 sums vs a literal double loop (tests/gpu/test_xc_vv10.py:263-292)
  * K assembly vs the dense tensor contraction
"""
import numpy
import pytest

import oracle
from independent import boys_closed, os_esp_shellpair, vv10_literal
from pyscf_wb97mv_fast_amd import cart2sph
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole


def _tiny_basis(ls, centers, exps, coefs):
    nprim = [len(e) for e in exps]
    norm = []
    for l, e, c in zip(ls, exps, coefs):
        norm.append(mole.normalize_contracted(l, numpy.array(e),
                                              numpy.array(c)[:, None])[:, 0])
    return oracle.OracleBasis(ls, nprim, numpy.array(centers, dtype=float),
                              numpy.concatenate(exps), numpy.concatenate(norm)), norm


def test_boys_vs_closed_form():
    T = numpy.concatenate([numpy.linspace(0, 50, 2001), numpy.geomspace(50, 1e6, 200),
                           [1e-14, 29.999, 30.0, 30.001]])
    F = oracle.boys(6, T)
    for n in range(7):
        ref = boys_closed(n, T)
        assert numpy.max(numpy.abs(F[:, n] - ref) / ref) < 2e-13, n


def test_boys_recursion_identity():
    # F_{n}(T) = (2T F_{n+1} + e^{-T}) / (2n+1): the downward relation
    T = numpy.geomspace(1e-3, 200, 300)
    F = oracle.boys(6, T)
    for n in range(6):
        lhs = F[:, n]
        rhs = (2 * T * F[:, n + 1] + numpy.exp(-T)) / (2 * n + 1)
        assert numpy.allclose(lhs, rhs, rtol=1e-12, atol=0)


def test_cart2sph_tables_agree_and_are_orthonormal():
    u, w = grids.angular_product(21)
    for l in range(4):
        T = cart2sph.cart2sph(l)
        assert numpy.allclose(T, oracle.C2S[l], rtol=0, atol=1e-15)
        comps = cart2sph.cart_components(l)
        mono = numpy.stack([u[:, 0] ** a * u[:, 1] ** b * u[:, 2] ** c
                            for a, b, c in comps], axis=1)
        Y = mono @ T.T
        assert numpy.allclose((Y * w[:, None]).T @ Y, numpy.eye(2 * l + 1), atol=1e-12)


@pytest.mark.parametrize('omega', [0.0, 0.3, 1.7])
def test_esp_vs_obara_saika(omega):
    rng = numpy.random.default_rng(3)
    for la in range(4):
        for lb in range(la + 1):
            A = rng.normal(size=3)
            B = A + rng.normal(size=3) * 0.8
            ea = list(rng.uniform(0.2, 6.0, size=2))
            eb = list(rng.uniform(0.2, 6.0, size=2 if lb < 3 else 1))
            ca = list(rng.uniform(0.3, 1.0, size=len(ea)))
            cb = list(rng.uniform(0.3, 1.0, size=len(eb)))
            basis, norm = _tiny_basis([la, lb], [A, B], [ea, eb], [ca, cb])
            pts = numpy.vstack([A + 0.3 * rng.normal(size=(3, 3)),
                                (A + B) / 2, B, A + [7.0, 1.0, -2.0]])
            dense = oracle.esp_dense(basis, pts, omega)
            na = 2 * la + 1
            for g, C in enumerate(pts):
                ref = os_esp_shellpair(la, ea, norm[0], A, lb, eb, norm[1], B, C,
                                       oracle.C2S[la], oracle.C2S[lb], omega)
                got = dense[g, :na, na:]
                assert numpy.allclose(got, ref, rtol=1e-10, atol=1e-12 * numpy.abs(ref).max()), \
                    (la, lb, g)


def test_esp_ss_analytic():
    from scipy.special import erf
    a, b = 1.3, 0.45
    A, B, C = numpy.array([0.1, 0.2, -0.3]), numpy.array([1.0, -0.4, 0.5]), numpy.array([0.7, 0.9, 0.2])
    basis, norm = _tiny_basis([0, 0], [A, B], [[a], [b]], [[1.0], [1.0]])
    p = a + b
    P = (a * A + b * B) / p
    Kab = numpy.exp(-a * b / p * ((A - B) ** 2).sum())
    y00 = 0.5 / numpy.sqrt(numpy.pi)
    for omega in (0.0, 0.3):
        th = 1.0 if omega == 0 else omega ** 2 / (omega ** 2 + p)
        T = th * p * ((P - C) ** 2).sum()
        F0 = 0.5 * numpy.sqrt(numpy.pi / T) * erf(numpy.sqrt(T))
        ref = norm[0][0] * norm[1][0] * y00 * y00 * 2 * numpy.pi / p * Kab * numpy.sqrt(th) * F0
        got = oracle.esp_dense(basis, C[None], omega)[0, 0, 1]
        assert abs(got - ref) < 1e-14 * abs(ref)


def test_esp_brute_force_quadrature():
    """int chi_a chi_b v(|r - C|) on a dense spherical grid centred on C."""
    A, B, C = numpy.array([0.0, 0.0, 0.0]), numpy.array([0.9, 0.3, -0.2]), numpy.array([0.4, -0.5, 0.6])
    basis, _ = _tiny_basis([2, 1], [A, B], [[0.9], [0.6]], [[1.0], [1.0]])
    r, wr = grids.radial_treutler(160, 1.5)
    u, wa = grids.angular_product(41)
    pts = (C[None, None, :] + r[:, None, None] * u[None, :, :]).reshape(-1, 3)
    w = (wr[:, None] * wa[None, :]).ravel()
    X = oracle.ao_eval(basis, pts)[0]
    dist = numpy.linalg.norm(pts - C, axis=1)
    from scipy.special import erf
    for omega in (0.0, 0.3):
        v = 1.0 / dist if omega == 0 else erf(omega * dist) / dist
        quad = (X * (w * v)[:, None]).T @ X
        got = oracle.esp_dense(basis, C[None], omega)[0]
        assert numpy.allclose(got, quad, atol=2e-8 * numpy.abs(quad).max())


def test_esp_invariants():
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    basis = oracle.OracleBasis.from_mol(mol)
    rng = numpy.random.default_rng(1)
    pts = rng.normal(size=(4, 3)) * 2
    A = oracle.esp_dense(basis, pts)
    assert numpy.allclose(A, A.transpose(0, 2, 1), atol=0)
    # translation invariance
    shift = numpy.array([3.3, -1.2, 0.7])
    mol2 = mole.Mole([(s, tuple(numpy.array(x) + shift * mole.BOHR)) for s, x in geometry.WATER_DIMER],
                     'def2-svp')
    A2 = oracle.esp_dense(oracle.OracleBasis.from_mol(mol2), pts + shift)
    assert numpy.allclose(A, A2, atol=1e-12)
    # far field: A -> S / d  (d = 400 bohr)
    S = guess.overlap_matrix(mol)
    far = mol.atom_coords().mean(axis=0) + numpy.array([400.0, 0, 0])
    Afar = oracle.esp_dense(basis, far[None])[0]
    d = numpy.linalg.norm(mol.atom_coords()[mol._bas[:, 0]].mean(0) - far)
    assert numpy.allclose(Afar * d, S, atol=2e-2)
    # erf -> 1/r as omega -> infinity
    Abig = oracle.esp_dense(basis, pts, omega=1e4)
    assert numpy.allclose(Abig, A, atol=1e-6 * numpy.abs(A).max())


def test_ao_normalization_and_gradient():
    mol = mole.Mole([('O', (0, 0, 0))], 'def2-tzvp')
    basis = oracle.OracleBasis.from_mol(mol)
    g = grids.Grids(mol, level=5).build()
    X = oracle.ao_eval(basis, g.coords)[0]
    Sq = (X * g.weights[:, None]).T @ X
    assert numpy.allclose(Sq, guess.overlap_matrix(mol), atol=1e-7)
    rng = numpy.random.default_rng(2)
    pts = rng.normal(size=(5, 3))
    ao = oracle.ao_eval(basis, pts, deriv=1)
    h = 1e-5
    for ax in range(3):
        e = numpy.zeros(3); e[ax] = h
        fd = (oracle.ao_eval(basis, pts + e)[0] - oracle.ao_eval(basis, pts - e)[0]) / (2 * h)
        assert numpy.allclose(ao[ax + 1], fd, atol=1e-7 * numpy.abs(ao[1:]).max())


def test_vv10_sum_vs_literal_loop():
    rng = numpy.random.default_rng(5)
    n = 140
    coords = rng.normal(size=(n, 3)) * 3
    q = rng.uniform(0.01, 1.0, n)
    W0 = rng.uniform(0.1, 2.0, n)
    kappa = rng.uniform(0.1, 1.0, n)
    got = oracle.vv10_sum(coords, q, W0, kappa)
    ref = vv10_literal(coords, q, W0, kappa)
    for a, b in zip(got, ref):
        assert numpy.allclose(a, b, rtol=1e-13)


@pytest.mark.parametrize('omega', [0.0, 0.3])
def test_k_build_vs_dense_contraction(omega):
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    basis = oracle.OracleBasis.from_mol(mol)
    rng = numpy.random.default_rng(7)
    pts = mol.atom_coords()[rng.integers(0, mol.natm, 60)] + rng.normal(size=(60, 3))
    w = rng.uniform(0.01, 0.3, 60)
    D = guess.huckel_density(mol)
    K = oracle.k_build(basis, pts, w, D, omega)
    X = oracle.ao_eval(basis, pts)[0]
    A = oracle.esp_dense(basis, pts, omega)
    F = X @ D
    G = numpy.einsum('gmn,gm->gn', A, F)
    ref = X.T @ (w[:, None] * G)
    assert numpy.allclose(K, ref, atol=1e-13 * numpy.abs(ref).max())


def test_screened_oracle_k_is_the_unscreened_within_the_tolerance():
    """The CPU baseline's screened K build (per-point pair skipping at the
    device build's 1e-11) leaves out only what the unscreened build adds
    below that scale, and skips a large share of the pair evaluations."""
    import oracle
    from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole
    mol = mole.Mole(geometry.water_cluster(4, seed=1), 'def2-svp')
    g = grids.Grids(mol, level=0).build()
    dm = guess.huckel_density(mol)
    ob = oracle.OracleBasis.from_mol(mol)
    sel = numpy.random.default_rng(3).choice(g.size, 64, replace=False)
    for omega in (0.0, 0.3):
        K0 = oracle.k_build(ob, g.coords[sel], g.weights[sel], dm, omega, pair_cut=1e-14)
        K1, kept = oracle.k_build_screened(ob, g.coords[sel], g.weights[sel], dm, omega,
                                           pair_cut=1e-14, tol=1e-11)
        npairs = mol.nbas * (mol.nbas + 1) // 2
        assert 0 < kept < 0.8 * npairs * len(sel), kept
        assert numpy.abs(K1 - K0).max() <= 1e-8 * numpy.abs(K0).max()
        # no skipping at tol 0
        K2, kept2 = oracle.k_build_screened(ob, g.coords[sel], g.weights[sel], dm, omega,
                                            pair_cut=1e-14, tol=0.0)
        assert kept2 > kept and numpy.allclose(K2, K0, rtol=0, atol=1e-13 * numpy.abs(K0).max())
