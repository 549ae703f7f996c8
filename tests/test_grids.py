"""Pins the host Stratmann-Scuseria-Frisch partition (systems/grids.py), the
reference for the device partition (wb_grid_partition), against checks
that do not restate it: a brute-force cell weight over all atoms with the
switching polynomial solved from its defining conditions, partition of
unity, the two-atom midplane, and molecular-grid quadrature of Gaussians
against their closed-form integrals."""
import numpy

from independent import ssf_brute, ssf_polynomial
from pyscf_wb97mv_fast_amd.systems import geometry, grids, mole


def test_ssf_polynomial_is_the_published_one():
    # Chem. Phys. Lett. 257, 213 (1996), eq. 14: (35x - 35x^3 + 21x^5 - 5x^7)/16
    assert numpy.allclose(ssf_polynomial() * 16, [35, -35, 21, -5], atol=1e-12)


def _random_case(seed, natm=6, n=400):
    rng = numpy.random.default_rng(seed)
    atoms = rng.uniform(-3, 3, (natm, 3))
    owner = rng.integers(0, natm, n)
    coords = atoms[owner] + rng.normal(size=(n, 3)) * rng.uniform(0.1, 3, (n, 1))
    return coords, atoms, owner


def test_ssf_host_matches_brute_force():
    for seed in range(3):
        coords, atoms, owner = _random_case(seed)
        got = grids._ssf_partition(coords, atoms, owner, chunk=97)
        ref = ssf_brute(coords, atoms, owner)
        assert numpy.allclose(got, ref, rtol=0, atol=1e-13), seed


def test_ssf_partition_of_unity_and_midplane():
    coords, atoms, _ = _random_case(7, natm=5, n=300)
    tot = sum(grids._ssf_partition(coords, atoms, numpy.full(len(coords), a)) for a in range(5))
    assert numpy.allclose(tot, 1.0, atol=1e-14)
    A, B = numpy.array([0.0, 0, 0]), numpy.array([1.7, 0.4, -0.3])
    rng = numpy.random.default_rng(2)
    # points on the perpendicular bisector plane of A-B: weight 1/2 each
    n = numpy.cross(B - A, rng.normal(size=(20, 3)))
    mid = 0.5 * (A + B) + n
    w = grids._ssf_partition(mid, numpy.stack([A, B]), numpy.zeros(20, dtype=int))
    assert numpy.allclose(w, 0.5, atol=1e-14)


def test_molecular_grid_integrates_gaussians():
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = grids.Grids(mol, level=5).build()
    atoms = mol.atom_coords()
    for alpha in (0.3, 1.0, 4.0):
        f = sum(numpy.exp(-alpha * ((g.coords - A) ** 2).sum(axis=1)) for A in atoms)
        exact = len(atoms) * (numpy.pi / alpha) ** 1.5
        # the quadrature error falls with the level: 1e-4 at level 1, 1e-7 at 5
        assert abs(g.weights @ f - exact) < 3e-7 * exact, alpha
