"""Generally contracted shells (nctr > 1) on the CPU: the packing of the C
ABI and the oracle against the closed relations to the uncontracted basis
(tests/gc_basis.py): X_gc = X_prim C and K_gc[D] = C^T K_prim[C D C^T] C.
The reference keeps such shells whole, with nctr (2l+1) <= 16
(gpu/shellpairs.py:43-48, :188-212); here they are split into one
segmented shell per column, which must give the same function set in the
same (contraction-major) order."""
import numpy

import oracle
from gc_basis import molecules
from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess


def _setup(npts=300, seed=0):
    mg, mp, C = molecules(geometry.WATER_DIMER)
    g = grids.Grids(mg, level=0).build()
    rng = numpy.random.default_rng(seed)
    sel = numpy.sort(rng.choice(g.size, npts, replace=False))
    return mg, mp, C, g.coords[sel], g.weights[sel]


def test_general_contraction_is_present_and_packed_per_column():
    mg, mp, C = molecules(geometry.WATER_DIMER)
    assert max(mg.bas_nctr(i) for i in range(mg.nbas)) == 3
    sa = pack_mol(mg)
    assert sa.nao == C.shape[1] == mg.ao_loc_nr()[-1]
    # every column of a shell shares its primitives and centre
    assert len(sa.l) == sum(mg.bas_nctr(i) for i in range(mg.nbas))


def test_oracle_ao_values_are_the_contracted_primitives():
    mg, mp, C, coords, _ = _setup()
    Xg = oracle.ao_eval(oracle.OracleBasis.from_mol(mg), coords, deriv=1)
    Xp = oracle.ao_eval(oracle.OracleBasis.from_mol(mp), coords, deriv=1)
    ref = Xp @ C
    assert numpy.abs(Xg - ref).max() < 1e-12 * numpy.abs(ref).max()


def test_oracle_k_is_the_contracted_primitive_k():
    mg, mp, C, coords, w = _setup(npts=120)
    dm = guess.huckel_density(mg)
    for omega in (0.0, 0.3):
        Kg = oracle.k_build(oracle.OracleBasis.from_mol(mg), coords, w, dm, omega)
        Kp = oracle.k_build(oracle.OracleBasis.from_mol(mp), coords, w, C @ dm @ C.T, omega)
        ref = C.T @ Kp @ C
        assert numpy.abs(Kg - ref).max() < 1e-11 * numpy.abs(ref).max(), omega
