"""The product path against what the mathematics fixes (the checks that pin
the oracle, tests/test_oracle_pins.py, applied to the HIP path directly):

* the K hook (get_k_only_gpu, both precisions) on an s-only molecule
  converges to the closed-form exchange K = sum (ml|ns) D_ls of the (ss|ss)
  integrals, for 1/r and erf(0.3 r)/r, and with the overlap-fitted
  projector Q = S_grid^-1 S (PJS) it is closer still;
* the device VV10 flow's vmat is the derivative of its energy,
  d E_nl(D + e S)/de = tr(vmat S) (central differences, FP64 form);
* the device VV10 kernel sums make beta + F/2 vanish for a uniform density
  (the defining property of beta).
"""
import numpy
import pytest

import oracle
from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.gpu import sgx_k, vv10
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole, sgx
from test_oracle_pins import s_only_system, ss_eri, uniform_gas_exc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    c = lib.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_k_hook_converges_to_the_analytic_exchange(ctx, precision):
    mol = s_only_system()
    rng = numpy.random.default_rng(1)
    D = rng.normal(size=(mol.nao, mol.nao))
    D = 0.5 * (D + D.T)
    S = guess.overlap_matrix(mol)
    ob = oracle.OracleBasis.from_mol(mol)
    kb = sgx_k.GpuKBuilder(mol, ctx, tile_tol=0.0, precision=precision)
    for omega in (0.0, 0.3):
        exact = numpy.einsum('mlns,ls->mn', ss_eri(mol, omega), D)
        scale = numpy.abs(exact).max()
        err = {}
        for lev in (2, 4):
            g = grids.Grids(mol, level=lev).build()
            X = oracle.ao_eval(ob, g.coords)[0]
            Q = numpy.linalg.solve((X * g.weights[:, None]).T @ X, S)
            for name, s in (('I', sgx.SyntheticSGX(mol, g)), ('Q', sgx.SyntheticSGX(mol, g, Q=Q))):
                with mol.with_range_coulomb(omega):
                    K = sgx_k.get_k_only_gpu(kb, s, D)
                err[lev, name] = numpy.abs(K - exact).max() / scale
        print(precision, omega, {k: '%.1e' % v for k, v in err.items()})
        assert err[4, 'I'] < 5e-7 and err[4, 'I'] < 0.1 * err[2, 'I']
        assert err[4, 'Q'] < 2e-7 and err[2, 'Q'] < err[2, 'I']


def test_nlc_vmat_is_the_derivative_of_its_energy(ctx):
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = grids.Grids(mol, level=0).build()
    rng = numpy.random.default_rng(0)
    sel = numpy.sort(rng.choice(g.size, 2500, replace=False))
    g.coords, g.weights = g.coords[sel], g.weights[sel]
    D = guess.huckel_density(mol)
    st = vv10.NlcState(ctx, mol)
    ni = sgx.SyntheticNumInt()
    _, _, vmat = vv10.gpu_nr_nlc_vxc(st, ni, mol, g, 'wb97m-v', D, precision='fp64')
    for seed in (1, 2):
        S = numpy.random.default_rng(seed).normal(size=D.shape)
        S = 0.005 * (S + S.T)
        h = 1e-3
        ep = vv10.gpu_nr_nlc_vxc(st, ni, mol, g, 'wb97m-v', D + h * S, precision='fp64')[1]
        em = vv10.gpu_nr_nlc_vxc(st, ni, mol, g, 'wb97m-v', D - h * S, precision='fp64')[1]
        fd = (ep - em) / (2 * h)
        an = numpy.einsum('ij,ij->', vmat, S)
        print('fd %.12e analytic %.12e' % (fd, an))
        assert abs(fd - an) < 1e-7 * abs(an)


@pytest.mark.parametrize('precision,tol', [('fp64', 1e-11), ('mixed', 2e-6)])
def test_vv10_sums_cancel_beta_for_a_uniform_gas(ctx, precision, tol):
    k = vv10.Vv10Kernel(ctx, precision=precision)
    for n in (1e-3, 0.3, 20.0):
        beta, coords, q, W0, kappa = uniform_gas_exc(n)
        F = k.kernel_sum(coords, q, W0, kappa)[0][0]
        print('%s n=%g beta %.6e F/2 %.6e' % (precision, n, beta, 0.5 * F))
        assert abs(beta + 0.5 * F) < tol * beta
