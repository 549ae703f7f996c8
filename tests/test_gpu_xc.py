"""Semilocal XC flow (wb_xc_rho / wb_xc_vmat, gpu/xc.py:gpu_nr_rks) against
dense FP64 numpy on the oracle's AO values (rho, grad rho, tau; the potential
matrix for a given potential), and a property that needs no restatement:
excsum(D) from gpu_nr_rks has the derivative tr(vmat dD) (central
differences) for LDA, GGA and meta-GGA forms of a test functional.  The
functional itself is PySCF's (libxc) in real use and is not part of this
path (the reference's gpu/xc.py:130-201 calls it the same way)."""
import numpy
import pytest

import oracle
from pyscf_wb97mv_fast_amd import lib
from pyscf_wb97mv_fast_amd.gpu import vv10, xc
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    return lib.Context(0)


@pytest.fixture(scope='module')
def system(ctx):
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = grids.Grids(mol, level=0).build()
    rng = numpy.random.default_rng(3)
    sel = numpy.sort(rng.choice(g.size, 1500, replace=False))
    g.coords, g.weights = g.coords[sel], g.weights[sel]
    dm = guess.huckel_density(mol)
    st = vv10.NlcState(ctx, mol)
    ao = oracle.ao_eval(oracle.OracleBasis.from_mol(mol), g.coords, deriv=1)
    return mol, g, dm, st, ao


def _rho_ref(ao, dm):
    t0 = ao[0] @ dm
    rho = numpy.zeros((5, ao.shape[1]))
    rho[0] = numpy.einsum('pm,pm->p', t0, ao[0])
    for d in range(3):
        rho[1 + d] = 2.0 * numpy.einsum('pm,pm->p', t0, ao[1 + d])
        rho[4] += 0.5 * numpy.einsum('pm,pm->p', ao[1 + d] @ dm, ao[1 + d])
    return rho


@pytest.mark.parametrize('mgga', [False, True])
def test_xc_rho_matches_dense(system, mgga):
    mol, g, dm, st, ao = system
    rho = xc.xc_rho(st, g, dm, mgga)
    ref = _rho_ref(ao, dm)
    if not mgga:
        ref[4] = 0.0
    for k in range(5):
        scale = max(numpy.abs(ref[k]).max(), 1e-300)
        assert numpy.abs(rho[k] - ref[k]).max() < 1e-12 * scale + 1e-300, k


@pytest.mark.parametrize('mgga', [False, True])
def test_xc_vmat_matches_dense(system, mgga):
    mol, g, dm, st, ao = system
    rng = numpy.random.default_rng(7)
    v = rng.normal(size=(5, g.size))
    w = g.weights
    ref = numpy.einsum('p,pm,pn->mn', w * v[0], ao[0], ao[0])
    for d in range(3):
        a = numpy.einsum('p,pm,pn->mn', w * v[1 + d], ao[1 + d], ao[0])
        ref += a + a.T
        if mgga:
            ref += 0.5 * numpy.einsum('p,pm,pn->mn', w * v[4], ao[1 + d], ao[1 + d])
    got = xc.xc_vmat(st, g, v, mgga)
    assert numpy.abs(got - ref).max() < 1e-12 * numpy.abs(ref).max()
    assert numpy.array_equal(got, got.T) or numpy.abs(got - got.T).max() < 1e-14 * numpy.abs(ref).max()


class ToyFunctional:
    """f(rho, sigma, tau) = -A rho^(4/3) + B sigma / (rho + d) + C tau^2 / (rho + d)
    with its exact first derivatives (libxc's convention: exc = f / rho)."""
    A, B, C, D = 0.75, 0.004, 0.02, 1e-3

    def __init__(self, xctype):
        self.xctype = xctype

    def _xc_type(self, xc_code):
        return self.xctype

    def energy_density(self, rho5):
        r = rho5[0]
        f = -self.A * numpy.cbrt(r) ** 4
        if self.xctype in ('GGA', 'MGGA'):
            f = f + self.B * (rho5[1:4] ** 2).sum(0) / (r + self.D)
        if self.xctype == 'MGGA':
            f = f + self.C * rho5[4] ** 2 / (r + self.D)
        return f

    def eval_xc(self, xc_code, rho, spin=0, deriv=1):
        rho = numpy.asarray(rho)
        r5 = numpy.zeros((5, rho.shape[-1]))
        if rho.ndim == 1:
            r5[0] = rho
        else:
            r5[:rho.shape[0]] = rho
        r = r5[0]
        sigma = (r5[1:4] ** 2).sum(0)
        tau = r5[4]
        f = self.energy_density(r5)
        exc = f / numpy.maximum(r, 1e-300)
        vrho = -4.0 / 3.0 * self.A * numpy.cbrt(r)
        vsigma = vtau = None
        if self.xctype in ('GGA', 'MGGA'):
            vrho = vrho - self.B * sigma / (r + self.D) ** 2
            vsigma = self.B / (r + self.D)
        if self.xctype == 'MGGA':
            vrho = vrho - self.C * tau ** 2 / (r + self.D) ** 2
            vtau = 2.0 * self.C * tau / (r + self.D)
        return exc, (vrho, vsigma, None, vtau), None, None


@pytest.mark.parametrize('xctype', ['LDA', 'GGA', 'MGGA'])
def test_gpu_nr_rks_energy_derivative(system, xctype):
    mol, g, dm, st, ao = system
    ni = ToyFunctional(xctype)
    nelec, exc, vmat = xc.gpu_nr_rks(st, ni, mol, g, 'toy', dm)
    ref_rho = _rho_ref(ao, dm)
    assert abs(nelec - float((ref_rho[0] * g.weights).sum())) < 1e-10
    assert abs(exc - float((ni.energy_density(ref_rho) * g.weights).sum())) < 1e-10 * abs(exc)
    rng = numpy.random.default_rng(11)
    S = rng.normal(scale=1e-2, size=dm.shape)
    S = S + S.T
    eps = 1e-4
    ep = xc.gpu_nr_rks(st, ni, mol, g, 'toy', dm + eps * S)[1]
    em = xc.gpu_nr_rks(st, ni, mol, g, 'toy', dm - eps * S)[1]
    fd = (ep - em) / (2 * eps)
    an = float(numpy.sum(vmat * S))
    print('%s dE/dD: finite difference %.10e, tr(vmat S) %.10e' % (xctype, fd, an))
    assert abs(fd - an) < 1e-7 * max(abs(an), 1e-3)


def test_nr_rks_hook_dispatches_to_the_device_flow(system):
    import warnings
    from pyscf_wb97mv_fast_amd.gpu import install
    from pyscf_wb97mv_fast_amd.systems import sgx
    mol, g, dm, st, ao = system

    class NI(sgx.SyntheticNumInt):
        toy = ToyFunctional('MGGA')

        def _xc_type(self, xc_code):
            return 'MGGA'

        def eval_xc(self, xc_code, rho, spin=0, relativity=0, deriv=1, omega=None,
                    verbose=None):
            return self.toy.eval_xc(xc_code, rho, spin, deriv)

    mf = sgx.SyntheticRKS(mol, g, g)
    mf._numint = NI()
    hooks = install.install_gpu(mf, backend='hip')
    try:
        with warnings.catch_warnings():
            warnings.simplefilter('error')            # no fallback to the CPU original
            n, e, v = mf._numint.nr_rks(mol, g, 'toy', dm)
    finally:
        hooks.restore_all()
    ref = xc.gpu_nr_rks(st, NI(), mol, g, 'toy', dm)
    assert n == ref[0] and e == ref[1] and numpy.array_equal(v, ref[2])
