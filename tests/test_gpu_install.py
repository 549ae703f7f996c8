"""The installed boundary on the device: PySCF's get_veff (here the
stand-in's, which combines the terms as rks.get_veff does) with the K hook
through the repaired RSH branch of SGX.get_jk, the nr_rks and nr_nlc_vxc
hooks, and the overlapped get_veff (XC and VV10 on a worker thread and the
XC context while the calling thread runs J and K on the K context) equals
the serial get_veff, without a fallback to the CPU originals (the
reference's tests/gpu/test_install.py:222-262)."""
import warnings

import numpy
import pytest

from pyscf_wb97mv_fast_amd.gpu import install
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole, sgx
from test_gpu_xc import ToyFunctional

pytestmark = pytest.mark.gpu


class _NI(sgx.SyntheticNumInt):
    toy = ToyFunctional('MGGA')

    def eval_xc(self, xc_code, rho, spin=0, relativity=0, deriv=1, omega=None, verbose=None):
        return self.toy.eval_xc(xc_code, rho, spin, deriv)


def _mf():
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = grids.Grids(mol, level=0).build()
    rng = numpy.random.default_rng(2)
    sel = numpy.sort(rng.choice(g.size, 2000, replace=False))
    g.coords, g.weights = g.coords[sel], g.weights[sel]
    mf = sgx.SyntheticRKS(mol, g, g)
    mf._numint = _NI()
    mf.with_df.j_builder = lambda s, dm: 0.5 * (dm + dm.T)
    return mf


def test_overlapped_get_veff_on_the_device_equals_serial():
    mf = _mf()
    dm = guess.huckel_density(mf.mol)
    out = {}
    for overlap in (False, True):
        mf.with_df._rsh_df.clear()
        mf.with_df._pjs_data = None
        hooks = install.install_gpu(mf, backend='hip', k=True, sgx_module=sgx.sgx_jk,
                                    overlap=overlap, rsh_fix=True, sgx_class=sgx.SyntheticSGX)
        try:
            with warnings.catch_warnings():
                warnings.simplefilter('error')            # no fallback to the CPU originals
                out[overlap] = mf.get_veff(mf.mol, dm)
            # one K call computed both operators; the LR copy's call was served
            assert len(hooks.k_stats) == 2 and hooks.k_stats[1].get('cached')
            assert mf.with_df._rsh_df['0.300000'].j_calls == 0
            if overlap:
                assert len(hooks.timings) == 1
        finally:
            hooks.restore_all()
    a, b = out[False], out[True]
    print('max |dV| %.2e, dE_xc %.2e' % (numpy.abs(numpy.asarray(a) - b).max(), a.exc - b.exc))
    assert numpy.allclose(numpy.asarray(a), numpy.asarray(b), rtol=0, atol=1e-12)
    assert abs(a.exc - b.exc) < 1e-12 and abs(a.ecoul - b.ecoul) < 1e-12
    assert numpy.array_equal(a.vk, b.vk)
