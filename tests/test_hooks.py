"""Hook mechanics on the CPU (tests/gpu/test_install.py:59-155, 287-381 of
the reference): install / restore identity, fallback with a RuntimeWarning
to the original PySCF callable, dispatch only for mf.with_df and its
long-range copies, the OOM halving sequence."""
import types
import warnings

import numpy
import pytest

from pyscf_wb97mv_fast_amd.core.hooks import HookSet
from pyscf_wb97mv_fast_amd.errors import GpuOOM
from pyscf_wb97mv_fast_amd.gpu import install
from pyscf_wb97mv_fast_amd.systems import geometry, grids, mole, sgx


class _NumInt(sgx.SyntheticNumInt):
    calls = 0
    rks_calls = 0

    def nr_nlc_vxc(self, mol, grids, xc_code, dm, relativity=0, hermi=1, max_memory=2000,
                   verbose=None):
        type(self).calls += 1
        return 1.0, 2.0, numpy.zeros_like(dm)

    def nr_rks(self, mol, grids, xc_code, dms, relativity=0, hermi=1, max_memory=2000,
               verbose=None):
        type(self).rks_calls += 1
        return 3.0, 4.0, numpy.zeros_like(dms)


def _mf():
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = grids.Grids(mol, level=0)
    g.coords, g.weights = numpy.zeros((64, 3)), numpy.ones(64)
    mf = sgx.SyntheticRKS(mol, g, g)
    mf._numint = _NumInt()
    return mf


def _fake_sgx_module():
    calls = []

    def get_k_only(sgx_obj, dm, hermi=1, direct_scf_tol=1e-13):
        calls.append(sgx_obj)
        return numpy.full_like(dm, 7.0)
    mod = types.SimpleNamespace(get_k_only=get_k_only)
    mod.calls = calls
    return mod


def test_hookset_restores_module_and_instance_attributes():
    obj = _NumInt()
    mod = types.SimpleNamespace(f=lambda: 1)
    orig_f = mod.f
    h = HookSet()
    h.wrap(obj, 'nr_nlc_vxc', lambda orig: (lambda *a, **k: 'hooked'))
    h.wrap(mod, 'f', lambda orig: (lambda: 2))
    assert obj.nr_nlc_vxc() == 'hooked' and mod.f() == 2
    h.restore_all()
    assert 'nr_nlc_vxc' not in vars(obj)          # class method visible again
    assert mod.f is orig_f
    with pytest.raises(RuntimeError):
        h.wrap(obj, 'nr_nlc_vxc', lambda o: o)


def test_install_without_gpu_falls_back_with_warning():
    mf = _mf()
    mod = _fake_sgx_module()
    hooks = install.install_gpu(mf, backend='hip', k=True, sgx_module=mod)
    try:
        dm = numpy.eye(mf.mol.nao)
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter('always')
            if install.gpu_available():
                pytest.skip('a GPU is present: the fallback path is not exercised')
            r = mf._numint.nr_nlc_vxc(mf.mol, mf.nlcgrids, 'wb97m-v', dm)
            x = mf._numint.nr_rks(mf.mol, mf.nlcgrids, 'wb97m-v', dm)
            k = mod.get_k_only(mf.with_df, dm)
        assert r[0] == 1.0 and _NumInt.calls >= 1
        assert x[0] == 3.0 and _NumInt.rks_calls >= 1
        assert numpy.all(k == 7.0) and mod.calls == [mf.with_df]
        msgs = [str(x.message) for x in w if issubclass(x.category, RuntimeWarning)]
        assert any(m.startswith(install.WARN_PREFIX + ' nr_nlc_vxc') for m in msgs)
        assert any(m.startswith(install.WARN_PREFIX + ' nr_rks') for m in msgs)
        assert any(m.startswith(install.WARN_PREFIX + ' get_k_only') for m in msgs)
    finally:
        hooks.restore_all()
    assert 'nr_nlc_vxc' not in vars(mf._numint) and 'nr_rks' not in vars(mf._numint)


def test_semilocal_false_leaves_nr_rks_alone():
    mf = _mf()
    hooks = install.install_gpu(mf, backend='hip', semilocal=False)
    try:
        assert 'nr_rks' not in vars(mf._numint) and 'nr_nlc_vxc' in vars(mf._numint)
    finally:
        hooks.restore_all()


def test_k_hook_leaves_other_sgx_objects_alone():
    mf = _mf()
    mod = _fake_sgx_module()
    hooks = install.install_gpu(mf, backend='hip', k=True, sgx_module=mod)
    try:
        other = sgx.SyntheticSGX(mf.mol, mf.with_df.grids)
        with warnings.catch_warnings():
            warnings.simplefilter('error')            # no fallback warning
            k = mod.get_k_only(other, numpy.eye(mf.mol.nao))
        assert numpy.all(k == 7.0) and mod.calls == [other]
    finally:
        hooks.restore_all()


def test_long_range_copy_is_dispatched():
    mf = _mf()
    mod = _fake_sgx_module()
    lr = sgx.SyntheticSGX(mf.mol, mf.with_df.grids)
    mf.with_df._rsh_df['0.300000'] = lr
    hooks = install.install_gpu(mf, backend='hip', k=True, sgx_module=mod)
    try:
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter('always')
            mod.get_k_only(lr, numpy.eye(mf.mol.nao))
        if not install.gpu_available():
            assert any(str(x.message).startswith(install.WARN_PREFIX) for x in w)
    finally:
        hooks.restore_all()


def test_oom_retry_halves_copies_from_current_and_keeps_them():
    state = {'n': 6, 'calls': []}

    def get():
        return state['n']

    def put(n):
        state['n'] = n

    def flow():
        state['calls'].append(state['n'])
        raise GpuOOM('out of memory')
    with pytest.raises(GpuOOM):
        install.run_with_oom_retry(flow, get, put)
    assert state['calls'] == [6, 3, 1]                 # starts from the context's count
    state.update(n=8, calls=[])

    def flaky():
        state['calls'].append(state['n'])
        if len(state['calls']) < 3:
            raise MemoryError('out of memory')
        return 'ok'
    assert install.run_with_oom_retry(flaky, get, put) == 'ok'
    assert state['calls'] == [8, 4, 2] and state['n'] == 2   # the working count stays
    state['calls'] = []
    with pytest.raises(ValueError):
        install.run_with_oom_retry(lambda: (_ for _ in ()).throw(ValueError('x')), get, put)
    assert state['n'] == 2


def test_no_gpu_installs_nothing():
    if install.gpu_available():
        pytest.skip('GPU present')
    mf = _mf()
    with pytest.warns(RuntimeWarning):
        hooks = install.install_gpu(mf)
    assert 'nr_nlc_vxc' not in vars(mf._numint) and 'nr_rks' not in vars(mf._numint)
    hooks.restore_all()


class _RecordingBuilder:
    """GpuKBuilder with the device call replaced by a record of what was
    asked (raw K = omega * I + trace(qd) so results are distinguishable)."""

    def __new__(cls):
        from pyscf_wb97mv_fast_amd.gpu import sgx_k

        class B(sgx_k.GpuKBuilder):
            def __init__(self):
                super().__init__(mol=None, ctx=object())
                self.calls = []

            def k_raw(self, grids, qd, omegas):
                self.calls.append(list(omegas))
                return numpy.array([w * numpy.eye(qd.shape[0]) + numpy.trace(qd)
                                    for w in omegas])
        return B()


def test_companion_operators_evaluated_together_and_cached():
    b = _RecordingBuilder()
    g1, g2 = object(), object()
    qd = numpy.diag([1.0, 2.0])
    k0 = b.k_raw_shared(g1, qd, 0.0, companions=(0.0, 0.3))
    assert b.calls == [[0.0, 0.3]]
    k1 = b.k_raw_shared(g1, qd.copy(), 0.3, companions=(0.0, 0.3))
    assert b.calls == [[0.0, 0.3]]                           # served from the first call
    assert numpy.allclose(k1, 0.3 * numpy.eye(2) + 3.0) and numpy.allclose(k0, 3.0)
    b.k_raw_shared(g1, qd, 0.3, companions=(0.0, 0.3))        # taken once only
    assert b.calls[-1] == [0.3, 0.0]
    b.k_raw_shared(g1, qd + 1e-15, 0.0, companions=(0.0, 0.3))  # another density
    assert b.calls[-1] == [0.0, 0.3]
    b.k_raw_shared(g2, qd + 1e-15, 0.3, companions=(0.0, 0.3))  # another grid
    assert b.calls[-1] == [0.3, 0.0]


def test_rsh_companions_from_functional_and_rsh_copies():
    from pyscf_wb97mv_fast_amd.gpu import sgx_k
    mf = types.SimpleNamespace(xc='wb97m-v', _numint=sgx.SyntheticNumInt(),
                               with_df=types.SimpleNamespace(_rsh_df={}))
    assert sgx_k.rsh_companions(mf) == (0.0, 0.3)
    mf.with_df._rsh_df['%.6f' % 0.4] = object()
    assert sgx_k.rsh_companions(mf) == (0.0, 0.3, 0.4)
    mf2 = types.SimpleNamespace(xc='pbe', _numint=sgx.SyntheticNumInt(), with_df=None)
    assert sgx_k.rsh_companions(mf2) == (0.0,)
