"""The drop-in boundary on the CPU: the patch stack (core/hooks.py), the
repaired RSH branch of SGX.get_jk (core/sgx_patch.py; the reference's
tests/core/test_sgx_patch.py checks the same three faults against live
PySCF), the overlapped get_veff (the reference's
tests/gpu/test_install.py:222-262 checks it against the serial get_veff),
and inputs outside the GPU path's scope, which must raise Unsupported before
any device work (the reference's tests/gpu/test_sgx_k.py:247-263,
tests/gpu/test_shellpairs.py:74-92)."""
import threading
import types
import warnings

import numpy
import pytest

from pyscf_wb97mv_fast_amd.basis import pack_mol
from pyscf_wb97mv_fast_amd.core import sgx_patch
from pyscf_wb97mv_fast_amd.core.hooks import HookSet
from pyscf_wb97mv_fast_amd.errors import Unsupported
from pyscf_wb97mv_fast_amd.gpu import install, sgx_k
from pyscf_wb97mv_fast_amd.systems import geometry, grids, mole, sgx


# ---------------------------------------------------------------- HookSet ----
class _Owner:
    def method(self):
        return 'class'

    @staticmethod
    def helper():
        return 'static'


def test_hookset_undoes_in_reverse_order_and_restores_kinds():
    obj = _Owner()
    mod = types.SimpleNamespace(f=lambda: 1)
    orig_f, orig_static = mod.f, vars(_Owner)['helper']
    order = []
    h = HookSet()
    h.on_restore(lambda: order.append('first registered'))
    h.wrap(obj, 'method', lambda orig: (lambda: 'patched ' + orig()))
    h.wrap(mod, 'f', lambda orig: (lambda: orig() + 1))
    h.wrap(_Owner, 'helper', lambda orig: staticmethod(lambda: 'patched'))
    h.on_restore(lambda: order.append('last registered'))
    assert obj.method() == 'patched class' and mod.f() == 2 and _Owner.helper() == 'patched'
    h.restore_all()
    assert order == ['last registered', 'first registered']
    assert 'method' not in vars(obj) and obj.method() == 'class'     # shadow removed
    assert mod.f is orig_f
    assert vars(_Owner)['helper'] is orig_static and _Owner.helper() == 'static'
    with pytest.raises(RuntimeError):
        h.wrap(obj, 'method', lambda o: o)


def test_hookset_keeps_undoing_after_a_failing_action():
    mod = types.SimpleNamespace(a=1, b=2)
    h = HookSet()
    h.wrap(mod, 'a', lambda o: 10)
    h.on_restore(lambda: (_ for _ in ()).throw(KeyError('boom')))
    h.wrap(mod, 'b', lambda o: 20)
    with pytest.raises(KeyError):
        h.restore_all()
    assert (mod.a, mod.b) == (1, 2)


def test_hookset_as_context_manager():
    mod = types.SimpleNamespace(g=lambda: 'orig')
    with HookSet() as h:
        h.wrap(mod, 'g', lambda o: (lambda: 'new'))
        assert mod.g() == 'new'
    assert mod.g() == 'orig'


# ------------------------------------------------------------- sgx_patch ----
@pytest.fixture
def rsh_setup(monkeypatch):
    """Water dimer stand-in SGX with a J builder that counts its calls and a
    get_k_only that records which object, omega and tolerance it saw (and
    builds the PJS data as sgx_jk.get_k_only does)."""
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = types.SimpleNamespace(coords=numpy.zeros((64, 3)), weights=numpy.ones(64))
    obj = sgx.SyntheticSGX(mol, g, j_builder=lambda s, dm: 2.0 * dm)
    calls = []

    def get_k_only(s, dm, hermi=1, direct_scf_tol=1e-13):
        pd = s._pjs_data
        if pd is None or pd._itol != direct_scf_tol:
            s._build_pjs(direct_scf_tol)
        calls.append((s, s.mol.omega, direct_scf_tol, s._pjs_data))
        return dm * (1.0 + s.mol.omega)

    monkeypatch.setattr(sgx.sgx_jk, 'get_k_only', get_k_only)
    return obj, numpy.eye(mol.nao), calls


def test_unpatched_rsh_branch_builds_an_extra_j_and_drops_the_tolerance(rsh_setup):
    """Documents the released PySCF 2.14 behaviour the patch repairs."""
    obj, dm, calls = rsh_setup
    obj.get_jk(dm, 1, None, with_j=False, with_k=True, direct_scf_tol=1e-9, omega=0.3)
    assert obj.j_calls == 0 and obj._rsh_df['0.300000'].j_calls == 1     # discarded J
    assert calls[-1][2] == 1e-13                                          # tol lost
    obj.get_jk(dm, 1, None, with_j=False, with_k=True)                    # full range
    assert obj._pjs_data is obj._rsh_df['0.300000']._pjs_data or calls[-1][3] is not None


def test_patched_rsh_branch(rsh_setup):
    obj, dm, calls = rsh_setup
    hooks = sgx_patch.apply(sgx.SyntheticSGX)
    try:
        assert sgx.SyntheticSGX.get_jk.rsh_fixed
        sgx_patch.apply(sgx.SyntheticSGX).restore_all()        # idempotent: no second layer
        assert sgx.SyntheticSGX.get_jk.rsh_fixed
        kf = obj.get_jk(dm, 1, None, with_j=False, with_k=True)[1]              # full K first
        vj, kl = obj.get_jk(dm, 1, None, with_j=False, with_k=True, direct_scf_tol=1e-9,
                            omega=0.3)
        twin = obj._rsh_df['%.6f' % 0.3]
        assert vj is None and twin.j_calls == 0 and obj.j_calls == 0           # no J built
        assert calls[-1][0] is twin and calls[-1][1] == 0.3                      # under omega
        assert calls[-1][2] == 1e-9 and twin._pjs_data._itol == 1e-9            # tol arrives
        assert twin._pjs_data is not obj._pjs_data                              # own bounds
        assert twin._rsh_df is None and twin._vjopt is None
        assert obj.mol.omega == 0.0                                             # restored
        assert numpy.allclose(kf, dm) and numpy.allclose(kl, 1.3 * dm)
        # the copy is reused, and with_j=True still builds J for the caller
        vj2, _ = obj.get_jk(dm, 1, None, with_j=True, with_k=True, omega=0.3)
        assert obj._rsh_df['%.6f' % 0.3] is twin and numpy.allclose(vj2, 2 * dm)
    finally:
        hooks.restore_all()
    assert not getattr(sgx.SyntheticSGX.get_jk, 'rsh_fixed', False)


def test_install_gpu_applies_and_removes_the_rsh_fix():
    mf = _mf()
    hooks = install.install_gpu(mf, backend='hip', rsh_fix=True, sgx_class=sgx.SyntheticSGX)
    assert sgx.SyntheticSGX.get_jk.rsh_fixed
    hooks.restore_all()
    assert not getattr(sgx.SyntheticSGX.get_jk, 'rsh_fixed', False)


# ------------------------------------------------------ overlapped get_veff ----
class _RecordingNumInt(sgx.SyntheticNumInt):
    """nr_rks / nr_nlc_vxc with fixed results; records the thread."""

    def __init__(self):
        self.threads = []

    def nr_rks(self, mol, grids, xc_code, dms, relativity=0, hermi=1, max_memory=2000,
               verbose=None):
        self.threads.append(('rks', threading.current_thread().name))
        return 10.0, -3.25, 0.01 * numpy.arange(dms.size).reshape(dms.shape)

    def nr_nlc_vxc(self, mol, grids, xc_code, dm, relativity=0, hermi=1, max_memory=2000,
                   verbose=None):
        self.threads.append(('nlc', threading.current_thread().name))
        return 10.0, 0.125, 0.5 * numpy.eye(dm.shape[0])


def _mf():
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = types.SimpleNamespace(coords=numpy.zeros((64, 3)), weights=numpy.ones(64),
                              build=lambda **kw: None)
    mf = sgx.SyntheticRKS(mol, g, g)
    mf._numint = _RecordingNumInt()
    mf.with_df.j_builder = lambda s, dm: 2.0 * dm
    return mf


def test_overlapped_get_veff_equals_serial(monkeypatch):
    if install.gpu_available():
        pytest.skip('a GPU is present: the hooked flows would run on it')
    monkeypatch.setattr(sgx.sgx_jk, 'get_k_only',
                        lambda s, dm, hermi=1, direct_scf_tol=1e-13: dm * (1 + s.mol.omega))
    mf = _mf()
    rng = numpy.random.default_rng(0)
    a = rng.normal(size=(mf.mol.nao, mf.mol.nao))
    dm = a + a.T
    serial = mf.get_veff(mf.mol, dm)
    hooks = install.install_gpu(mf, backend='hip', overlap=True, rsh_fix=True,
                                sgx_class=sgx.SyntheticSGX)
    try:
        assert 'get_veff' in vars(mf)
        mf._numint.threads.clear()
        with warnings.catch_warnings():
            warnings.simplefilter('ignore', RuntimeWarning)    # no GPU: the flows fall back
            ov = mf.get_veff(mf.mol, dm)
        assert numpy.allclose(numpy.asarray(ov), numpy.asarray(serial), rtol=0, atol=1e-13)
        assert ov.exc == pytest.approx(serial.exc, abs=1e-12)
        assert ov.ecoul == pytest.approx(serial.ecoul, abs=1e-12)
        assert numpy.array_equal(ov.vk, serial.vk) and numpy.array_equal(ov.vj, serial.vj)
        assert sorted(k for k, _ in mf._numint.threads) == ['nlc', 'rks']
        assert all(t == 'wb97mv-amd-xc' for _, t in mf._numint.threads)   # on the worker
        assert len(hooks.timings) == 1 and set(hooks.timings[0]) == {'jk', 'xc', 'wait'}
    finally:
        hooks.restore_all()
    assert 'get_veff' not in vars(mf) and 'nr_rks' not in vars(mf._numint)


def test_overlapped_get_veff_reraises_worker_errors(monkeypatch):
    if install.gpu_available():
        pytest.skip('a GPU is present')
    monkeypatch.setattr(sgx.sgx_jk, 'get_k_only',
                        lambda s, dm, hermi=1, direct_scf_tol=1e-13: dm)
    mf = _mf()
    hooks = install.install_gpu(mf, backend='hip', overlap=True, semilocal=False)
    try:
        def bad(*a, **k):
            raise ZeroDivisionError('worker')
        mf._numint.nr_nlc_vxc = bad           # the worker's VV10 term raises
        with warnings.catch_warnings():
            warnings.simplefilter('ignore', RuntimeWarning)
            with pytest.raises(ZeroDivisionError):
                mf.get_veff(mf.mol, numpy.eye(mf.mol.nao))
    finally:
        hooks.restore_all()


# ------------------------------------------------------------ out of scope ----
def test_out_of_scope_basis_raises_unsupported():
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    mol.cart = True
    with pytest.raises(Unsupported, match='spherical'):
        pack_mol(mol)
    g_basis = {'H': [(0, [(1.0, 1.0)]), (4, [(0.8, 1.0)])]}
    with pytest.raises(Unsupported, match='l = 4'):
        pack_mol(mole.Mole([('H', (0, 0, 0)), ('H', (0, 0, 0.74))], g_basis))


@pytest.mark.parametrize('case', ['hermi0', 'two_dms', 'nonsym', 'omega_negative',
                                  'wrong_nao'])
def test_out_of_scope_k_input_raises_before_device(case):
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = types.SimpleNamespace(coords=numpy.zeros((64, 3)), weights=numpy.ones(64))
    s = sgx.SyntheticSGX(mol, g)
    # no device context: any device call would raise something else than
    # Unsupported, so the checks provably run first
    kb = sgx_k.GpuKBuilder(mol, ctx=object())
    n = mol.nao
    dm, hermi = numpy.eye(n), 1
    if case == 'hermi0':
        hermi = 0
    elif case == 'two_dms':
        dm = numpy.stack([dm, dm])
    elif case == 'nonsym':
        dm = dm + numpy.triu(numpy.ones((n, n)), 1)
    elif case == 'wrong_nao':
        kb._basis = types.SimpleNamespace(nao=n + 1)
    if case == 'omega_negative':
        with mol.with_range_coulomb(-0.3), pytest.raises(Unsupported, match='short-range'):
            sgx_k.get_k_only_gpu(kb, s, dm, hermi)
    else:
        with pytest.raises(Unsupported):
            sgx_k.get_k_only_gpu(kb, s, dm, hermi)


def test_out_of_scope_k_falls_back_with_the_warning_prefix(monkeypatch):
    """hermi = 0 through the installed hook: Unsupported, a RuntimeWarning
    with the prefix, and the original's result (tests/gpu/test_sgx_k.py:
    247-263 of the reference)."""
    seen = []
    monkeypatch.setattr(sgx.sgx_jk, 'get_k_only',
                        lambda s, dm, hermi=1, direct_scf_tol=1e-13: seen.append(hermi) or dm)
    mf = _mf()
    hooks = install.install_gpu(mf, backend='hip', k=True, sgx_module=sgx.sgx_jk,
                                overlap=False)
    try:
        with warnings.catch_warnings(record=True) as w:
            warnings.simplefilter('always')
            out = sgx.sgx_jk.get_k_only(mf.with_df, numpy.eye(mf.mol.nao), 0)
        assert seen == [0] and numpy.array_equal(out, numpy.eye(mf.mol.nao))
        msgs = [str(x.message) for x in w if issubclass(x.category, RuntimeWarning)]
        assert any(m.startswith(install.WARN_PREFIX + ' get_k_only failed (Unsupported')
                   for m in msgs)
    finally:
        hooks.restore_all()
