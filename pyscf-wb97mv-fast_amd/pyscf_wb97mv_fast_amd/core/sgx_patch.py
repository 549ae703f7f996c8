"""Repair of the range-separated branch of PySCF 2.14's ``SGX.get_jk``.

What PySCF 2.14's ``pyscf.sgx.sgx.SGX.get_jk(dm, hermi, vhfopt, with_j,
with_k, direct_scf_tol, omega)`` does when an attenuated operator is asked
for (``omega`` given), and what goes wrong (the reference documents and
fixes the same two faults, core/sgx_patch.py:1-30, :33-62):

* It evaluates the operator on a per-omega copy of the SGX object kept in
  ``_rsh_df['%.6f' % omega]`` and calls that copy's ``get_jk`` with
  positional arguments that leave out ``vhfopt``.  ``with_j`` then receives
  ``with_k``'s value, ``with_k`` receives ``direct_scf_tol`` and the
  tolerance falls back to its default: each attenuated K also builds a J
  that is discarded, and the caller's ``direct_scf_tol`` never arrives.
* The copy is shallow, so it inherits ``_pjs_data`` (and the J optimizer and
  overlap correction) built for the full-range kernel.  Its K is then
  screened with the 1/r bounds instead of bounds rebuilt under
  ``mol.with_range_coulomb(omega)``.

``rsh_get_jk(original)`` returns a ``get_jk`` that passes every argument by
keyword and gives each per-omega copy caches of its own; with ``omega =
None`` it is the original.  ``apply()`` installs it on PySCF's class (or a
given class) and returns the ``HookSet`` that removes it again;
``install_gpu(..., rsh_fix=True)`` does this as part of its hooks, so the
drop-in needs nothing from the reference.
"""
import functools

from .hooks import HookSet

# the per-omega copy must not inherit these from the full-range object
_OWN_CACHES = ('_rsh_df', '_vjopt', '_overlap_correction_matrix', '_pjs_data')


def _default_tol():
    try:
        from pyscf import __config__
    except ImportError:
        return 1e-13
    return getattr(__config__, 'scf_hf_SCF_direct_scf_tol', 1e-13)


def range_copy(sgx, omega):
    """sgx's object for the operator erf(omega r)/r (erfc for omega < 0),
    made on first use under PySCF's key and with empty caches."""
    table = sgx._rsh_df
    if table is None:                 # an RSH copy itself: keep it local
        table = sgx._rsh_df = {}
    key = '%.6f' % omega
    twin = table.get(key)
    if twin is None:
        twin = sgx.copy()
        for attr in _OWN_CACHES:
            setattr(twin, attr, None)
        table[key] = twin
    return twin


def rsh_get_jk(original):
    """SGX.get_jk with the range-separated branch repaired (see the module
    docstring); ``original`` is PySCF's unpatched method."""
    @functools.wraps(original)
    def get_jk(self, dm, hermi=1, vhfopt=None, with_j=True, with_k=True,
               direct_scf_tol=None, omega=None):
        tol = _default_tol() if direct_scf_tol is None else direct_scf_tol
        if omega is None:
            return original(self, dm, hermi, vhfopt, with_j, with_k, tol)
        twin = range_copy(self, omega)
        with twin.mol.with_range_coulomb(omega):
            return original(twin, dm, hermi=hermi, vhfopt=vhfopt, with_j=with_j,
                            with_k=with_k, direct_scf_tol=tol)
    get_jk.rsh_fixed = True
    return get_jk


def apply(cls=None, hooks=None):
    """Install the repaired get_jk on ``cls`` (PySCF's SGX class by default)
    and return the HookSet that removes it (``hooks`` to add it to one).
    Installing on a class that already carries it changes nothing."""
    if cls is None:
        from pyscf.sgx import sgx as sgx_module
        cls = sgx_module.SGX
    hooks = HookSet() if hooks is None else hooks
    if not getattr(cls.get_jk, 'rsh_fixed', False):
        hooks.wrap(cls, 'get_jk', rsh_get_jk)
    return hooks
