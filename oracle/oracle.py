"""FP64 CPU oracle of the omega-B97M-V / COSX hot path (Python side).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's CPU-baseline leg, always as the checker.  The product package
never imports it.

* ``OracleBasis.from_mol(mol)`` packs a molecule's shells (PySCF accessor
  names: ``bas_angular``, ``bas_exp``, ``_libcint_ctr_coeff``, ...).
* ``boys``, ``esp_dense``, ``ao_eval``, ``k_build``, ``vv10_sum`` wrap the C
  restatement in wb_oracle.c (liboracle.so, built by oracle/Makefile).
* ``get_k_only`` adds the post-processing of PySCF's sgx_jk.get_k_only as the
  reference applies it (gpu/sgx_k.py:736-741): K = Q^T K if sym_ovlp, then
  (K + K^T)/2.
* ``nr_nlc_vxc`` restates dft.numint.nr_nlc_vxc / _vv10nlc as mirrored by
  the reference (gpu/vv10.py:243-342) in numpy FP64 on top of ``ao_eval`` and
  ``vv10_sum``.
"""
import ctypes
import os

import numpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

THRESH = 1e-8                     # gpu/vv10.py:40 (numint._vv10nlc)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, 'liboracle.so')
        if not os.path.exists(path):
            raise RuntimeError('oracle not built: run make -C oracle')
        _LIB = ctypes.CDLL(path)
    return _LIB


class _CBasis(ctypes.Structure):
    _fields_ = [('nshell', ctypes.c_int), ('nao', ctypes.c_int),
                ('l', ctypes.c_void_p), ('nprim', ctypes.c_void_p),
                ('prim0', ctypes.c_void_p), ('ao0', ctypes.c_void_p),
                ('center', ctypes.c_void_p), ('exps', ctypes.c_void_p),
                ('coefs', ctypes.c_void_p), ('c2s', ctypes.c_void_p * 4)]


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# real solid harmonics, l <= 3, rows m = -l..l (p: x, y, z), libcint cart
# order; written out independently of the package's table
def _c2s_tables():
    s, pi = numpy.sqrt, numpy.pi
    t0 = numpy.array([[0.5 / s(pi)]])
    t1 = s(3 / (4 * pi)) * numpy.eye(3)
    # cart: xx xy xz yy yz zz
    a, b, c = s(15 / (4 * pi)), s(5 / (16 * pi)), s(15 / (16 * pi))
    t2 = numpy.array([[0, a, 0, 0, 0, 0],
                      [0, 0, 0, 0, a, 0],
                      [-b, 0, 0, -b, 0, 2 * b],
                      [0, 0, a, 0, 0, 0],
                      [c, 0, 0, -c, 0, 0]])
    # cart: xxx xxy xxz xyy xyz xzz yyy yyz yzz zzz
    f3, f2, f1 = s(35 / (32 * pi)), s(105 / (4 * pi)), s(21 / (32 * pi))
    f0, g2 = s(7 / (16 * pi)), s(105 / (16 * pi))
    t3 = numpy.zeros((7, 10))
    t3[0, 1], t3[0, 6] = 3 * f3, -f3
    t3[1, 4] = f2
    t3[2, 8], t3[2, 1], t3[2, 6] = 4 * f1, -f1, -f1
    t3[3, 9], t3[3, 2], t3[3, 7] = 2 * f0, -3 * f0, -3 * f0
    t3[4, 5], t3[4, 0], t3[4, 3] = 4 * f1, -f1, -f1
    t3[5, 2], t3[5, 7] = g2, -g2
    t3[6, 0], t3[6, 3] = f3, -3 * f3
    return [numpy.ascontiguousarray(t) for t in (t0, t1, t2, t3)]


C2S = _c2s_tables()


class OracleBasis:
    """Segmented basis; a generally contracted shell becomes one shell per
    contraction column (same primitives), which is the same function set."""

    def __init__(self, l, nprim, center, exps, coefs):
        self.l = numpy.ascontiguousarray(l, dtype=numpy.int32)
        self.nprim = numpy.ascontiguousarray(nprim, dtype=numpy.int32)
        self.prim0 = numpy.zeros(len(l), dtype=numpy.int32)
        self.prim0[1:] = numpy.cumsum(self.nprim)[:-1]
        dims = 2 * self.l + 1
        self.ao0 = numpy.zeros(len(l), dtype=numpy.int32)
        self.ao0[1:] = numpy.cumsum(dims)[:-1]
        self.nao = int(dims.sum())
        self.center = numpy.ascontiguousarray(center, dtype=numpy.float64)
        self.exps = numpy.ascontiguousarray(exps, dtype=numpy.float64)
        self.coefs = numpy.ascontiguousarray(coefs, dtype=numpy.float64)
        c = _CBasis()
        c.nshell, c.nao = len(self.l), self.nao
        c.l, c.nprim, c.prim0, c.ao0 = (_ptr(self.l), _ptr(self.nprim),
                                        _ptr(self.prim0), _ptr(self.ao0))
        c.center, c.exps, c.coefs = _ptr(self.center), _ptr(self.exps), _ptr(self.coefs)
        for i in range(4):
            c.c2s[i] = C2S[i].ctypes.data
        self._c = c

    @classmethod
    def from_mol(cls, mol):
        l, nprim, cen, exps, coefs = [], [], [], [], []
        for ib in range(mol.nbas):
            e = numpy.asarray(mol.bas_exp(ib), dtype=numpy.float64)
            cc = numpy.asarray(mol._libcint_ctr_coeff(ib), dtype=numpy.float64)
            for k in range(cc.shape[1]):
                l.append(mol.bas_angular(ib))
                nprim.append(len(e))
                cen.append(mol.bas_coord(ib))
                exps.append(e)
                coefs.append(cc[:, k])
        return cls(l, nprim, numpy.asarray(cen), numpy.concatenate(exps),
                   numpy.concatenate(coefs))

    @property
    def cref(self):
        return ctypes.byref(self._c)


def boys(nmax, T):
    T = numpy.ascontiguousarray(T, dtype=numpy.float64).ravel()
    out = numpy.empty((T.size, nmax + 1))
    lib().orc_boys_batch(ctypes.c_int(nmax), ctypes.c_int(T.size), _ptr(T), _ptr(out))
    return out


def esp_dense(basis, coords, omega=0.0):
    coords = numpy.ascontiguousarray(coords, dtype=numpy.float64).reshape(-1, 3)
    out = numpy.empty((len(coords), basis.nao, basis.nao))
    lib().orc_esp_dense(basis.cref, ctypes.c_int(len(coords)), _ptr(coords),
                        ctypes.c_double(omega), _ptr(out))
    return out


def ao_eval(basis, coords, deriv=0):
    coords = numpy.ascontiguousarray(coords, dtype=numpy.float64).reshape(-1, 3)
    out = numpy.empty((1 + 3 * deriv, len(coords), basis.nao))
    lib().orc_ao_eval(basis.cref, ctypes.c_int(len(coords)), _ptr(coords),
                      ctypes.c_int(deriv), _ptr(out))
    return out


def k_build(basis, coords, weights, QD, omega=0.0, pair_cut=0.0):
    """K_raw = X^T (w o G) for the contracted dm QD (no post-processing)."""
    coords = numpy.ascontiguousarray(coords, dtype=numpy.float64).reshape(-1, 3)
    weights = numpy.ascontiguousarray(weights, dtype=numpy.float64)
    QD = numpy.ascontiguousarray(QD, dtype=numpy.float64)
    K = numpy.empty((basis.nao, basis.nao))
    lib().orc_k_build(basis.cref, ctypes.c_int(len(coords)), _ptr(coords),
                      _ptr(weights), _ptr(QD), ctypes.c_double(omega),
                      ctypes.c_double(pair_cut), _ptr(K))
    return K


def k_build_screened(basis, coords, weights, QD, omega=0.0, pair_cut=0.0, tol=1e-11):
    """k_build with per-point pair screening (the CPU baseline's cost model of
    an SGX build's screening; not a rigorous bound): returns (K, pair
    evaluations kept)."""
    coords = numpy.ascontiguousarray(coords, dtype=numpy.float64).reshape(-1, 3)
    weights = numpy.ascontiguousarray(weights, dtype=numpy.float64)
    QD = numpy.ascontiguousarray(QD, dtype=numpy.float64)
    K = numpy.empty((basis.nao, basis.nao))
    kept = ctypes.c_longlong()
    lib().orc_k_build_screened(basis.cref, ctypes.c_int(len(coords)), _ptr(coords),
                               _ptr(weights), _ptr(QD), ctypes.c_double(omega),
                               ctypes.c_double(pair_cut), ctypes.c_double(tol), _ptr(K),
                               ctypes.byref(kept))
    return K, int(kept.value)


def get_k_only(basis, coords, weights, dm, Q=None, sym_ovlp=False, omega=0.0,
               pair_cut=0.0):
    """sgx_jk.get_k_only's contract: QD = Q dm, K = X^T(w o A(X QD)),
    K = Q^T K if sym_ovlp, then (K + K^T)/2 (gpu/sgx_k.py:709-741)."""
    Q = numpy.eye(basis.nao) if Q is None else Q
    K = k_build(basis, coords, weights, Q @ dm, omega, pair_cut)
    if sym_ovlp:
        K = Q.T @ K
    return 0.5 * (K + K.T)


def vv10_sum(coords, q, W0, kappa):
    coords = numpy.ascontiguousarray(coords, dtype=numpy.float64).reshape(-1, 3)
    n = len(coords)
    args = [numpy.ascontiguousarray(a, dtype=numpy.float64) for a in (q, W0, kappa)]
    F, U, W = numpy.empty(n), numpy.empty(n), numpy.empty(n)
    lib().orc_vv10_sum(ctypes.c_int(n), _ptr(coords), *[_ptr(a) for a in args],
                       _ptr(F), _ptr(U), _ptr(W))
    return F, U, W


def vv10_sum_outer(coords, q, W0, kappa, outer):
    """F/U/W for the outer points ``outer`` (indices) against all points."""
    coords = numpy.ascontiguousarray(coords, dtype=numpy.float64).reshape(-1, 3)
    n = len(coords)
    outer = numpy.ascontiguousarray(outer, dtype=numpy.int32)
    args = [numpy.ascontiguousarray(a, dtype=numpy.float64) for a in (q, W0, kappa)]
    m = len(outer)
    F, U, W = numpy.empty(m), numpy.empty(m), numpy.empty(m)
    lib().orc_vv10_sum_outer(ctypes.c_int(n), _ptr(coords), *[_ptr(a) for a in args],
                             ctypes.c_int(m), _ptr(outer), _ptr(F), _ptr(U), _ptr(W))
    return F, U, W


def vv10_pointwise(rho0, grad, weights, nlc_pars):
    """numint._vv10nlc's per-point quantities (gpu/vv10.py:243-265)."""
    b, C = nlc_pars
    pi = numpy.pi
    pi43 = 4.0 * pi / 3.0
    kvv = b * 1.5 * pi * (9.0 * pi) ** (-1.0 / 6.0)
    beta = ((3.0 / (b * b)) ** 0.75) / 32.0
    G = (grad ** 2).sum(axis=0)
    w0tmp = C * (G / (rho0 * rho0)) ** 2
    W0 = numpy.sqrt(w0tmp + pi43 * rho0)
    dW0dR = (0.5 * pi43 * rho0 - 2.0 * w0tmp) / W0
    dW0dG = w0tmp * rho0 / (G * W0)
    kappa = kvv * rho0 ** (1.0 / 6.0)
    dKdR = kappa / 6.0
    return rho0 * weights, W0, kappa, dW0dR, dW0dG, dKdR, beta


def nr_nlc_vxc(basis, coords, weights, dm, nlc_pars=(6.0, 0.01)):
    """(nelec, excsum, vmat) of the VV10 term for one RKS dm (FP64)."""
    ao = ao_eval(basis, coords, deriv=1)
    tmp = ao[0] @ dm
    rho = numpy.empty((4, len(coords)))
    rho[0] = numpy.einsum('pm,pm->p', tmp, ao[0])
    for d in range(3):
        rho[d + 1] = 2.0 * numpy.einsum('pm,pm->p', tmp, ao[d + 1])
    nelec = float((rho[0] * weights).sum())
    ind = rho[0] >= THRESH
    q, W0, kappa, dW0dR, dW0dG, dKdR, beta = vv10_pointwise(
        rho[0][ind], rho[1:4][:, ind], weights[ind], nlc_pars)
    F, U, W = vv10_sum(coords[ind], q, W0, kappa)
    exc = numpy.zeros(len(coords))
    v0 = numpy.zeros(len(coords))
    v1 = numpy.zeros(len(coords))
    exc[ind] = beta + 0.5 * F
    v0[ind] = beta + F + 1.5 * (U * dKdR + W * dW0dR)
    v1[ind] = 1.5 * W * dW0dG
    excsum = float((rho[0] * weights * exc).sum())
    wv = numpy.empty((4, len(coords)))
    wv[0] = 0.5 * v0 * weights
    wv[1:4] = 2.0 * v1 * rho[1:4] * weights
    aow = numpy.einsum('dpm,dp->pm', ao, wv)
    vmat = ao[0].T @ aow
    return nelec, excsum, vmat + vmat.T
