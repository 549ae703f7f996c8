"""The generated McMurchie-Davidson code (csrc/gen_md.py -> md_gen.h),
compiled for the host with g++, against the FP64 oracle for every class
(la, lb) <= (f, f) and both operators.  This checks the generator's Hermite
recursion, its factorized E contraction and the cart<->sph transforms the
device kernel uses, with FP64 arithmetic (no GPU needed)."""
import ctypes
import os
import subprocess
import sys

import numpy
import pytest

import oracle
from independent import boys_closed
from pyscf_wb97mv_fast_amd.systems import mole

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, 'pyscf-wb97mv-fast_amd')


@pytest.fixture(scope='module')
def md_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp('mdhost')
    subprocess.check_call([sys.executable, os.path.join(PKG, 'csrc', 'gen_md.py'),
                           str(d / 'md_gen.h')])
    so = d / 'md_host.so'
    subprocess.check_call(['g++', '-O1', '-std=c++17', '-shared', '-fPIC', '-I', str(d),
                           os.path.join(ROOT, 'tests', 'host', 'md_host.cpp'), '-o', str(so)])
    lib = ctypes.CDLL(str(so))
    lib.md_block.restype = ctypes.c_int
    return lib


def _hermite(la, lb, p, xpa, xpb):
    E = numpy.zeros((la + 1, lb + 1, la + lb + 1))
    E[0, 0, 0] = 1.0
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            if j == 0:
                src, x = E[i - 1, 0], xpa
                tp = i - 1
            else:
                src, x = E[i, j - 1], xpb
                tp = i + j - 1
            for t in range(i + j + 1):
                v = 0.0
                if t > 0:
                    v += src[t - 1] / (2 * p)
                if t <= tp:
                    v += x * src[t]
                if t + 1 <= tp:
                    v += (t + 1) * src[t + 1]
                E[i, j, t] = v
    return E


@pytest.mark.parametrize('onec', [False, True])
@pytest.mark.parametrize('omega', [0.0, 0.3])
def test_generated_md_blocks(md_lib, omega, onec):
    """Two-centre variant on A != B; one-centre variant (which drops the E
    terms that vanish when A = B = P) on two shells of one centre.  Blocks
    that vanish by symmetry (one-centre, C = A) are compared on the scale of
    the pair's R0 values."""
    rng = numpy.random.default_rng(11)
    for la in range(4):
        for lb in range(la + 1):
            A = rng.normal(size=3)
            B = A.copy() if onec else A + rng.normal(size=3)
            ea, eb = rng.uniform(0.3, 4.0, 2), rng.uniform(0.3, 4.0, 2)
            ca = mole.normalize_contracted(la, ea, numpy.ones((2, 1)))[:, 0]
            cb = mole.normalize_contracted(lb, eb, numpy.ones((2, 1)))[:, 0]
            basis = oracle.OracleBasis([la, lb], [2, 2], numpy.array([A, B]),
                                       numpy.concatenate([ea, eb]), numpy.concatenate([ca, cb]))
            L = la + lb
            for C in (A + 0.2, (A + B) / 2, A + [3.0, -1.0, 2.0]):
                d, R0, E = [], [], []
                for i in range(2):
                    for j in range(2):
                        p = ea[i] + eb[j]
                        P = (ea[i] * A + eb[j] * B) / p
                        K = numpy.exp(-ea[i] * eb[j] / p * ((A - B) ** 2).sum())
                        pref = ca[i] * cb[j] * K * 2 * numpy.pi / p
                        th = 1.0 if omega == 0 else omega ** 2 / (omega ** 2 + p)
                        T = th * p * ((P - C) ** 2).sum()
                        F = [boys_closed(n, numpy.array([T]))[0] for n in range(L + 1)]
                        R0 += [pref * numpy.sqrt(th) * (-2 * th * p) ** n * F[n]
                               for n in range(L + 1)]
                        d += list(P - C)
                        for ax in range(3):
                            E += list(_hermite(la, lb, p, P[ax] - A[ax], P[ax] - B[ax]).ravel())
                d, R0, E = (numpy.ascontiguousarray(v, dtype=numpy.float64) for v in (d, R0, E))
                out = numpy.zeros((2 * la + 1) * (2 * lb + 1))
                rc = md_lib.md_block(la, lb, int(onec), 4, d.ctypes.data_as(ctypes.c_void_p),
                                     R0.ctypes.data_as(ctypes.c_void_p),
                                     E.ctypes.data_as(ctypes.c_void_p),
                                     out.ctypes.data_as(ctypes.c_void_p))
                assert rc == 0
                ref = oracle.esp_dense(basis, C[None], omega)[0][:2 * la + 1, 2 * la + 1:]
                got = out.reshape(2 * la + 1, 2 * lb + 1)
                assert numpy.allclose(got, ref, rtol=1e-11,
                                      atol=1e-13 * max(numpy.abs(ref).max(), numpy.abs(R0).max())), \
                    (la, lb, omega, numpy.abs(got - ref).max())
