"""Device Boys function (FP32 and FP64) against the closed form
F_n(T) = Gamma(n+1/2) P(n+1/2, T) / (2 T^(n+1/2)) -- the checks of the
reference's tests/gpu/test_boys.py:21-44, with its gates: maximum relative
error 2e-7 for every order over [0, 50] and a geometric tail to 1e6, and
mean signed error below 1e-8 over 1e6 uniform points in [0, 50] (no bias).

The FP32 routine for highest order L evaluates only F_L by Taylor series
(F_0 and F_1 both for L = 1) and the lower orders by the downward
recursion, carried in FP64 from FP64 leading values; beyond T = 36 the
upward product F_n = F_{n-1} (n - 1/2) / T also runs in FP64.  Each order
is rounded to FP32 once, so the gate holds for every n.  A separate mean
bound (3e-9) covers the asymptotic branch alone."""
import numpy
import pytest

from independent import boys_closed

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def ctx():
    from pyscf_wb97mv_fast_amd import lib
    return lib.Context(0)


def _ts():
    return numpy.concatenate([[0.0, 1e-30, 1e-8], numpy.linspace(0, 50, 20001),
                              numpy.geomspace(50, 1e6, 2000)])


@pytest.mark.parametrize('nmax', [0, 1, 2, 3, 4, 5, 6])
def test_boys_fp32_relative_error(ctx, nmax):
    from pyscf_wb97mv_fast_amd.gpu.boys import boys_eval
    T = _ts().astype(numpy.float32).astype(numpy.float64)     # the FP32 inputs exactly
    F = boys_eval(ctx, T, nmax, 'mixed')
    assert F.shape == (nmax + 1, T.size)
    for n in range(nmax + 1):
        ref = boys_closed(n, T)
        rel = numpy.abs(F[n] - ref) / ref
        assert numpy.all(numpy.isfinite(F[n]))
        print('nmax %d n %d max rel %.2e' % (nmax, n, rel.max()))
        assert rel.max() < 2e-7, (nmax, n, rel.max(), T[rel.argmax()])


@pytest.mark.parametrize('nmax', [0, 1, 2, 6])
def test_boys_fp32_unbiased(ctx, nmax):
    from pyscf_wb97mv_fast_amd.gpu.boys import boys_eval
    T = numpy.random.default_rng(0).uniform(0, 50, 1_000_000).astype(numpy.float32)
    T = T.astype(numpy.float64)
    F = boys_eval(ctx, T, nmax, 'mixed')
    for n in range(nmax + 1):
        ref = boys_closed(n, T)
        assert abs(numpy.mean((F[n] - ref) / ref)) < 1e-8, (nmax, n)


@pytest.mark.parametrize('nmax', [0, 1, 3, 6])
def test_boys_fp32_asymptotic_unbiased(ctx, nmax):
    from pyscf_wb97mv_fast_amd.gpu.boys import boys_eval
    rng = numpy.random.default_rng(1)
    T = numpy.exp(rng.uniform(numpy.log(36.0), numpy.log(1e5), 400_000))
    T = T.astype(numpy.float32).astype(numpy.float64)
    F = boys_eval(ctx, T, nmax, 'mixed')
    for n in range(nmax + 1):
        ref = boys_closed(n, T)
        assert abs(numpy.mean((F[n] - ref) / ref)) < 3e-9, (nmax, n)


@pytest.mark.parametrize('nmax', [0, 3, 6])
def test_boys_fp64(ctx, nmax):
    from pyscf_wb97mv_fast_amd.gpu.boys import boys_eval
    T = _ts()
    F = boys_eval(ctx, T, nmax, 'fp64')
    for n in range(nmax + 1):
        ref = boys_closed(n, T)
        rel = numpy.abs(F[n] - ref) / ref
        assert rel.max() < 2e-13, (nmax, n, rel.max(), T[rel.argmax()])


def test_boys_nmax_range(ctx):
    from pyscf_wb97mv_fast_amd.gpu.boys import boys_eval
    with pytest.raises(ValueError):
        boys_eval(ctx, numpy.zeros(4), 7)
