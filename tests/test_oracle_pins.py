"""Pins of the oracle functions that restate PySCF formulas (VV10 pointwise
terms and assembly, get_k_only post-processing) against what the
mathematics fixes, so that a dropped term, a wrong sign or factor or a
transposed operand fails here rather than being carried into both sides of
a parity test:

* VV10 uniform electron gas: beta is defined so that the non-local energy
  per electron of a uniform density vanishes, beta + (1/2) int n Phi(R) d^3R
  = 0 for every n (Vydrov and Van Voorhis, J. Chem. Phys. 133, 244103
  (2010)).  oracle.vv10_pointwise's beta, W0 and kappa and the kernel
  convention of oracle.vv10_sum_outer (F = -3/2 sum T) must satisfy it.
* vv10_pointwise's derivative terms against central differences of its own
  W0 and kappa (dW0dR = rho dW0/drho, dW0dG = rho dW0/dG, dKdR = rho dK/drho).
* oracle.nr_nlc_vxc: vmat is the derivative of excsum with respect to the
  density matrix, d E_nl(D + e S)/de = tr(vmat S), central differences.
* oracle.get_k_only, end to end: for an s-only basis the exact exchange
  K_mn = sum_ls (ml|ns) D_ls has the closed form of the (ss|ss) integral
  (with erf(w r)/r: the factor sqrt(w^2/(w^2+a)) and the attenuated Boys
  argument).  The seminumerical K converges to it as the grid grows; with
  PJS's overlap-fitted projector Q = S_grid^-1 S (COSX) it is closer still,
  and with the transposed Q it is not.
"""
import numpy
import pytest
from scipy.special import erf

import oracle
from pyscf_wb97mv_fast_amd.systems import geometry, grids, guess, mole

NLC = (6.0, 0.01)


def uniform_gas_exc(n, nr=400, b=NLC[0], C=NLC[1]):
    """beta + F/2 at a point of a uniform density n: the inner points sit
    on a line through it (the kernel depends on distances only), at +R and
    -R with half the weights of int 4 pi R^2 dR each (so the points'
    centroid, the FP32 frame's origin on the device, is the outer point),
    Gauss-Legendre in t = R / (R + s), s = sqrt(kappa/W0)."""
    q, W0, kappa, _, _, _, beta = oracle.vv10_pointwise(
        numpy.array([n]), numpy.full((3, 1), 1e-20), numpy.array([1.0]), (b, C))
    s = numpy.sqrt(kappa[0] / W0[0])
    x, w = numpy.polynomial.legendre.leggauss(nr)
    t, wt = 0.5 * (x + 1), 0.5 * w
    R = s * t / (1 - t)
    wr = 4 * numpy.pi * R ** 2 * s / (1 - t) ** 2 * wt
    coords = numpy.zeros((2 * nr + 1, 3))
    coords[1:nr + 1, 2] = R
    coords[nr + 1:, 2] = -R
    qq = numpy.concatenate([[0.0], 0.5 * n * wr, 0.5 * n * wr])   # the outer point adds nothing
    return beta, coords, qq, numpy.full(2 * nr + 1, W0[0]), numpy.full(2 * nr + 1, kappa[0])


@pytest.mark.parametrize('n', [1e-5, 1e-2, 0.3, 20.0])
def test_vv10_beta_cancels_the_uniform_gas(n):
    beta, coords, q, W0, kappa = uniform_gas_exc(n)
    F = oracle.vv10_sum_outer(coords, q, W0, kappa, numpy.array([0]))[0][0]
    assert abs(beta + 0.5 * F) < 1e-12 * beta, (beta, F)
    # a wrong kernel factor would not cancel: the test discriminates
    assert abs(beta + 0.25 * F) > 1e-3 * beta


def test_vv10_beta_closed_form():
    """beta = (3 / b^2)^(3/4) / 32 is what the uniform-gas integral gives:
    beta = 3 pi^2 / (32 (kappa W0 / n^(2/3))^(3/2)) with W0 = sqrt(4 pi n / 3)
    and kappa = b (3 pi / 2) (9 pi)^(-1/6) n^(1/6)."""
    for b in (5.9, 6.0, 6.3):
        n = 0.7
        _, W0, kappa, _, _, _, beta = oracle.vv10_pointwise(
            numpy.array([n]), numpy.full((3, 1), 1e-20), numpy.array([1.0]), (b, 0.01))
        ref = 3 * numpy.pi ** 2 / (32 * (kappa[0] * W0[0] / n ** (2 / 3)) ** 1.5)
        assert beta == pytest.approx(ref, rel=1e-14)
        assert W0[0] == pytest.approx(numpy.sqrt(4 * numpy.pi * n / 3), rel=1e-15)


def test_vv10_pointwise_derivatives_by_finite_differences():
    rng = numpy.random.default_rng(0)
    rho = rng.uniform(1e-3, 2.0, 50)
    grad = rng.normal(size=(3, 50)) * rho
    G = (grad ** 2).sum(axis=0)
    _, W0, kappa, dW0dR, dW0dG, dKdR, _ = oracle.vv10_pointwise(rho, grad, numpy.ones(50), NLC)

    def W0_of(r, g):
        gs = numpy.sqrt(g / G)[None, :] * grad
        return oracle.vv10_pointwise(r, gs, numpy.ones(50), NLC)[1]

    kap = lambda r, g: oracle.vv10_pointwise(r, grad, numpy.ones(50), NLC)[2]   # noqa: E731

    def rich(f, var):
        # central differences in the relative step h, Richardson-extrapolated
        def d(h):
            if var == 'r':
                return (f(rho * (1 + h), G) - f(rho * (1 - h), G)) / (2 * h)
            return (f(rho, G * (1 + h)) - f(rho, G * (1 - h))) / (2 * h)
        return (4 * d(5e-5) - d(1e-4)) / 3

    fd_r = rich(W0_of, 'r')                      # rho dW0/drho
    fd_g = rich(W0_of, 'g') * rho / G            # rho dW0/dG
    fd_k = rich(kap, 'r')                        # rho dK/drho
    tol = 1e-10 * numpy.abs(W0).max()            # round-off of the differences
    assert numpy.allclose(dW0dR, fd_r, rtol=1e-8, atol=tol)
    assert numpy.allclose(dW0dG, fd_g, rtol=1e-8, atol=tol * (rho / G).max())
    assert numpy.allclose(dKdR, fd_k, rtol=1e-8, atol=1e-10 * numpy.abs(kappa).max())


def test_nr_nlc_vxc_is_the_derivative_of_its_energy():
    mol = mole.Mole(geometry.WATER_DIMER, 'def2-svp')
    g = grids.Grids(mol, level=0).build()
    rng = numpy.random.default_rng(0)
    sel = numpy.sort(rng.choice(g.size, 2500, replace=False))
    c, w = g.coords[sel], g.weights[sel]
    ob = oracle.OracleBasis.from_mol(mol)
    D = guess.huckel_density(mol)
    _, _, vmat = oracle.nr_nlc_vxc(ob, c, w, D)
    assert numpy.array_equal(vmat, vmat.T)
    for seed in (1, 2):
        S = numpy.random.default_rng(seed).normal(size=D.shape)
        S = 0.005 * (S + S.T)
        h = 1e-3
        ep = oracle.nr_nlc_vxc(ob, c, w, D + h * S)[1]
        em = oracle.nr_nlc_vxc(ob, c, w, D - h * S)[1]
        fd = (ep - em) / (2 * h)
        an = numpy.einsum('ij,ij->', vmat, S)
        assert abs(fd - an) < 1e-7 * abs(an), (fd, an)


# ------------------------------------------------ exchange, closed form ----
# s-only basis: STO-3G-like contracted s shells plus one diffuse s per atom
S_BASIS = {
    'H': [(0, [(3.42525091, 0.15432897), (0.62391373, 0.53532814), (0.16885540, 0.44463454)]),
          (0, [(0.12, 1.0)])],
    'O': [(0, [(6.36242139, 0.15432897), (1.15892300, 0.53532814), (0.31364979, 0.44463454)]),
          (0, [(0.5, 1.0)])],
}
S_ATOMS = [('H', (0.0, 0.0, 0.0)), ('H', (0.0, 0.0, 0.74)), ('O', (1.3, 0.4, 0.2))]


def s_only_system():
    return mole.Mole(S_ATOMS, S_BASIS)


def _boys0(T):
    T = numpy.asarray(T, dtype=float)
    Ts = numpy.maximum(T, 1e-300)
    return numpy.where(T < 1e-12, 1.0 - T / 3.0, 0.5 * numpy.sqrt(numpy.pi / Ts) * erf(numpy.sqrt(Ts)))


def ss_eri(mol, omega):
    """(ab|cd) of the s-only AOs in closed form: for primitives of exponents
    a, b, c, d (p = a + b, q = c + d, alpha = pq / (p + q)),
    2 pi^(5/2) / (p q sqrt(p + q)) K_ab K_cd sqrt(th) F0(th alpha |PQ|^2),
    th = w^2 / (w^2 + alpha) for erf(w r)/r, 1 for 1/r; the AO is the
    libcint-normalized radial function times Y00 = 1 / (2 sqrt(pi))."""
    y00 = 0.5 / numpy.sqrt(numpy.pi)
    prims = []
    for ib in range(mol.nbas):
        assert mol.bas_angular(ib) == 0
        cc = mol._libcint_ctr_coeff(ib)[:, 0] * y00
        prims.append((cc, mol.bas_exp(ib), numpy.asarray(mol.bas_coord(ib))))
    n = len(prims)
    # primitive-pair tables per AO pair
    pairs = {}
    for i in range(n):
        for j in range(n):
            ci, ei, A = prims[i]
            cj, ej, B = prims[j]
            p = ei[:, None] + ej[None, :]
            P = (ei[:, None, None] * A + ej[None, :, None] * B) / p[..., None]
            K = numpy.exp(-ei[:, None] * ej[None, :] / p * ((A - B) ** 2).sum())
            pairs[i, j] = ((ci[:, None] * cj[None, :] * K).ravel(), p.ravel(), P.reshape(-1, 3))
    out = numpy.zeros((n,) * 4)
    for (i, j), (cab, p, P) in pairs.items():
        for (k, l), (ccd, q, Q) in pairs.items():
            al = p[:, None] * q[None, :] / (p[:, None] + q[None, :])
            th = 1.0 if omega == 0 else omega ** 2 / (omega ** 2 + al)
            R2 = ((P[:, None, :] - Q[None, :, :]) ** 2).sum(-1)
            v = (2 * numpy.pi ** 2.5 / (p[:, None] * q[None, :] * numpy.sqrt(p[:, None] + q[None, :]))
                 * numpy.sqrt(th) * _boys0(th * al * R2))
            out[i, j, k, l] = cab @ v @ ccd
    return out


def test_ss_eri_closed_form_is_symmetric_and_attenuated():
    mol = s_only_system()
    E = ss_eri(mol, 0.0)
    assert numpy.allclose(E, E.transpose(1, 0, 2, 3)) and numpy.allclose(E, E.transpose(2, 3, 0, 1))
    El = ss_eri(mol, 0.3)
    assert numpy.all(El < E) and numpy.all(El > 0)                  # erf(w r) < 1
    assert numpy.allclose(ss_eri(mol, 1e5), E, rtol=1e-8)            # erf -> 1/r
    # the (aa|aa) self-repulsion of a single normalized s Gaussian: 2 sqrt(a/pi)
    one = mole.Mole([('H', (0, 0, 0))], {'H': [(0, [(0.8, 1.0)])]})
    assert ss_eri(one, 0.0)[0, 0, 0, 0] == pytest.approx(2 * numpy.sqrt(0.8 / numpy.pi), rel=1e-13)


@pytest.fixture(scope='module')
def s_system():
    mol = s_only_system()
    rng = numpy.random.default_rng(1)
    D = rng.normal(size=(mol.nao, mol.nao))
    D = 0.5 * (D + D.T)
    exact = {w: numpy.einsum('mlns,ls->mn', ss_eri(mol, w), D) for w in (0.0, 0.3)}
    gr = {lev: grids.Grids(mol, level=lev).build() for lev in (2, 4)}
    return mol, D, exact, gr


@pytest.mark.parametrize('omega', [0.0, 0.3])
def test_get_k_only_converges_to_the_analytic_exchange(s_system, omega):
    mol, D, exact, gr = s_system
    ob = oracle.OracleBasis.from_mol(mol)
    S = guess.overlap_matrix(mol)
    ref = exact[omega]
    scale = numpy.abs(ref).max()
    err = {}
    for lev, g in gr.items():
        X = oracle.ao_eval(ob, g.coords)[0]
        Q = numpy.linalg.solve((X * g.weights[:, None]).T @ X, S)   # S_grid^-1 S
        for name, kw in (('I', {}), ('Q', {'Q': Q}), ('Q^T', {'Q': Q.T}),
                         ('Q sym', {'Q': Q, 'sym_ovlp': True}),
                         ('Q^T sym', {'Q': Q.T, 'sym_ovlp': True})):
            K = oracle.get_k_only(ob, g.coords, g.weights, D, omega=omega, **kw)
            assert numpy.array_equal(K, K.T)
            err[lev, name] = numpy.abs(K - ref).max() / scale
    print(omega, {k: '%.1e' % v for k, v in err.items()})
    assert err[4, 'I'] < 4e-7 and err[4, 'I'] < 0.1 * err[2, 'I']      # converges
    assert err[4, 'Q'] < 1e-7
    assert err[2, 'Q'] < err[2, 'I'] and err[4, 'Q'] < err[4, 'I']      # overlap fitting helps
    for lev in (2, 4):
        assert err[lev, 'Q^T'] > 3 * err[lev, 'Q']                      # transposed Q does not
        assert err[lev, 'Q^T sym'] > 3 * err[lev, 'Q sym']
