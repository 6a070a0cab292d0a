"""Independent pins of oracle pieces that no reference vector covers.

* P1D (oracle/chain.py p1d, the 1-D power behind the slice covariance): the
  projection of the isotropic 3-D power onto the line of sight,
  (1/2 pi)^2 int d^2k_perp P0 W^2 at |k| = sqrt(k_par^2 + k_perp^2), by
  scipy's adaptive quadrature in k_perp (a different variable, integrand and
  rule from the oracle's running trapezoid in |k|).
* The Gaussian quadratic form (chain.quadform): the characteristic-function
  identity E exp(sum_i a_i delta_i) = exp(a^T C a / 2) for delta ~ N(0, C) and
  complex a, the expectation by a tensor Gauss-Hermite rule.
* The chain's composition on a tiny table: cells -> compensated Lambda
  (oracle_lnphi) -> phi with a Gaussian part -> the Filon inversion
  (oracle_invert), against the exact compound-Poisson density, the sum over
  every jump count n_c of prod_c Pois(n_c; w_c) N(x; sum n_c xi_c -
  sum w_c xi_c, s^2).
* The surrogate's FP64 MLP (oracle/mlp.py): backward against central
  finite differences of the forward.
"""
import itertools
import math

import numpy as np
from scipy import integrate, special, stats

from oracle import chain, mlp
from oracle.cosmo import Cosmology


def test_p1d_against_perpendicular_quadrature():
    cos = Cosmology()
    Rs = 20.0
    kq = np.array([2e-3, 1e-2, 0.05, 0.2])

    def W2(k):
        x = k * Rs
        w = 1 - x ** 2 / 10 * (1 - x ** 2 / 28) if x < 1e-2 else 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
        return w * w

    got = chain.p1d(cos, kq, Rs)
    for k, g in zip(kq, got):
        f = lambda kp: kp * float(cos.P0(np.array([math.hypot(k, kp)]))[0]) * W2(math.hypot(k, kp))
        # d^2k_perp / (2 pi)^2 = k_perp dk_perp / (2 pi); the window confines
        # the integrand below ~ 1 / Rs
        val = sum(integrate.quad(f, a, b, limit=400, epsabs=0, epsrel=1e-11)[0]
                  for a, b in ((0, 0.05), (0.05, 0.5), (0.5, 10.0 / Rs * 20)))
        assert abs(g / (val / (2 * np.pi)) - 1) < 1e-4, (k, g, val / (2 * np.pi))


def test_quadform_is_the_gaussian_characteristic_function():
    rng = np.random.default_rng(4)
    n = 3
    A = rng.normal(size=(n, n))
    C = A @ A.T / n + 0.2 * np.eye(n)
    a = (rng.normal(size=(n, 2)) + 1j * rng.normal(size=(n, 2))) * 0.6
    Lc = np.linalg.cholesky(C)
    x, w = np.polynomial.hermite_e.hermegauss(40)
    w = w / w.sum()
    Z = np.array(list(itertools.product(x, repeat=n)))          # standard normal nodes
    Wt = np.prod(np.array(list(itertools.product(w, repeat=n))), axis=1)
    delta = Z @ Lc.T
    E = (Wt[:, None] * np.exp(delta @ a)).sum(0)
    np.testing.assert_allclose(np.exp(chain.quadform(C, a)), E, rtol=1e-11)


def test_composition_against_exact_compound_poisson():
    xi = np.array([0.11, -0.04, 0.27])
    w = np.array([1.3, 0.7, 0.4])
    s = 0.05
    kq = np.concatenate([[0.0], np.linspace(1e-3, 160.0, 32000)])
    lam, _ = chain.lnphi(xi.reshape(1, 1, -1), w.reshape(1, 1, -1), np.ones((1, 1)), kq[1:])
    phi = np.concatenate([[1.0 + 0j], np.exp(lam - 0.5 * s * s * kq[1:] ** 2)])
    t = np.linspace(-0.3, 0.6, 181)
    p = chain.invert(t, kq, phi)
    # exact: every jump count up to where the Poisson tails are < 1e-16
    nmax = 22
    n = np.arange(nmax)
    P = [stats.poisson.pmf(n, wc) for wc in w]
    ref = np.zeros_like(t)
    shift = float(w @ xi)
    for i, j, k in itertools.product(n, n, n):
        pr = P[0][i] * P[1][j] * P[2][k]
        if pr < 1e-18:
            continue
        mu = i * xi[0] + j * xi[1] + k * xi[2] - shift
        ref += pr * np.exp(-0.5 * ((t - mu) / s) ** 2) / (s * math.sqrt(2 * math.pi))
    assert np.max(np.abs(p - ref)) < 1e-6 * ref.max()


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(9)
    sizes = [5, 7, 6, 4]
    params = [(rng.normal(size=(o, i)) / math.sqrt(i), rng.normal(size=o) * 0.1)
              for i, o in zip(sizes[:-1], sizes[1:])]
    x = rng.normal(size=(3, sizes[0]))
    dy = rng.normal(size=(3, sizes[-1]))
    y, acts, pre = mlp.forward(x, params)
    grads = mlp.backward(dy, params, acts, pre)

    def loss(ps):
        return float(np.sum(mlp.forward(x, ps)[0] * dy))

    h = 1e-6
    for li, (W, b) in enumerate(params):
        for idx in [(0, 0), (1, 2), (W.shape[0] - 1, W.shape[1] - 1)]:
            Wp, Wm = W.copy(), W.copy()
            Wp[idx] += h
            Wm[idx] -= h
            pp = [(Wp, b) if j == li else q for j, q in enumerate(params)]
            pm = [(Wm, b) if j == li else q for j, q in enumerate(params)]
            fd = (loss(pp) - loss(pm)) / (2 * h)
            assert abs(grads[li][0][idx] - fd) < 1e-7 * max(1.0, abs(fd)), (li, idx)
        for j in (0, b.size - 1):
            bp, bm = b.copy(), b.copy()
            bp[j] += h
            bm[j] -= h
            pp = [(W, bp) if q_ == li else q for q_, q in enumerate(params)]
            pm = [(W, bm) if q_ == li else q for q_, q in enumerate(params)]
            fd = (loss(pp) - loss(pm)) / (2 * h)
            assert abs(grads[li][1][j] - fd) < 1e-7 * max(1.0, abs(fd)), (li, j)
    # the SiLU derivative itself
    z = np.linspace(-8, 8, 33)
    np.testing.assert_allclose(mlp.dsilu(z), (mlp.silu(z + 1e-6) - mlp.silu(z - 1e-6)) / 2e-6,
                               rtol=1e-6, atol=1e-9)
