"""The oracle's clustering="limber" term pinned without its own formulas.

* limber_lss_var (sgl_full.py:2541-2572) against the same variance in the
  order the reference's docstring derives it, int dl l / (2 pi) C_l with
  C_l = int dchi W^2 Dg^2 / f_K^2 P0(l / f_K): the l integral outside, a
  Gauss-Legendre chi quadrature inside, chi(z) from a fine cumulative
  integral of c / H, and the lensing efficiency from the critical surface
  density, W = rho_m0 (1 + z)^2 / Sigma_cr(z, z_s), instead of the
  3/2 Omega_m H0^2 form.  A dropped growth or (1 + z) factor, a wrong power
  of H0 or 2 pi, or swapped l limits fails it by far more than the 3e-5
  tolerance (the reference's own node counts put its discretisation near
  1e-5).
* lss_term (sgl_full.py:3104-3105) against the characteristic function of a
  Gaussian convergence of variance v with an isotropic Gaussian shear of the
  same total variance: E[e^{i k_k kappa}] E[J0(k_g |gamma|)] with |gamma|
  Rayleigh distributed, by quadrature.
"""
import os
import sys

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.interpolate import CubicSpline
from scipy.special import j0

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import ref2d as oref  # noqa: E402
from oracle.cosmo import CKMS, Cosmology  # noqa: E402


def ell_space_variance(cos, zs, n_chi=256, n_ell=4000):
    zf = np.linspace(0.0, zs, 40001)
    g = CKMS / (cos.H0 * cos.E(zf))
    chif = np.concatenate([[0.0], np.cumsum(0.5 * (g[1:] + g[:-1]) * np.diff(zf))])
    chis = chif[-1]
    z_of_chi = CubicSpline(chif, zf)
    x, w = np.polynomial.legendre.leggauss(n_chi)
    a, b = chis * 1e-4, chis * (1 - 1e-4)
    chi = a + 0.5 * (b - a) * (x + 1)
    wc = 0.5 * (b - a) * w
    z = z_of_chi(chi)
    W = np.array([cos.rhom * (1 + zz) ** 2 / cos.Sigma_cr(zz, zs) for zz in z])
    D = np.array([cos.Dg(zz) for zz in z])
    fk = np.asarray(cos.f_K(chi), float)
    ell = np.geomspace(1.0, 1e7, n_ell)
    k = ell[:, None] / fk[None, :]
    inside = (k >= 1e-7) & (k <= 1e5)
    pk = np.where(inside, cos.P0(np.clip(k, 1e-7, 1e5)), 0.0)
    cl = pk @ (wc * W ** 2 * D ** 2 / fk ** 2)
    return float(np.trapezoid(ell ** 2 * cl / (2 * np.pi), np.log(ell)))


@pytest.mark.parametrize("zs,kw", [(1.0, {}), (0.5, dict(growth_mode="legacy")),
                                   (2.0, dict(Ok=0.05, w0=-0.9))])
def test_limber_variance_matches_ell_space_integral(zs, kw):
    cos = Cosmology(**kw)
    v = oref.limber_lss_var(cos, zs)
    ref = ell_space_variance(cos, zs)
    print(f"z_s={zs} {kw}: limber_lss_var {v:.8e}, l-space {ref:.8e}, rel {v / ref - 1:.2e}")
    assert abs(v / ref - 1) < 3e-5


def test_lss_term_is_the_gaussian_characteristic_function():
    v = 3.7e-4
    kk = np.array([0.0, 12.0, 55.0])
    kg = np.array([0.0, 9.0, 40.0, 90.0])
    got = oref.lss_term(v, kk, kg)
    s2 = 0.5 * v                      # variance of each shear component
    for j, g in enumerate(kg):
        rayleigh = quad(lambda r: j0(g * r) * r / s2 * np.exp(-r * r / (2 * s2)), 0.0,
                        40.0 * np.sqrt(s2), limit=400, epsabs=1e-14)[0]
        for i, k in enumerate(kk):
            want = -0.5 * v * k * k + np.log(rayleigh)
            assert abs(got[i, j] - want) < 1e-9 * max(1.0, abs(want)), (k, g)


def test_engine_shells():
    """z_shells(zint="engine") (sgl_full.py:480-518): the engine nodes z_j
    below z_s with j >= 1 as the nodes, the backward differences as widths
    (so they telescope to z_last - z_0), two shells fewer than the midpoint
    rule, and the lensing weight sum dchi (1 + z)^2 / Sigma_cr a few per
    cent below the midpoint rule's, as the reference states ("~3.4% low in
    lensing weight": it drops [0, z_0] and [z_last, z_s])."""
    zl, _ = oref.engine_grids()
    cos = Cosmology()
    for zs in (0.5, 1.0, 3.0):
        eng, mid = oref.z_shells(zs, "engine"), oref.z_shells(zs)
        below = zl[zl < zs]
        assert [s[2] for s in eng] == list(below[1:])
        assert abs(sum(s[3] for s in eng) - (below[-1] - below[0])) < 1e-14
        assert len(eng) == len(mid) - 2

        def weight(sh):
            return sum(CKMS / cos.Hz(z) * dz * (1 + z) ** 2 / cos.Sigma_cr(z, zs)
                       for _, _, z, dz in sh)
        ratio = weight(eng) / weight(mid)
        print(f"z_s={zs}: engine / midpoint lensing weight {ratio:.4f}")
        assert 0.95 < ratio < 0.98


@pytest.mark.parametrize("n", [1, 2, 3])
def test_gauss_nodes_integrate_polynomials_per_shell(n):
    """z_shells(zint="gauss", zint_n=n) (sgl_full.py:503-518): the n nodes
    of every midpoint shell integrate z^p exactly for p < 2n, and n = 1 is
    the midpoint rule."""
    mid = oref.z_shells(1.0)
    g = oref.z_shells(1.0, f"gauss{n}")
    assert len(g) == n * len(mid)
    for s, (a, b, _, _) in enumerate(mid):
        nodes = g[n * s:n * (s + 1)]
        assert all(x[0] == a and x[1] == b and a < x[2] < b for x in nodes)
        for p in range(2 * n):
            got = sum(dz * z ** p for _, _, z, dz in nodes)
            assert abs(got - (b ** (p + 1) - a ** (p + 1)) / (p + 1)) < 1e-14 * max(1.0, b ** (p + 1))
    if n == 1:
        assert g == mid
