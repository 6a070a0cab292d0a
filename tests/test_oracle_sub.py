"""Pins of the subhalo oracle (oracle/sub.py) on the CPU.

The ingredient as a whole is pinned by the explorer's "+ subhalos" ladder row
(tests/test_oracle_ref2d.py).  Here its pieces are checked against what the
mathematics fixes, each chosen so that a plausible slip (a dropped term, a
wrong node, a swapped index) fails:

* the 3-node Gauss rule reproduces the standardised moments tau_0..tau_5 of
  the law it was built from; the 2-node fallback reproduces tau_0..tau_3;
  a known three-point law is recovered node for node;
* standardised moments from cumulants: a Poisson law (every cumulant lambda)
  against its raw moments by direct summation of the probability mass;
* the rare-mark cut against its definition by brute force;
* the cancellation-free determinant c2 sS2 - c11^2 = cross + sS2 W against
  exact rational arithmetic on a random clump table;
* the ring average (a midpoint rule in the angle) against adaptive
  quadrature, on a profile continuous at its outer edge;
* the clump-number profile: the fraction inside r200 by adaptive quadrature,
  the projected density normalised to one clump over the disc;
* on a real host (z_s = 1): the nine bulk points reproduce the mean of S and
  the bulk variance of K and S, and the shear terms are non-negative.
"""
from fractions import Fraction

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.stats import poisson

from oracle import ref2d, sub
from oracle.cosmo import Cosmology


def _raw_from_nodes(t, p, order):
    return np.array([np.sum(p * t ** j) for j in range(order + 1)])


def test_gauss3_reproduces_moments():
    rng = np.random.default_rng(1)
    for _ in range(20):
        # a 4-point law (not representable by 3 nodes), standardised
        x = rng.normal(size=4)
        w = rng.uniform(0.1, 1.0, 4)
        w /= w.sum()
        mu = np.sum(w * x)
        sd = np.sqrt(np.sum(w * (x - mu) ** 2))
        z = (x - mu) / sd
        tau = np.array([[np.sum(w * z ** j) for j in range(6)]])
        t, p = sub.gauss3(tau, np.array([-np.inf]))
        m = _raw_from_nodes(t[0], p[0], 5)
        assert np.allclose(m, tau[0], atol=1e-9), (m, tau[0])


def test_gauss3_recovers_three_point_law():
    z = np.array([-1.3, 0.2, 1.9])
    w = np.array([0.2, 0.5, 0.3])
    mu = np.sum(w * z)
    sd = np.sqrt(np.sum(w * (z - mu) ** 2))
    zs = (z - mu) / sd
    tau = np.array([[np.sum(w * zs ** j) for j in range(6)]])
    t, p = sub.gauss3(tau, np.array([-10.0]))
    assert np.allclose(t[0], zs, atol=1e-7) and np.allclose(p[0], w, atol=1e-7)


def test_gauss3_two_node_fallback():
    """A lower bound above the third node forces the 2-node rule: exact to tau_3."""
    g3 = 0.7
    tau = np.array([[1.0, 0.0, 1.0, g3, 3.5, 6.0]])
    t, p = sub.gauss3(tau, np.array([-0.72]))
    assert p[0, 2] == 0.0 and t[0, 1] == t[0, 2]
    m = _raw_from_nodes(t[0], p[0], 3)
    assert np.allclose(m, [1.0, 0.0, 1.0, g3], atol=1e-12)


def test_std_moments_poisson():
    lam = 2.7
    sd, tau = sub.std_moments(*(np.array([lam]) for _ in range(5)))
    n = np.arange(200)
    pm = poisson.pmf(n, lam)
    z = (n - lam) / np.sqrt(lam)
    want = [np.sum(pm * z ** j) for j in range(6)]
    assert abs(sd[0] - np.sqrt(lam)) < 1e-15
    assert np.allclose(tau[0], want, rtol=1e-10, atol=1e-12)


def test_cut_for_definition():
    rng = np.random.default_rng(2)
    for _ in range(20):
        k = rng.uniform(0, 1, 50)
        nu = rng.exponential(1e-3, 50)
        eps = 3e-3
        kstar = sub.cut_for(nu, k, eps)
        # marks strictly above the cut carry total rate <= eps, and adding the
        # mark at the cut would exceed it (or every mark is above: cut 0)
        above = nu[k > kstar].sum()
        assert above <= eps
        if kstar > 0:
            assert above + nu[k == kstar].sum() > eps


def test_det_ks_exact():
    """c2 sS2 - c11^2 (the (K, S) covariance determinant of a clump table with
    mass depending on the mass bin only) equals cross + sS2 W exactly."""
    rng = np.random.default_rng(3)
    nb, nd = 4, 6
    ring = rng.uniform(0, 1, (1, nd))
    d = np.repeat(np.arange(nd), nb)
    ib = np.tile(np.arange(nb), nd)
    pref = rng.uniform(0.1, 1, d.size)
    k = rng.uniform(0, 1, d.size)
    mbin = rng.uniform(0.5, 2, nb)
    F = Fraction
    nu = [F(ring[0, d[c]]) * F(pref[c]) for c in range(d.size)]
    c2 = sum(nu[c] * F(k[c]) ** 2 for c in range(d.size))
    sS2 = sum(nu[c] * F(mbin[ib[c]]) ** 2 for c in range(d.size))
    c11 = sum(nu[c] * F(k[c]) * F(mbin[ib[c]]) for c in range(d.size))
    det = float(c2 * sS2 - c11 * c11)
    # the oracle's form (as inside SubhaloModel.coarse)
    P = np.zeros((nd, nb))
    Pa = np.zeros((nd, nb))
    np.add.at(P, (d, ib), pref)
    np.add.at(Pa, (d, ib), pref * k)
    n_, a_ = ring @ P, ring @ Pa
    ab = a_ / n_
    W = np.sum(ring[:, d] * pref[None, :] * (k[None, :] - ab[:, ib]) ** 2, axis=1)
    X = mbin[None, None, :] * ab[:, :, None] - mbin[None, :, None] * ab[:, None, :]
    cross = 0.5 * np.einsum("ri,rj,rij->r", n_, n_, X ** 2)
    assert abs(cross[0] + float(sS2) * W[0] - det) < 1e-12 * abs(det)


def test_ring_against_quadrature():
    from oracle import chain
    xp = np.linspace(0.05, 2.0, 64)
    fp = np.exp(-xp) * (1 + np.sin(3 * xp)) * (2.0 - xp)    # continuous at the outer edge
    r = np.array([0.1, 0.7, 1.5])
    dd = np.array([0.05, 0.4, 1.1, 2.5])
    nt = 4096
    cth = np.cos((np.arange(nt) + 0.5) * np.pi / nt)
    ring = np.empty((r.size, dd.size))
    ecd = np.empty_like(ring)
    chain.lib().oracle_sub_ring(r.size, chain._p(r), dd.size, chain._p(dd), nt, chain._p(cth),
                                xp.size, chain._p(xp), chain._p(fp), chain._p(ring),
                                chain._p(ecd))
    f = lambda x: np.interp(x, xp, fp, right=0.0)
    for i, ri in enumerate(r):
        for j, dj in enumerate(dd):
            R = lambda th: np.sqrt(ri * ri + dj * dj - 2 * ri * dj * np.cos(th))
            pts = [th for th in np.linspace(0, np.pi, 400)]
            num = quad(lambda th: f(R(th)), 0, np.pi, limit=400, points=pts[1:-1:20])[0] / np.pi
            cos_num = quad(lambda th: f(R(th)) * np.cos(th), 0, np.pi, limit=400,
                           points=pts[1:-1:20])[0] / np.pi
            assert abs(ring[i, j] - num) < 1e-6 * max(num, 1e-2)
            if num > 1e-3:
                assert abs(ecd[i, j] - cos_num / num) < 1e-5


def test_clump_profile_normalisation():
    C, eta = 6.0, 1.25
    x0 = sub.X0_RVIR * eta
    g = lambda x: x / np.sqrt((x / x0) ** -sub.BIAS_EXP + 1.0) / (1 + C * x) ** 2
    want = quad(g, 0, 1.0)[0] / quad(g, 0, eta)[0]
    assert abs(sub.frac_inside(1.0, eta, C, x0) - want) < 1e-7
    Rh, n3 = sub.n3_proj(1.0, C, x0)
    # projected density of the 3-D profile rho B / X (per volume): compare one
    # radius with an adaptive line-of-sight integral
    R = Rh[40]
    um = np.sqrt(1 - R * R)
    p = lambda u: 1.0 / np.sqrt((np.hypot(R, u) / x0) ** -2.5 + 1.0) / (
        np.hypot(R, u) * (1 + C * np.hypot(R, u)) ** 2)
    want_n3 = 2 * quad(p, 0, um)[0]
    assert abs(n3[40] / want_n3 - 1) < 1e-4


@pytest.fixture(scope="module")
def host():
    cos = Cosmology()
    zs = 1.0
    shells = ref2d.z_shells(zs)
    zl = np.array([s[2] for s in shells])
    Ml = ref2d.engine_grids()[1]
    model = sub.SubhaloModel(cos, zs, zl, Ml)
    jz, jM = 40, 80
    h = model.host(jz, jM)
    r = [q for q in ref2d.rows(cos, zs, shells={jz}) if q["jM"] == jM][0]
    grid = ref2d.host_grid(r["ks"], r["xt"], r["xm"], r["Nbar"])
    x_sub = (r["xt"] * h["rs"] + h["D"].max()) / h["rs"] * 1.001
    xref = np.concatenate([grid["xc"], grid["xo"]])
    cm = model.coarse(h, xref[xref <= x_sub])
    return model, h, cm


def test_host_tables(host):
    model, h, cm = host
    assert h is not None and 0 < h["f"] < 0.95
    assert cm["nq"] in (9, 15)
    assert np.all(cm["wp_c"] >= 0) and np.all(cm["A2_c"] >= -1e-300)
    # points sum to the no-miss probability (bulk weights e^{-nu_rare} and
    # the rare points' (1 - e^{-nu})): total 1 per coarse radius
    assert np.allclose(cm["wp_c"].sum(axis=1), 1.0, atol=1e-10)


def test_bulk_points_reproduce_moments(host):
    """The nine bulk points at each coarse radius: mean K = c1, var K = c2
    (the 3-node K rule), mean S = sS1 ms, var S = sS2 ms^2 (regression on K
    plus the residual rule) wherever no node was clamped at S = 0."""
    model, h, cm = host
    C, ms = cm["C"], cm["ms"]
    w = cm["wp_c"][:, :9]
    pw = w / w.sum(axis=1, keepdims=True)
    K, S = cm["Kp_c"][:, :9], cm["Sp_c"][:, :9]
    mK = np.sum(pw * K, axis=1)
    vK = np.sum(pw * (K - mK[:, None]) ** 2, axis=1)
    mS = np.sum(pw * S, axis=1)
    vS = np.sum(pw * (S - mS[:, None]) ** 2, axis=1)
    ok = (C[:, 1] > 0) & np.all(S > 0, axis=1)
    assert ok.sum() > 64
    assert np.allclose(mK[ok], C[ok, 0], rtol=1e-9)
    assert np.allclose(vK[ok], C[ok, 1], rtol=1e-7)
    assert np.allclose(mS[ok], C[ok, 5] * ms, rtol=1e-9)
    assert np.allclose(vS[ok], C[ok, 6] * ms ** 2, rtol=1e-6)
