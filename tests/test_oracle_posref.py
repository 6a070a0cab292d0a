"""Pins of the positive ('vp') closure on the reference chain (oracle/posref.py).

* vp rates with signed loadings: E_w lam = 1 and Var_w lam = v on the 48
  nodes, the rates of -phi the mirror image of those of +phi, and the
  interval form (rate_intervals) reproducing D = lam - 1.
* ln R (oracle_vp_lnr: group sums through the interval form, prefix sums
  over the states, log-scaled forward recursion) against a brute-force
  expectation over every path of the latent chain on a small random
  population: E_U exp(sum_s z_s(U_s)) with z_s(u_j) = sum_g D_gj Y_g and
  Y_g = sum_c W_c (e^{i k kappa_c} J0(q gamma_c) - 1), the path
  probabilities w_{j0} K_0[j0, j1] K_1[j1, j2] ...
* eta: the model variance at the calibrated eta equals the target, and the
  chain variance equals the brute-force variance over every path.
* The whole chain at z_s = 1 against the stored pdf_zs1_vp.json: P_S to 1e-9
  of the peak (the reference's gate, tests/test_reference.py:29-45), eta,
  the two clustering targets, the grid (with the closure's coarse grid
  405 x 101); z_s = 0.5 with SGL_SLOW_ORACLE=1.  The record's cell and group
  counts come from the code that wrote the fixtures and are counted there
  slightly differently (2,450,076 / 553,466 against 2,450,071 / 553,570
  here for the same PDF), so they are compared to 1e-3 only.
"""
import itertools
import json
import os

import numpy as np
import pytest
from scipy.special import j0

from oracle import closure, posref

HERE = os.path.dirname(os.path.abspath(__file__))
PDFS = np.load(os.path.join(HERE, "golden", "reference_pdfs.npz"))


def _fixture(name):
    """The closure record and moments of a stored vp PDF (tests/golden)."""
    gold = json.load(open(os.path.join(HERE, "golden", "reference_scalars.json")))
    return gold["fixtures_vp"][name]


def test_signed_rates_moments_and_intervals():
    u, w = closure.gh_nodes(48)
    v = np.array([0.0, 1e-6, 0.05, 0.3, 0.9, 1.7, 2.5])
    for sgn in (1.0, -1.0):
        lam = posref.vp_rates(v, np.full(v.size, sgn), u, w)
        assert np.all(lam >= 0)
        assert np.abs(lam @ w - 1).max() < 1e-13
        assert np.abs((lam - 1) ** 2 @ w - v).max() < 1e-12 * (1 + v).max()
        a, b, c, lo, hi = posref.rate_intervals((lam - 1).T, u)
        anchor = np.where(c == 0, 0.0, np.where(a >= 0, u[lo], u[hi]))
        act = (np.arange(48)[:, None] >= lo) & (np.arange(48)[:, None] <= hi)
        model = c[None, :] + np.where(act, (u[:, None] - anchor) * a + b, 0.0)
        assert np.abs(model - (lam - 1).T).max() < 1e-12
    lp = posref.vp_rates(v, np.ones(v.size), u, w)
    ln = posref.vp_rates(v, -np.ones(v.size), u, w)
    assert np.array_equal(ln, lp[:, ::-1])
    # a clipped rate (phi > 0) is zero up to a node and positive from there to the top
    act = lp[-1] > 0
    first = int(np.argmax(act))
    assert 0 < first and np.all(act[first:]) and not np.any(act[:first])


def _small_population(seed, n_sh=3):
    rng = np.random.default_rng(seed)
    cells = []
    for s in range(n_sh):
        n = 30 + 7 * s
        ph_vals = np.array([0.0, 0.4, -0.7, 1.3, 2.2]) * (1 + 0.1 * s)
        cells.append(dict(W=rng.uniform(0.1, 3.0, n), kap=rng.normal(0, 0.05, n),
                          gam=np.abs(rng.normal(0, 0.04, n)) + 1e-4,
                          phi=ph_vals[rng.integers(0, ph_vals.size, n)]))
    # share some (kappa, gamma) points between shells, as pooled centroids do
    cells[1]["kap"][:5], cells[1]["gam"][:5] = cells[0]["kap"][:5], cells[0]["gam"][:5]
    return dict(cells=cells, shells=list(range(n_sh)), rho=rng.uniform(-0.3, 0.9, n_sh - 1),
                fvar=rng.uniform(0.3, 1.0, n_sh), target_kappa=0.0, target_gamma=0.01)


def _brute_lnR(clo, pp, kv, qv):
    u, w = clo.u, clo.w
    ns = u.size
    out = np.empty((kv.size, qv.size), complex)
    Ds = []
    for s, cl in enumerate(pp["cells"]):
        ph, gidx, _ = clo.grp[s]
        Ds.append((posref.vp_rates(clo.eta * pp["fvar"][s] * ph ** 2, np.sign(ph), u, w) - 1.0,
                   gidx))
    for a, k in enumerate(kv):
        for b, q in enumerate(qv):
            zs = []
            for s, cl in enumerate(pp["cells"]):
                D, gidx = Ds[s]
                F = np.exp(1j * k * cl["kap"]) * j0(q * cl["gam"]) - 1.0
                Y = np.bincount(gidx, cl["W"] * F.real, minlength=D.shape[0]) + 1j * np.bincount(
                    gidx, cl["W"] * F.imag, minlength=D.shape[0])
                zs.append(D.T @ Y)
            # E over every path j0 -> j1 -> ... of the stationary chain
            E = 0.0
            for path in itertools.product(range(ns), repeat=len(zs)):
                p = w[path[0]]
                for s in range(1, len(zs)):
                    p *= clo.K[s - 1][path[s - 1], path[s]]
                E += p * np.exp(sum(z[j] for z, j in zip(zs, path)))
            out[a, b] = E
    return out


def test_lnR_matches_brute_force_paths():
    pp = _small_population(7)
    clo = posref.Closure(pp, nstate=6, eta=0.8)
    kv = np.array([0.0, 3.0, 40.0, 210.0])
    qv = np.array([0.0, 5.0, 77.0])
    got = clo.lnR(kv, qv)
    ref = _brute_lnR(clo, pp, kv, qv)
    np.testing.assert_allclose(np.exp(got), ref, rtol=1e-11, atol=1e-14)


def test_calibration_and_chain_variance():
    pp = _small_population(11)
    clo = posref.Closure(pp, nstate=6, eta=1.0)
    target = 0.7 * clo.variance(1.0)
    eta = clo.calibrate(target)
    assert abs(clo.variance(eta) / target - 1) < 1e-10
    # the chain variance of sum_s m_s(U_s) against every path
    u, w = clo.u, clo.w
    ms = [posref.vp_rates(eta * f * ph ** 2, np.sign(ph), u, w).T @ wk
          for (ph, _, wk), f in zip(clo.grp, pp["fvar"])]
    vals, probs = [], []
    for path in itertools.product(range(u.size), repeat=len(ms)):
        p = w[path[0]]
        for s in range(1, len(ms)):
            p *= clo.K[s - 1][path[s - 1], path[s]]
        probs.append(p)
        vals.append(sum(m[j] for m, j in zip(ms, path)))
    vals, probs = np.array(vals), np.array(probs)
    mean = probs @ vals
    assert abs(closure.chain_variance(ms, w, clo.K) / (probs @ (vals - mean) ** 2) - 1) < 1e-11


def _check_vp_fixture(r, key, zs):
    fx = _fixture("pdf_zs%g_vp" % zs)
    P0, T0 = PDFS[key + "_P_s"], PDFS[key + "_P_s_tail"]
    nb = P0.size
    d = np.max(np.abs(r.P[:nb] - P0)) / P0.max()
    dt = np.max(np.abs(r.P[nb:] - T0)) / P0.max()
    assert d <= 1e-9 and dt <= 1e-9, (d, dt)
    assert abs(r.closure.eta / fx["eta"] - 1) < 1e-9
    assert abs(r.pos["target_kappa"] / fx["target_kappa"] - 1) < 1e-8
    assert abs(r.pos["target_gamma"] / fx["target_gamma"] - 1) < 1e-8
    for k in ("nk_k", "nk_g", "n_kap", "n_gam", "xi_coarse", "xi_coarse_k", "window"):
        assert r.grid[k] == fx["grid"][k], k
    for k in ("mean", "var", "skew", "sigmaDL"):
        assert abs(getattr(r, k) / fx[k] - 1) < 1e-8, k
    assert abs(r.closure.n_cells / fx["cells"] - 1) < 1e-3
    assert abs(r.closure.n_groups / fx["groups"] - 1) < 1e-3
    assert len(r.pos["shells"]) == fx["shells"]


def test_vp_pdf_zs1_matches_fixture():
    _check_vp_fixture(posref.run(1.0), "zs1_vp", 1.0)


@pytest.mark.skipif(os.environ.get("SGL_SLOW_ORACLE") != "1",
                    reason="minutes; SGL_SLOW_ORACLE=1 runs it")
def test_vp_pdf_zs05_matches_fixture():
    _check_vp_fixture(posref.run(0.5), "zs0p5_vp", 0.5)
