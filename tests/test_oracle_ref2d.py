"""Pins of the 2-D (k_kappa, k_gamma) oracle (oracle/ref2d.py) on the CPU.

* The explorer notebook's ingredient ladder at z_s = 1 (Planck): halos,
  + ellipticity, + filaments, + subhalos; the printed 5 significant figures of
  sigma_kappa, the clipped variance and skewness, sigma_DL/D_L and
  P(ln mu > 1) (tests/golden/reference_scalars.json, extracted from the
  notebook's output by tests/golden/make_golden.py).
* The stored full-config fixture's N_expl (the halo rows' expected count,
  independent of the ingredients) to 1e-13.
* The shells of the walkthrough (cell 10): 67 shells at z_s = 1, first nodes
  (0.005, 0.01) and (0.0104, 0.0007).
* The inversion of a Gaussian characteristic function against the closed
  form of the 3-D Gaussian density, and the arc integral against an
  independent adaptive quadrature of that density along the arc.
* Lambda against a direct complex sum on a random table and the
  compound-Poisson cumulants (curvature at 0 = -sum w kappa^2 along k_kappa,
  -sum w gamma^2 / 2 along k_gamma).
"""
import json
import os

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import j0

from oracle import ref2d
from oracle.cosmo import Cosmology

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_scalars.json")))
LADDER = GOLD["explorer_ladder_zs1"]


def _rel5(v, target):
    """Agreement to the printed 5 significant figures (half a unit in the last place)."""
    digits = 5 - 1 - int(np.floor(np.log10(abs(target))))
    return abs(v - target) <= 0.5 * 10.0 ** (-digits) * 1.0001


@pytest.fixture(scope="module")
def cos():
    return Cosmology()


@pytest.mark.parametrize("model,kw", [("halos", {}), ("+ ellipticity", dict(ell=True)),
                                      ("+ filaments", dict(ell=True, fil=True)),
                                      ("+ subhalos", dict(ell=True, fil=True, sub=True))])
def test_ladder_row(model, kw):
    r = ref2d.run(1.0, **kw)
    row = ref2d.ladder_row(r)
    tgt = LADDER[model]
    for key in ("sigma_kappa", "clipped_var", "clipped_skew", "sigmaDL", "p_tail1"):
        assert _rel5(row[key], tgt[key]), (model, key, row[key], tgt[key])


def test_n_expl_matches_fixture(cos):
    tot = sum(r["Nbar"] for r in ref2d.rows(cos, 1.0))
    ref = GOLD["fixtures"]["pdf_zs1"]["N_expl"]
    assert abs(tot / ref - 1) < 1e-13


def test_shells_walkthrough():
    sh = ref2d.z_shells(1.0)
    assert len(sh) == 67
    assert abs(sh[0][2] - 0.005) < 1e-12 and abs(sh[0][3] - 0.01) < 1e-12
    assert round(sh[1][2], 4) == 0.0104 and round(sh[1][3], 4) == 0.0007
    assert sh[-1][1] == 1.0
    # the shells tile [0, z_s]
    assert abs(sum(s[3] for s in sh) - 1.0) < 1e-12


def test_gaussian_inversion_closed_form():
    """Phi = exp(-s_k^2 k_k^2 / 2 - s_g^2 k_g^2 / 2): kappa ~ N(0, s_k^2),
    (gamma_1, gamma_2) ~ N(0, s_g^2 I) -> P3 = N(kappa) e^{-gamma^2/(2 s_g^2)}/(2 pi s_g^2)."""
    sk, sg = 0.03, 0.025
    lamfn = lambda a, b: (-0.5 * sk ** 2 * np.asarray(a)[:, None] ** 2
                          - 0.5 * sg ** 2 * np.asarray(b)[None, :] ** 2) + 0j
    g = ref2d.grid_plan(lamfn, sk)
    k_k = np.linspace(0.0, g["k_k_max"], g["nk_k"])
    k_g = np.linspace(0.0, g["k_g_max"], g["nk_g"])
    kap = np.linspace(g["kap_min"], g["kap_max"], g["n_kap"])
    gam = np.linspace(0.0, g["gam_max"], g["n_gam"])
    P3 = ref2d.p3_joint(k_k, k_g, np.exp(lamfn(k_k, k_g)), kap, gam)

    def p3(k, gm):
        return (np.exp(-0.5 * k ** 2 / sk ** 2) / np.sqrt(2 * np.pi * sk ** 2)
                * np.exp(-0.5 * gm ** 2 / sg ** 2) / (2 * np.pi * sg ** 2))

    exact = p3(kap[:, None], gam[None, :])
    assert np.max(np.abs(P3 - exact)) < 1e-7 * exact.max()
    # arc integral of the closed form along constant-mu arcs (adaptive quadrature)
    xi = np.array([-0.08, -0.02, 0.0, 0.03, 0.1, 0.2])
    Pi = ref2d.arc_integral(kap, gam, exact, g["g_def"], xi)
    for x, v in zip(xi, Pi):
        mu = np.exp(x)
        hi = 1 - 1 / np.sqrt(mu)
        lo = max(kap[0], 1 - np.sqrt(1 / mu + g["g_def"] ** 2))
        f = lambda k: p3(k, np.sqrt(max((hi - k) * (1 - k + 1 / np.sqrt(mu)), 0.0)))
        ref = np.pi / mu ** 2 * quad(f, lo, hi, limit=400, epsabs=0, epsrel=1e-12)[0]
        gb = min(g["gam_max"], np.sqrt(max((hi - kap[0]) * (1 - kap[0] + 1 / np.sqrt(mu)), 0)))
        if gb > g["g_def"]:
            ft = lambda gm: p3(1 - np.sqrt(1 / mu + gm ** 2), gm) * gm / np.sqrt(1 / mu + gm ** 2)
            ref += np.pi / mu ** 2 * quad(ft, g["g_def"], gb, limit=400)[0]
        # the scheme's own discretisation (bilinear P3, 2400 points per arc): ~1e-4
        assert abs(v - ref) < 1e-3 * abs(ref) + 1e-9, (x, v, ref)


def test_lambda_direct_and_cumulants():
    rng = np.random.default_rng(7)
    n = 300
    pop = ref2d.Population(w=rng.uniform(0.1, 50.0, n), k=rng.normal(0, 0.02, n) * rng.choice(
        [1e-4, 1e-2, 1.0], n), g=np.abs(rng.normal(0, 0.05, n)) + 1e-9)
    kk = np.array([0.0, 1e-3, 0.7, 13.0, 250.0, 4000.0])
    kg = np.array([0.0, 2e-3, 1.1, 40.0, 900.0])
    L = ref2d.lam(pop, kk, kg)
    direct = (pop.w[None, None, :] * (np.exp(1j * kk[:, None, None] * pop.k[None, None, :])
                                      * j0(kg[None, :, None] * pop.g[None, None, :])
                                      - 1 - 1j * kk[:, None, None] * pop.k[None, None, :])).sum(-1)
    assert np.max(np.abs(L - direct)) < 1e-12 * pop.w.sum()
    # curvature at the origin: the compound-Poisson second cumulants
    h = 1e-3
    Lk = ref2d.lam(pop, np.array([h]), np.array([0.0]))[0, 0]
    Lg = ref2d.lam(pop, np.array([0.0]), np.array([h]))[0, 0]
    assert abs(Lk.real / (-0.5 * h * h * np.sum(pop.w * pop.k ** 2)) - 1) < 1e-6
    assert abs(Lg.real / (-0.25 * h * h * np.sum(pop.w * pop.g ** 2)) - 1) < 1e-6
