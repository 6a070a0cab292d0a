"""Pins of the oracle's positive ('vp') closure (CPU): the rates' exact mean and
variance, the transition's stationarity and reversibility, the calibration,
the recursion against brute-force expectations, the k -> 0 limit against the
Gaussian closure, and |Phi| <= 1."""
import itertools

import numpy as np
import pytest

from oracle import chain, closure
from oracle.cosmo import Cosmology


def test_vp_rates_moments():
    u, w = closure.gh_nodes()
    v = np.array([1e-4, 1e-2, 0.05, 0.3, 1.0, 3.0])
    lam = closure.vp_rates(v, u, w)
    assert np.all(lam >= 0)
    np.testing.assert_allclose(lam @ w, 1.0, rtol=1e-12)
    np.testing.assert_allclose((lam - 1) ** 2 @ w, v, rtol=1e-10)
    # small variance: exactly the linear rate
    np.testing.assert_allclose(lam[0], 1 + np.sqrt(v[0]) * u, rtol=1e-14)


def test_transition_stationary_reversible():
    u, w = closure.gh_nodes()
    for rho in (0.1, 0.6, 0.95):
        K = closure.transition(u, w, rho)
        np.testing.assert_allclose(w @ K, w, atol=1e-13)
        np.testing.assert_allclose(K.sum(axis=1), 1.0, atol=1e-12)
        np.testing.assert_allclose(w[:, None] * K, (w[:, None] * K).T, atol=1e-15)
        # the copula's lag-one correlation of the node values
        corr = float(u @ (w[:, None] * K) @ u) / float(w @ (u * u))
        assert abs(corr - rho) < 1e-3


def test_recursion_against_brute_force():
    """Three slices, 5 states: E exp(sum_s z_s(U_s)) summed over every path."""
    rng = np.random.default_rng(3)
    u, w = closure.gh_nodes(5)
    K = [closure.transition(u, w, 0.5), closure.transition(u, w, 0.3)]
    z = rng.normal(size=(3, 5)) * 0.3 + 1j * rng.normal(size=(3, 5))
    ref = 0.0
    for p in itertools.product(range(5), repeat=3):
        ref += w[p[0]] * K[0][p[0], p[1]] * K[1][p[1], p[2]] * np.exp(z[0, p[0]] + z[1, p[1]] + z[2, p[2]])
    vp = closure.VP.__new__(closure.VP)
    vp.u, vp.w, vp.K = u, w, K
    vp.D = np.stack([np.eye(5)] * 3)                        # z_s(u_j) = Y_s[j]
    got = vp.lnR(z[:, :, None])[0]
    assert abs(got - np.log(ref)) < 1e-12


@pytest.mark.parametrize("zs", [0.3, 1.0])
def test_vp_chain_limits(zs):
    """Calibrated model variance = the Gaussian closure's target; ln R(k) ->
    Psi_2(k) ~ -k^2 target / 2 as k -> 0; |Phi| <= 1 everywhere."""
    osh = chain.Shape(n_z=8, n_m=32, n_b=64, n_k=1024, n_out=256)
    cos = Cosmology()
    T = chain.rows(cos, zs, osh)
    R = T["rows"]
    xi, ws, _, _ = chain.cells(R, osh)
    C = chain.slice_covariance(cos, T["chi_edges"], T["chis"])
    vp = closure.VP(R[:, :, 5], chain.vp_inputs(xi, ws, R[:, :, 5]), C)
    assert abs(vp.variance(vp.eta) / vp.target - 1) < 1e-9
    n_z, n_m = R.shape[:2]
    # small k: both closures -> -k^2 target / 2 (deviations ~ k^2 from the
    # fourth cumulants, 1.5e-3 at k^2 target / 2 = 1e-7); then a wide k range
    k0 = np.sqrt(2 * np.array([1e-9, 1e-8]) / vp.target)
    kq = np.concatenate([k0, np.geomspace(1e-2, 1e4, 64)])
    lam, amp = chain.lnphi(xi, ws, R[:, :, 5], kq)
    _, Yr = chain.lnphi(xi.reshape(n_z * n_m, 1, -1), ws.reshape(n_z * n_m, 1, -1),
                        np.ones((n_z * n_m, 1)), kq)
    lnR = vp.lnR(Yr.reshape(n_z, n_m, -1))
    psi2 = chain.quadform(C, amp)
    ref = -0.5 * k0 ** 2 * vp.target
    np.testing.assert_allclose(lnR[:2].real / ref, 1.0, rtol=1e-3)
    np.testing.assert_allclose(psi2[:2].real / ref, 1.0, rtol=1e-3)
    assert np.all(np.abs(np.exp(lam + lnR)) <= 1 + 1e-12)


def test_vp_filament_groups_end_to_end():
    """With the filament rows as further groups: the calibration still meets
    the Gaussian closure's k^2 coefficient (halo + filament A_s), phi stays a
    characteristic function (|phi| <= 1), and the PDF keeps the grid mass and
    mean of the Gaussian closure to the closures' higher-cumulant difference."""
    osh = chain.Shape(n_z=6, n_m=16, n_b=32, n_k=512, n_out=256, filaments=True)
    rv = chain.pdf(0.5, osh, closure="vp", keep=True)
    rg = chain.pdf(0.5, osh, closure="gauss", keep=True)
    T = rv.info["tables"]
    F = chain.fil_rows(rv.info["cos"], 0.5, osh)
    fxi, fws, _, _ = chain.cells_fil(F)
    A = (np.sum(T["rows"][:, :, 5] * chain.vp_inputs(rv.info["xi"], rv.info["ws"], None), axis=1)
         + np.sum(F[:, :, 3] * chain.vp_inputs(fxi, fws, None), axis=1))
    assert abs(A @ rv.info["C"] @ A / rv.info["vp_target"] - 1) < 1e-12
    assert np.all(np.abs(rv.info["phi"]) <= 1 + 1e-12)
    dx = rv.x[1] - rv.x[0]
    mv, mg = rv.p.sum() * dx, rg.p.sum() * dx
    assert abs(mv - mg) < 1e-3
    assert abs(rv.info["moments"]["mu_mean"] - rg.info["moments"]["mu_mean"]) < 1e-3
    assert 1e-7 < np.sum(np.abs(rv.p - rg.p)) * dx < 0.05


def test_vp_calibration_domain():
    """The calibration keeps eta inside [1e-3, 1e4] (pos_closure.py:575-591): a
    target that needs eta far below the range raises, one inside calibrates to
    the model variance."""
    osh = chain.Shape(n_z=8, n_m=32, n_b=64, n_k=1024, n_out=256)
    cos = Cosmology()
    T = chain.rows(cos, 1.0, osh)
    R = T["rows"]
    xi, ws, _, _ = chain.cells(R, osh)
    C = chain.slice_covariance(cos, T["chi_edges"], T["chis"])
    vp = closure.VP(R[:, :, 5], chain.vp_inputs(xi, ws, R[:, :, 5]), C)
    eta0 = vp.eta
    vp.target *= 1e-6
    with pytest.raises(ValueError):
        vp.calibrate()
    vp.target *= 1e6 * 3.0
    eta = vp.calibrate()
    assert 1e-3 <= eta <= 1e4 and eta > eta0
    assert abs(vp.variance(eta) / vp.target - 1) < 1e-9
