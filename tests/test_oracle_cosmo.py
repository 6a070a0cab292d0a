"""Pins of the oracle's cosmology against the reference's stored values and
closed forms (CPU only)."""
import json
import os

import numpy as np
import pytest
from scipy.integrate import quad, solve_ivp
from scipy.special import hyp2f1

from oracle.cosmo import CKMS, Cosmology

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_scalars.json")))


@pytest.fixture(scope="module")
def planck():
    return Cosmology()


def test_fixture_scalars(planck):
    """tests/reference/pdf_zs*.json dark_energy block (Omega_r exact, growth amplitude)."""
    for name, fx in GOLD["fixtures"].items():
        assert planck.Or == fx["Omega_r"], name
        assert abs(planck.primordial_growth_amplitude() / fx["primordial_growth"] - 1) < 1e-12


def test_walkthrough_cell3(planck):
    w = GOLD["walkthrough"]
    assert round(planck.sigma_tophat8(), 4) == w["sigma8_tophat"]
    assert round(planck.D(1.0), 4) == w["D1"]
    assert round(planck.D(3.0), 4) == w["D3"]


def test_walkthrough_cell7_nfw(planck):
    """M200 = 1e14 Msun at z = 0.4, z_s = 1: c, r_s, kappa_s, r_vir/r_s."""
    w = GOLD["walkthrough"]["nfw_1e14_z04"]
    C, rs, ks, fC = planck.nfw_params(1e14, 0.4, 1.0)
    assert round(C, 2) == w["c"]
    assert round(rs * 1e3, 1) == w["rs_kpc"]
    assert round(ks, 4) == w["kappa_s"]
    assert round(planck.eta(C, 0.4) * C, 2) == w["rvir_over_rs"]


@pytest.mark.parametrize("w", [-1.2, -1.0, -0.8])
def test_constant_w_growth_hypergeometric(w):
    """Radiation-free constant-w growth has the closed form
    D ~ a 2F1(-1/(3w), (w-1)/(2w); 1-5/(6w); -(OL/Om) a^(-3w))."""
    c = Cosmology(w0=w, radiation=None)
    z = np.array([0.0, 0.1, 0.5, 1.0, 3.0, 10.0, 30.0])

    def raw(a):
        return a * hyp2f1(-1 / (3 * w), (w - 1) / (2 * w), 1 - 5 / (6 * w), -c.OL / c.Om * a ** (-3 * w))

    np.testing.assert_allclose(c.D(z), raw(1 / (1 + z)) / raw(1.0), rtol=2e-8)
    np.testing.assert_allclose(c.primordial_growth_amplitude(), raw(1.0), rtol=2e-8)


def test_growth_independent_integrator(planck):
    """Same ODE in the variable a with an implicit solver (Radau)."""
    c = planck

    def rhs(a, y):
        E2 = c.E2(1 / a - 1)
        # dE2/da
        dE2 = -3 * c.Om / a ** 4 - 4 * c.Or / a ** 5
        dlnE = 0.5 * dE2 / E2
        return (y[1], -(3 / a + dlnE) * y[1] + 1.5 * c.Om / (a ** 5 * E2) * y[0])

    a0 = 1e-6
    sol = solve_ivp(rhs, (a0, 1.0), (a0 + 2 / 3 * c.Or / c.Om, 1.0), method="Radau",
                    rtol=1e-11, atol=1e-16, dense_output=True)
    for z in (0.5, 1.0, 3.0):
        a = 1 / (1 + z)
        assert abs(sol.sol(a)[0] / sol.y[0, -1] / c.D(z) - 1) < 1e-7


def test_distance_quadrature_and_curvature():
    c = Cosmology(Ok=0.05)
    chi = c.chi(2.0)
    ref = quad(lambda z: CKMS / c.H0 / np.sqrt(c.E2(z)), 0, 2.0, epsabs=0, epsrel=1e-13)[0]
    assert abs(chi / ref - 1) < 1e-12
    u = np.sqrt(0.05) * chi * c.H0 / CKMS
    assert abs(c.f_K(chi) / (np.sinh(u) / u * chi) - 1) < 1e-13


def test_mass_function_normalisation():
    """Sheth-Tormen normalisation: int_0^inf f(nu) dln(nu) = 1 with the oracle's
    (A, p, q); the oracle's dn/dlnM is rho/M f(nu) dln(nu)/dlnM."""
    from oracle import cosmo as oc
    f = lambda nu: (oc.A_ST * (1 + (oc.Q_ST * nu) ** -oc.P_ST) * np.sqrt(oc.Q_ST * nu / (2 * np.pi))
                    * np.exp(-oc.Q_ST * nu / 2) / nu)
    total = quad(f, 0, 1, limit=200)[0] + quad(f, 1, np.inf, limit=200)[0]
    assert abs(total - 1) < 1e-9


def test_dndlnm_chain_rule(planck):
    """dn/dlnM = rho/M f(nu) dln(nu)/dlnM, with nu from sigma(M, z) and the
    derivative taken by finite differences of sigma (independent of the spline's)."""
    from oracle import cosmo as oc
    M, z = 3e12, 0.7
    nu = (oc.DC / planck.sigmaM(M, z)) ** 2
    h = 1e-4
    dlnnu = (np.log((oc.DC / planck.sigmaM(M * np.exp(h), z)) ** 2)
             - np.log((oc.DC / planck.sigmaM(M * np.exp(-h), z)) ** 2)) / (2 * h)
    f = (oc.A_ST * (1 + (oc.Q_ST * nu) ** -oc.P_ST) * np.sqrt(oc.Q_ST * nu / (2 * np.pi))
         * np.exp(-oc.Q_ST * nu / 2))
    assert abs(planck.dndlnM(M, z) / (planck.rhom / M * f * dlnnu) - 1) < 1e-6


def test_legacy_growth_pins():
    """growth_mode="legacy" (the Heath integral, sgl.py:270-306): for matter +
    Lambda + curvature without radiation it is the exact growing mode, so it
    agrees with the growth ODE run without radiation; Einstein-de Sitter gives
    D = a; it never carries radiation (the reference's
    test_heath_integral_excludes_radiation)."""
    z = np.array([0.05, 0.5, 1.0, 3.0, 6.0, 10.0])
    for Ok in (0.0, 0.05, -0.05):
        leg = Cosmology(radiation=None, growth_mode="legacy", Ok=Ok)
        ode = Cosmology(radiation=None, Ok=Ok)
        np.testing.assert_allclose(leg.D(z), ode.D(z), rtol=1e-9)
    np.testing.assert_array_equal(Cosmology(growth_mode="legacy").D(z),
                                  Cosmology(radiation=None, growth_mode="legacy").D(z))
    eds = Cosmology(Om=1.0, Ob=0.05, radiation=None, growth_mode="legacy")
    np.testing.assert_allclose(eds.D(z), 1.0 / (1.0 + z), rtol=1e-13)
    with pytest.raises(ValueError):
        Cosmology(w0=-0.9, growth_mode="legacy")


def test_bias_growth_pins():
    """Dg (sgl_full.py:521-538): the ODE growth on growth_mode="ode"; the
    Carroll-Press-Turner fit on "legacy", which is exact for Einstein-de
    Sitter (Omega_m(z) = 1 gives g = 1, Dg = 1/(1+z)) and within 0.3 % of the
    exact LambdaCDM growth (the fit's published accuracy is ~1 %)."""
    z = np.array([0.05, 0.5, 1.0, 3.0, 6.0, 10.0])
    ode = Cosmology()
    np.testing.assert_array_equal(ode.Dg(z), ode.D(z))
    eds = Cosmology(Om=1.0, Ob=0.05, radiation=None, growth_mode="legacy")
    np.testing.assert_allclose(eds.Dg(z), 1.0 / (1.0 + z), rtol=1e-14)
    for rad in ("engine", None):
        leg = Cosmology(radiation=rad, growth_mode="legacy")
        assert np.max(np.abs(leg.Dg(z) / leg.D(z) - 1)) < 3e-3
        assert leg.Dg(0.0) == 1.0


def test_no_radiation_background():
    """radiation=None: E^2 = Om (1+z)^3 + 1 - Om for flat LambdaCDM, exactly
    (the reference's test_library_default_has_no_radiation_and_old_expression)."""
    z = np.array([0.1, 0.5, 1.0, 3.0, 6.0, 10.0])
    c = Cosmology(radiation=None)
    assert c.Or == 0.0 and c.OL == 1.0 - c.Om
    np.testing.assert_allclose(c.E(z), np.sqrt(c.Om * (1 + z) ** 3 + 1.0 - c.Om), rtol=1e-15)
