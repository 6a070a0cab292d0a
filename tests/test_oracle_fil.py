"""Oracle of the filament population (cylinder filaments of the reference's
population builder, sgl_full.py:1984-2007, filaments.py:39-77): each piece
pinned by mathematics the reference's formulas must satisfy."""
import math

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import chain
from oracle import cosmo as oc
from oracle.cosmo import Cosmology


@pytest.fixture(scope="module")
def planck():
    return Cosmology()


def test_orientation_rule_integrates_the_kink_exactly():
    """sgl_full.py:1828-1845: sum_n w_n g(phi_n) ~ (2/pi) int_0^{pi/2} g dphi
    for the amplitude g = 1 / max(rc, cos phi) (the kink at acos rc), whose
    integral is ln(sec + tan)(phi_c) + (pi/2 - phi_c) / rc."""
    ph, wp = chain.fil_phi_rule()
    assert ph.size == 13 and abs(wp.sum() - 1) < 1e-15
    rc = chain.FIL_RC
    pc = math.acos(rc)
    exact = (math.log(1 / math.cos(pc) + math.tan(pc)) + (0.5 * math.pi - pc) / rc) / (0.5 * math.pi)
    approx = float(np.sum(wp / np.maximum(rc, np.abs(np.cos(ph)))))
    assert abs(approx / exact - 1) < 1e-5          # the reference's "~1e-5 at n = 12"
    ph24, wp24 = chain.fil_phi_rule(24)            # and the rule converges to the integral
    assert abs(float(np.sum(wp24 / np.maximum(rc, np.abs(np.cos(ph24))))) / exact - 1) < 1e-10


def test_cylinder_shear_is_mean_enclosed_minus_local():
    """For an axially symmetric lens gamma = kbar - kappa, kbar the mean
    convergence inside u: integrate the oracle's kappa(u) = 2 pi k0 / sqrt(1 + u^2)
    numerically and compare with kappa + gamma of the oracle's cells."""
    rs, k0base = 0.3, 0.02
    F = np.array([[[k0base, rs, 1.0, 1.0]]])
    xi, ws, kap, gam = chain.cells_fil(F)
    ph, _ = chain.fil_phi_rule()
    ue = chain.fil_u_edges()
    uc = np.where(ue[:-1] > 0, np.sqrt(ue[:-1] * ue[1:]), 0.5 * ue[1])
    k0 = k0base / np.maximum(2 * rs, np.abs(np.cos(ph)) * 20 * rs)
    K = kap.reshape(uc.size, ph.size)
    G = gam.reshape(uc.size, ph.size)
    for li in (0, 40, 90, 119):
        u = uc[li]
        for n in (0, 7, 12):
            kap_fn = lambda v: 2 * math.pi * k0[n] / math.sqrt(1 + v * v)
            assert abs(K[li, n] / kap_fn(u) - 1) < 1e-14
            kbar = 2 / u ** 2 * quad(lambda v: kap_fn(v) * v, 0, u, epsabs=0, epsrel=1e-13)[0]
            assert abs((K[li, n] + G[li, n]) / kbar - 1) < 1e-10


def test_filament_cells_weight_and_tilt():
    """Cell weight = w_row x annulus area pi r_s^2 (u1^2 - u0^2) x orientation
    weight x detA (the source-plane tilt), and the u bins tile the unit disc."""
    rs, k0base, wrow = 0.5, 1e-3, 2.0
    F = np.array([[[k0base, rs, wrow, 1.0]]])
    xi, ws, kap, gam = chain.cells_fil(F)
    detA = (1 - kap) ** 2 - gam ** 2
    _, wp = chain.fil_phi_rule()
    ue = chain.fil_u_edges()
    area = np.pi * rs ** 2 * np.diff(ue ** 2)
    expect = wrow * area[:, None] * wp[None, :] * detA.reshape(area.size, wp.size)
    np.testing.assert_allclose(ws.reshape(expect.shape), expect, rtol=1e-14)
    np.testing.assert_allclose(xi.reshape(expect.shape), -np.log(detA.reshape(expect.shape)),
                               rtol=1e-14)
    assert abs(np.sum(area) - np.pi * rs ** 2) < 1e-14


def test_filament_mass_function_normalised_and_bias_consistent():
    """Flat barrier (p, q, A) = (0, 0.7, 1/2): int f dln(nu) = 1 and the
    peak-background bias b = 1 + (q nu - 1) / delta_c satisfies
    int b f dln(nu) = 1 (q nu is chi-square with one degree of freedom)."""
    q, A = chain.Q_FIL, chain.A_FIL
    f = lambda nu: A * 2 * np.sqrt(q * nu / (2 * np.pi)) * np.exp(-q * nu / 2) / nu
    b = lambda nu: 1 + (q * nu - 1) / oc.DC0
    tot = quad(f, 0, 1, limit=200)[0] + quad(f, 1, np.inf, limit=200)[0]
    totb = (quad(lambda v: b(v) * f(v), 0, 1, limit=200)[0]
            + quad(lambda v: b(v) * f(v), 1, np.inf, limit=200)[0])
    assert abs(tot - 1) < 1e-9 and abs(totb - 1) < 1e-9


def test_filament_rows_chain_rule_and_bias(planck):
    """dn_fil/dlnM = rho/M f(nu) dln(nu)/dlnM with nu at (M, z) and the
    derivative by finite differences; beta = D b(nu0) with nu0 = (dc0 / D sigma0)^2."""
    M, z = 2e13, 0.6
    nu = (oc.DC / planck.sigmaM(M, z)) ** 2
    h = 1e-4
    dlnnu = (np.log((oc.DC / planck.sigmaM(M * np.exp(h), z)) ** 2)
             - np.log((oc.DC / planck.sigmaM(M * np.exp(-h), z)) ** 2)) / (2 * h)
    f = chain.A_FIL * 2 * np.sqrt(chain.Q_FIL * nu / (2 * np.pi)) * np.exp(-chain.Q_FIL * nu / 2)
    assert abs(chain.dndlnM_fil(planck, M, z) / (planck.rhom / M * f * dlnnu) - 1) < 1e-6
    nu0 = (oc.DC0 / planck.D(z) / planck.sigmaM(M, 0.0)) ** 2
    assert abs(chain.fil_bias(planck, M, z) - (1 + (0.7 * nu0 - 1) / oc.DC0)) < 1e-14


def test_pdf_with_filaments_small_shape_properties():
    """With the filament population added: unit mass, zero mean (compensated),
    variance = the summed jump variance of both populations + the two-point term."""
    sh = chain.Shape(n_z=8, n_m=32, n_b=64, n_k=2048, n_out=4001, out_lo=-1.5, out_hi=4.0,
                     filaments=True)
    r = chain.pdf(3.0, sh, keep=True)
    m = r.info["moments"]
    assert r.info["n_cells_fil"] > 0
    assert abs(m["norm"] - 1) < 1e-4
    assert abs(m["mean"]) < 1e-3
    var_pred = (r.info["var_xi"] + r.info["var_xi_fil"]
                + 2 * abs(r.info["psi2"][0].real) / r.info["kq"][1] ** 2)
    assert abs(m["var"] / var_pred - 1) < 0.05
