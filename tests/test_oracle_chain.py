"""Pins of the oracle's chain (CPU): each against an independent computation
that a plausible mistake (dropped term, wrong sign or index) would fail."""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import chain, nfw
from oracle.cosmo import Cosmology

rng = np.random.default_rng(7)


def _tiny_table(n_slices=3, per=40):
    xi = 10 ** rng.uniform(-6, 0.5, size=(n_slices, 1, per))
    xi[0, 0, :3] = -10 ** rng.uniform(-5, -1, size=3)      # a few negative jumps
    ws = 10 ** rng.uniform(-4, 1, size=xi.shape)
    beta = rng.uniform(0.5, 3.0, size=(n_slices, 1))
    return xi, ws, beta


def test_lnphi_against_extended_precision():
    """Compensated FP64 forms vs the direct e^{ik xi} - 1 - ik xi in 40-digit
    arithmetic (no cancellation issue there)."""
    xi, ws, beta = _tiny_table()
    k = np.array([1e-3, 0.7, 13.0, 850.0, 2.1e5])
    lam, amp = chain.lnphi(xi, ws, beta, k)
    with mp.workdps(40):
        _check_lnphi_extended(xi, ws, beta, k, lam, amp)


def _check_lnphi_extended(xi, ws, beta, k, lam, amp):
    for n, kk in enumerate(k):
        L = mp.mpc(0)
        A = [mp.mpc(0)] * xi.shape[0]
        for s in range(xi.shape[0]):
            for c in range(xi.shape[2]):
                ph = mp.mpf(kk) * mp.mpf(xi[s, 0, c])
                e = mp.expj(ph)
                L += mp.mpf(ws[s, 0, c]) * (e - 1 - 1j * ph)
                A[s] += mp.mpf(beta[s, 0] * ws[s, 0, c]) * (e - 1)
        # FP64 phase rounding: |d ph| <= eps k |xi|, times the total weight
        tol_ph = 4e-16 * kk * np.max(np.abs(xi)) * np.sum(ws * beta[:, :, None])
        assert abs(complex(L) - lam[n]) <= 1e-13 * max(abs(complex(L)), 1e-6) + tol_ph
        for s in range(xi.shape[0]):
            assert abs(complex(A[s]) - amp[s, n]) <= 1e-13 * max(abs(complex(A[s])), 1e-6) + tol_ph


def test_lnphi_small_k_cumulants():
    """ln phi(k) = -k^2 m2 / 2 - i k^3 m3 / 6 + k^4 m4 / 24 + ...  with
    m_n = sum w xi^n (compound-Poisson cumulants)."""
    xi, ws, beta = _tiny_table()
    k = np.array([1e-2])
    lam, _ = chain.lnphi(xi, ws, beta, k)
    m = [float(np.sum(ws * xi ** n)) for n in range(5)]
    kk = k[0]
    re = -kk ** 2 * m[2] / 2 + kk ** 4 * m[4] / 24
    im = -kk ** 3 * m[3] / 6
    assert abs(lam[0].real - re) < 1e-6 * abs(re)
    assert abs(lam[0].imag - im) < 1e-4 * abs(im)


def test_quadform_bruteforce():
    n, nk = 7, 5
    X = rng.normal(size=(n, n))
    C = X @ X.T
    a = rng.normal(size=(n, nk)) + 1j * rng.normal(size=(n, nk))
    ref = np.zeros(nk, complex)
    for q in range(nk):
        for i in range(n):
            for j in range(n):
                ref[q] += 0.5 * C[i, j] * a[i, q] * a[j, q]     # complex square, no conjugate
    np.testing.assert_allclose(chain.quadform(C, a), ref, rtol=1e-13)


def test_filon_inverts_gaussian():
    """phi = exp(i mu k - s^2 k^2 / 2) on the production k grid must give the
    normal density pointwise (p(t) = (1/pi) Re int_0^inf phi e^{-ikt} dk)."""
    sh = chain.Shape()
    kq = sh.k_nodes()
    for mu, s in ((0.02, 0.05), (-0.1, 0.3)):
        phi = np.exp(1j * mu * kq - 0.5 * s * s * kq * kq)
        t = np.linspace(mu - 6 * s, mu + 6 * s, 97)
        p = chain.invert(t, kq, phi)
        ref = np.exp(-0.5 * ((t - mu) / s) ** 2) / (s * math.sqrt(2 * math.pi))
        assert np.max(np.abs(p - ref)) < 2e-6 * ref.max()   # linear-phi interpolation error


def test_filon_matrix_matches_direct():
    sh = chain.Shape(n_k=300)
    kq = sh.k_nodes()
    phi = np.exp(-0.5 * (0.05 * kq) ** 2 + 0.01j * kq)
    t = np.linspace(-0.5, 3.0, 17)
    W = chain.filon_weights(t, kq)
    p1 = W @ np.stack([phi.real, phi.imag], -1).reshape(-1)
    np.testing.assert_allclose(p1, chain.invert(t, kq, phi), rtol=1e-12, atol=1e-14)


def test_slice_covariance_against_numerical_projection():
    """C_ij = sum_q 2 P1D(k_q)/L Re[s_i(k_q) conj(s_j(k_q))], s_i(k) the slice
    average of e^{ik chi}: here s_i by quadrature, not the closed form."""
    cos = Cosmology()
    zs, nz = 1.0, 6
    z, mid, dchi, edges = chain.slice_nodes(cos, zs, nz)
    chis = cos.chi(zs)
    C = chain.slice_covariance(cos, edges, chis)
    assert np.allclose(C, C.T)
    assert np.all(np.linalg.eigvalsh(C) > -1e-12 * np.max(C))
    # the slice projection: P1D (pinned on its own against a perpendicular
    # quadrature, tests/test_oracle_pins.py) at the modes, the slice averages
    # by Gauss-Legendre instead of the closed form
    Rs, L = 20.0, 1.05 * chis
    nq = max(4, int(np.floor(L / Rs)))
    kq = 2 * np.pi * np.arange(1, nq + 1) / L
    P1D = chain.p1d(cos, kq, Rs)
    g, gw = np.polynomial.legendre.leggauss(64)
    S = np.zeros((nz, nq), complex)
    for i in range(nz):
        a, b = edges[i], edges[i + 1]
        c = 0.5 * (a + b) + 0.5 * (b - a) * g
        S[i] = (np.exp(1j * np.outer(kq, c)) * gw).sum(1) * 0.5
    Cn = (2 * P1D / L * S[:, None, :] * np.conj(S[None, :, :])).real.sum(-1)
    np.testing.assert_allclose(C, Cn, rtol=1e-9, atol=1e-12 * np.max(np.abs(C)))


def test_cells_match_python_restatement():
    """oracle_cells (C) vs nfw.py + the jump/tilt rule written out here."""
    R = np.array([[[0.08, 0.15, 5.0, 0.9, 3.0, 1.2], [2.5, 0.05, 9.0, 0.5, 1e-3, 2.0]]])
    sh = chain.Shape(n_b=40)
    xi, ws, kap, gam = chain.cells(R, sh)
    ue = sh.u_edges()
    for j in range(2):
        ks, rs, xt, rv, w, beta = R[0, j]
        b0, b1 = rv * ue[:-1], rv * ue[1:]
        x = np.sqrt(b0 * b1) / rs
        k, g = nfw.nfw_trunc_kg(x, ks, xt)
        det = (1 - k) ** 2 - g ** 2
        ok = det > 0
        xr = np.where(ok, -np.log(np.where(ok, det, 1.0)), 0.0)
        keep = ok & (xr <= 30)
        wr = np.where(keep, w * np.pi * (b1 ** 2 - b0 ** 2) * det, 0.0)
        np.testing.assert_allclose(kap[0, j], k, rtol=1e-13, atol=1e-300)
        np.testing.assert_allclose(xi[0, j][keep], xr[keep], rtol=1e-12)
        np.testing.assert_allclose(ws[0, j], wr, rtol=1e-12)
    assert np.any(ws[0, 1] == 0)       # the strong lens has cells inside its critical curves


def test_pdf_small_shape_properties():
    """Unit mass of the discretised transform (phi(0) = 1) and consistency of
    the inversion with the exponent: grid mass on a wide fine grid ~ 1 at z_s = 3."""
    sh = chain.Shape(n_z=8, n_m=32, n_b=64, n_k=2048, n_out=4001, out_lo=-1.5, out_hi=4.0)
    r = chain.pdf(3.0, sh, keep=True)
    m = r.info["moments"]
    assert abs(m["norm"] - 1) < 1e-4
    assert abs(m["mean"]) < 1e-3                  # compensated: <xi> = 0
    var_pred = r.info["var_xi"] + 2 * abs(r.info["psi2"][0].real) / r.info["kq"][1] ** 2
    assert abs(m["var"] / var_pred - 1) < 0.05


def test_edge_rvir_rows():
    """edge="rvir": x_t = r_vir / r_s = eta c, above c where Delta_vir < 200
    (eta > 1) and equal to c where the r_vir root sits at the bracket end
    (eta = 1); every other row entry is the r200 row's."""
    from oracle import chain
    from oracle.cosmo import Cosmology
    cos = Cosmology(h=0.674, Om=0.315, s8=0.811, Ob=0.0493, ns=0.965, zeq=3402.0)
    sh200 = chain.Shape(n_z=6, n_m=8, n_b=16, n_k=64, n_out=32)
    shv = chain.Shape(n_z=6, n_m=8, n_b=16, n_k=64, n_out=32, edge="rvir")
    a = chain.rows(cos, 2.0, sh200)["rows"]
    b = chain.rows(cos, 2.0, shv)["rows"]
    assert np.array_equal(np.delete(a, 2, axis=2), np.delete(b, 2, axis=2))
    eta = b[:, :, 3] / b[:, :, 1] / a[:, :, 2]          # r_vir / (r_s c)
    assert np.allclose(b[:, :, 2], eta * a[:, :, 2], rtol=1e-14)
    assert np.all(eta >= 1.0 - 1e-15) and np.any(eta > 1.05)
    with pytest.raises(ValueError):
        chain.rows(cos, 2.0, chain.Shape(n_z=6, n_m=8, n_b=16, n_k=64, n_out=32, edge="r500"))


def _psi_derivs(k0, g0, X1, X2, eps):
    """Second derivatives of the pseudo-elliptical potential psi0(xi),
    xi^2 = (1 - eps) X1^2 + (1 + eps) X2^2, from the circular profile at xi:
    psi0'' = k0 - g0 and psi0'/xi = k0 + g0 (chain rule, no lensing formula)."""
    a1, a2 = 1.0 - eps, 1.0 + eps
    xi2 = a1 * X1 ** 2 + a2 * X2 ** 2
    p2, p1x = k0 - g0, k0 + g0
    u1, u2 = a1 * X1, a2 * X2
    p11 = p2 * u1 ** 2 / xi2 + p1x * (a1 - u1 ** 2 / xi2)
    p22 = p2 * u2 ** 2 / xi2 + p1x * (a2 - u2 ** 2 / xi2)
    p12 = (p2 - p1x) * u1 * u2 / xi2
    return p11, p22, p12


def test_pseudo_elliptical_profile_pins():
    """kg_eps (sgl_full.py:582-613): eps = 0 is the circular profile; kappa is
    exactly half the Laplacian of the pseudo-elliptical potential; gamma^2
    agrees with the potential's shear to first order in eps (the engine's
    eps^2 term is its documented cos^2 form, so the difference is O(eps^2):
    halving eps quarters it); beyond the truncation it is the monopole at x."""
    ks, xt = 0.3, 6.0
    x = np.geomspace(1e-3, 5.0, 41)
    phi = np.linspace(0.03, 1.5, 41)
    k, g = chain.kg_eps(x, phi, 0.0, ks, xt)
    kc, gc = chain.trunc_kg(x, ks, xt)
    np.testing.assert_allclose(k, kc, rtol=1e-14)
    np.testing.assert_allclose(g, gc, rtol=1e-14)
    diffs = []
    for eps in (0.2, 0.1):
        k, g = chain.kg_eps(x, phi, eps, ks, xt)
        X1, X2 = x * np.cos(phi), x * np.sin(phi)
        xe = np.sqrt((1 - eps) * X1 ** 2 + (1 + eps) * X2 ** 2)
        k0, g0 = chain.trunc_kg(xe, ks, xt)
        p11, p22, p12 = _psi_derivs(k0, g0, X1, X2, eps)
        np.testing.assert_allclose(k, 0.5 * (p11 + p22), rtol=1e-12, atol=1e-15)
        gex2 = 0.25 * (p11 - p22) ** 2 + p12 ** 2
        diffs.append(np.max(np.abs(g ** 2 - gex2) / (k0 ** 2 + g0 ** 2)))
    print("O(eps^2) shear differences", diffs)
    assert diffs[0] < 0.1 and 3.0 < diffs[0] / diffs[1] < 5.0
    xo = np.array([7.0, 9.0, 20.0])
    k, g = chain.kg_eps(xo, 0.7, 0.25, ks, xt)
    mut = np.log(1 + xt) - xt / (1 + xt)
    np.testing.assert_allclose(k, 0.0)
    np.testing.assert_allclose(g, 4 * ks * mut / xo ** 2, rtol=1e-14)


def test_mstar_and_epsilon():
    """M*(z) solves sigma(M*, 0) Dg(z) = delta_c0 on the reference's table, and
    eps(M*) = (1 - 0.54) / (1 + 0.54) (Allgood+06)."""
    from oracle.cosmo import DC0
    c = Cosmology()
    for z in (0.0, 0.5, 2.0):
        m = c.Mstar(z)
        assert abs(np.log(c.sigmaM(m, 0.0) * c.Dg(z) / DC0)) < 1e-5
        assert abs(c.epsilon(m, z) - 0.46 / 1.54) < 1e-5
    assert c.Mstar(0.0) > c.Mstar(2.0)
    assert c.epsilon(1e6, 0.0) < c.epsilon(1e14, 0.0)
