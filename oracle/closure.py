"""Positive ('vp') clustering closure on the kernel-shape chain (checker only).

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module.

The reference's positive closure (sgl_analytic/pos_closure.py:1-733) replaces
the Gaussian two-point term Psi_2 by the characteristic function of a
positive random lens intensity:

  ln R(k) = ln E_U exp( sum_s sum_g (lam_g(U_s) - 1) Y_sg(k) ),
  Y_sg(k) = sum_{c in slice s, group g} w_c (e^{i k xi_c} - 1),

U_s a stationary Markov chain of 48 Gauss-Hermite states along the line of
sight (Gaussian-copula kernel with the correlation rho_s of neighbouring
slices), lam_g the variance-preserving rectified rates with E_w lam = 1 and
Var_w lam = v_g = eta f_s phi_g^2 exactly, f_s = (1 - rho_bar)/(1 + rho_bar),
and eta fitted so that the model's jump clustering variance equals the
Gaussian closure's.  The reference takes the loadings phi and the rho_s from
its Limber (xi_lin) machinery; this chain's clustering is the slice
covariance C_ij of the smoothed 1-D field (sgl_full.py:2514-2538), in which
every cell of slice s is modulated by the same Gaussian delta_s with weight
beta_c.  So here: groups = mass rows, phi_g = beta_g sqrt(C_ss),
rho_s = C_{s,s+1} / sqrt(C_ss C_{s+1,s+1}), and the target variance is
A^T C A with A_s = sum_{c in s} beta_c w_c xi_c (the k^2 coefficient of
Psi_2).  Phi = exp(Lambda + ln R) replaces exp(Lambda + Psi_2); the 1-D chain
has no separate shear term.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import brentq

NSTATE = 48


def gh_nodes(n: int = NSTATE):
    """Probabilists' Gauss-Hermite nodes with weights summing to 1 (pos_closure.py:465-467)."""
    u, w = np.polynomial.hermite_e.hermegauss(n)
    return u, w / w.sum()


def transition(u, w, rho):
    """Reversible Gaussian-copula transition on the nodes with exactly stationary w
    (pos_closure.py:251-267: symmetric balancing of J_ij = w_i w_j exp(copula log-ratio))."""
    x, y = u[:, None], u[None, :]
    J = w[:, None] * w[None, :] * np.exp((2 * rho * x * y - rho * rho * (x * x + y * y))
                                         / (2 * (1 - rho * rho)))
    v = np.ones(u.size)
    for _ in range(10000):
        marg = v * (J @ v)
        if np.max(np.abs(marg / w - 1)) < 3e-14:
            break
        v *= np.sqrt(w / marg)
    else:
        raise RuntimeError("stationary transition balancing did not converge")
    return (v[:, None] * J * v[None, :]) / w[:, None]


def vp_rates(v, u, w):
    """Rates (G, n) with E_w lam = 1 and Var_w lam = v per group for positive
    loadings (pos_closure.py:273-308): lam = (u - t)^+ / E_w[(u - t)^+], the
    threshold t from the quadratic on the node segment where {u > t} is fixed;
    all nodes positive gives exactly the linear rate 1 + sqrt(v) u."""
    v = np.asarray(v, float)
    A = np.cumsum((w * u)[::-1])[::-1]
    B = np.cumsum(w[::-1])[::-1]
    C = np.cumsum((w * u * u)[::-1])[::-1]
    tb = u[:-1]
    E1 = A[1:] - tb * B[1:]
    E2 = C[1:] - 2 * tb * A[1:] + tb * tb * B[1:]
    rv = E2 / E1 ** 2 - 1.0          # relative variance at t = u[j-1], increasing in j
    if np.any(v >= rv[-1]):
        raise ValueError(f"vp rates: relative variance {v.max():.3g} beyond the cap {rv[-1]:.3g}")
    j = np.searchsorted(rv, v, side="right")
    a, b, c = A[j], B[j], C[j]
    q2 = b - (1 + v) * b * b
    q1 = 2 * a * ((1 + v) * b - 1)
    q0 = c - (1 + v) * a * a
    lo = np.where(j > 0, u[np.maximum(j - 1, 0)], -np.inf)
    hi = u[j]
    with np.errstate(divide="ignore", invalid="ignore"):
        disc = np.sqrt(np.maximum(q1 * q1 - 4 * q2 * q0, 0.0))
        r1 = np.where(q2 != 0, (-q1 - disc) / (2 * q2), -q0 / q1)
        r2 = np.where(q2 != 0, (-q1 + disc) / (2 * q2), -q0 / q1)
    t = np.where((r1 >= lo) & (r1 < hi), r1, r2)
    x = np.maximum(u[None, :] - t[:, None], 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):   # the linear rows are set below
        lam = x / (x @ w)[:, None]
    lin = j == 0
    lam[lin] = 1.0 + np.sqrt(v[lin])[:, None] * u[None, :]
    return lam


def chain_variance(ms, w, K):
    """Var of sum_s m_s(U_s) for the stationary chain (pos_closure.py:389-403)."""
    h, var = None, 0.0
    for s, m in enumerate(ms):
        mu = m - float(w @ m)
        var += float(w @ (mu * mu))
        if s == 0:
            h = w * mu
        else:
            hp = K[s - 1].T @ h
            var += 2 * float(mu @ hp)
            h = hp + w * mu
    return var


class VP:
    """Closure set-up for one instance: beta [n_z, n_g], wk [n_z, n_g] =
    sum over a group's cells of w xi, C [n_z, n_z]."""

    def __init__(self, beta, wk, C, nstate: int = NSTATE):
        self.u, self.w = gh_nodes(nstate)
        n_z = beta.shape[0]
        d = np.sqrt(np.diag(C))
        self.rho = np.array([C[s, s + 1] / (d[s] * d[s + 1]) for s in range(n_z - 1)])
        rbar = np.array([np.mean([self.rho[j] for j in (s - 1, s) if 0 <= j < self.rho.size])
                         if self.rho.size else 0.0 for s in range(n_z)])
        self.f = (1 - rbar) / (1 + rbar)
        self.phi = beta * d[:, None]
        self.wk = wk
        self.K = [transition(self.u, self.w, float(r)) for r in self.rho]
        A = np.sum(beta * wk, axis=1)
        self.target = float(A @ C @ A)
        self.eta = self.calibrate()
        self.D = np.stack([vp_rates(self.eta * self.f[s] * self.phi[s] ** 2, self.u, self.w) - 1.0
                           for s in range(n_z)])            # [n_z, n_g, nstate]

    def variance(self, eta):
        ms = [vp_rates(eta * self.f[s] * self.phi[s] ** 2, self.u, self.w).T @ self.wk[s]
              for s in range(self.wk.shape[0])]
        return chain_variance(ms, self.w, self.K)

    def calibrate(self, lo=1e-3, hi=1e4):
        """eta with model variance = target (pos_closure.py:575-591: brentq in ln eta,
        the bracket widened by 1.5 per side from [1/1.5, 1.5] while it stays
        inside [lo, hi]; a root outside raises ValueError)."""
        fun = lambda le: self.variance(np.exp(le)) - self.target
        a, b = -np.log(1.5), np.log(1.5)
        fa, fb = fun(a), fun(b)
        while fa * fb > 0:
            if fa > 0 and a > np.log(lo):
                b, fb, a = a, fa, a - np.log(1.5)
                fa = fun(a)
            elif fb < 0 and b < np.log(hi):
                a, fa, b = b, fb, b + np.log(1.5)
                fb = fun(b)
            else:
                raise ValueError(f"positive closure: clustering target {self.target:.3e} "
                                 f"outside eta in [{lo}, {hi}]")
        return float(np.exp(brentq(fun, a, b, xtol=2e-12)))

    def lnR(self, Y):
        """ln E exp(sum_s z_s(U_s)), z_s(u_j) = sum_g D_sgj Y_sg, for Y [n_z, n_g, n_k]
        (pos_closure.py:183-240: forward recursion, rescaled every slice)."""
        n_z, _, n_k = Y.shape
        alpha = np.broadcast_to(self.w[:, None], (self.w.size, n_k)).astype(complex)
        scale = np.zeros(n_k)
        for s in range(n_z):
            z = np.einsum("gj,gk->jk", self.D[s], Y[s])
            if s:
                alpha = self.K[s - 1].T @ alpha
            mx = z.real.max(axis=0)
            alpha = alpha * np.exp(z - mx)
            nrm = np.max(np.abs(alpha), axis=0)
            alpha = alpha / nrm
            scale += mx + np.log(nrm)
        return scale + np.log(alpha.sum(axis=0))
