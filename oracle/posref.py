"""Positive ('vp') clustering closure on the reference 2-D chain (checker only).

TEST INFRASTRUCTURE: imported only by tests/, never by the product.

pipeline.run(..., xi_closure="vp") (pipeline.py:154-207, 231-236) builds the
xi_lin population with xi['pos'] = True and replaces the Gaussian pair term
Psi_2 of Lam._Qxi by the positive closure of pos_closure.install
(pos_closure.py:708-733):

  Phi = exp(Lambda) * R(k_kappa, k_gamma) * exp(-shear_var k_gamma^2 / 4),
  ln R = ln E_U exp( sum_s sum_c W_sc (lam_sc(U_s) - 1) F_c ),
  F_c = e^{i k_kappa kappa_c} J0(k_gamma gamma_c) - 1,

with, per lens shell s (pos_closure.population, pos_closure.py:388-449):

* cells: the shell's own fine-lattice bins (weights W = the shell's h0), at
  the pooled centroids (kappa, gamma) that Lambda uses;
* loadings phi: per coarse bin the shell's m = 0 xi group's 'kappa' mode
  r0 = (m / |m|) . rows, m_r = rows_r . kappa_c + res_r, divided by the bin's
  lens rate W_b and signed so that r0 . W_b >= 0; a cell takes its coarse
  parent's loading by its own (per-shell) centroid;
* rho between consecutive shells from the top-hat line-of-sight correlation
  of the linear field at the k_perp nodes, weighted by the groups' leading
  k_perp profiles |v0| (los_rho, pos_closure.py:372-385);
* f_s = (1 - rho_bar) / (1 + rho_bar);
* target_kappa = the clustering variance of Lam.sigma, target_gamma = the
  same sum over the m = 1 groups (the shear variance);

then (Closure, pos_closure.py:453-591): 48 Gauss-Hermite states, the
stationary transitions, groups of equal loading per shell, eta calibrated by
brentq in ln eta so that the chain's kappa clustering variance equals
target_kappa; ln R by the rates' interval form (FusedVP,
pos_closure_fused.py:18-140) and the log-scaled forward recursion, computed
in oracle/chain.c (oracle_vp_lnr); qxi unwraps Im ln R along k_kappa at the
first k_gamma, then along k_gamma, and subtracts shear_var k_gamma^2 / 4
(pos_closure.py:690-705).  In invert (sgl_full.py:226-238) the closure's
coarse grid follows the box: n = min(401, max(101, 1 + ceil(12.65
sqrt(k_g_max sigma)))) nodes along k_gamma, nk = min(4001, max(n, 1 +
ceil(66.7 sqrt(k_k_max sigma)))) along k_kappa; the positive-definite window
applies only when the reach does not converge (window "auto",
sgl_full.py:83-94, 328-333); the normalisation grid has 4000 points.

Pinned by the stored pdf_zs0.5_vp.json / pdf_zs1_vp.json (P_S to 1e-9 of
the peak, eta, target_kappa, target_gamma, cell and group counts, the grid)
and the closed-form checks of tests/test_oracle_posref.py.
"""
from __future__ import annotations

import math

import numpy as np
from numpy import pi, sqrt
from scipy.optimize import brentq

from . import chain, closure, ref2d, xilin

NSTATE = 48
N_XI_WIDE_VP = 4000                    # pipeline.py:236
XI_COARSE_G_MAX, XI_COARSE_K_MAX = 401, 4001


def los_power(cos, kp, kpar_max=3.0, n=30001):
    """The linear P0 table of los_rho on (kp, kpar = linspace(0, kpar_max, n))."""
    kpar = np.linspace(0.0, kpar_max, n)
    return kpar, cos.P0(np.sqrt(kp[:, None] ** 2 + kpar[None, :] ** 2))


def los_rho(kpar, P, chi1, d1, chi2, d2):
    """pos_closure.los_rho (pos_closure.py:372-385): correlation of the top-hat
    shell averages of the linear field at transverse wavenumbers kp, with P
    the P0 table of los_power."""
    s1, s2 = np.sinc(kpar * d1 / (2 * pi)), np.sinc(kpar * d2 / (2 * pi))
    c12 = np.trapezoid(P * (s1 * s2 * np.cos(kpar * (chi2 - chi1)))[None, :], kpar, axis=1)
    return c12 / np.sqrt(np.trapezoid(P * (s1 * s1)[None, :], kpar, axis=1)
                         * np.trapezoid(P * (s2 * s2)[None, :], kpar, axis=1))


def population(pop, cos):
    """pos_closure.population (pos_closure.py:388-449) from an oracle build
    made with xi=True, pos=True."""
    ekc, egc = xilin.coarse_edges()
    kmc, gmc = xilin.coarse_centres()
    ngc = egc.size - 1
    g0 = {}
    for g in pop.xi_groups:
        if g["m"] == 0:
            if g["shell"] in g0:
                raise RuntimeError("one m = 0 Limber node per shell")
            g0[g["shell"]] = g
    shells = sorted(set(pop.pieces) & set(g0))
    flatW = pop.H0
    with np.errstate(divide="ignore", invalid="ignore"):
        flatk = np.where(flatW > 0, pop.H0k / np.where(flatW > 0, flatW, 1.0), 0.0)
        flatg = np.where(flatW > 0, pop.H0g / np.where(flatW > 0, flatW, 1.0), 0.0)
    cells = []
    for s in shells:
        p, g = pop.pieces[s], g0[s]
        W = p["h0"]
        with np.errstate(divide="ignore", invalid="ignore"):
            a = np.searchsorted(ekc, p["h0k"] / W, side="right") - 1
            b = np.searchsorted(egc, p["h0g"] / W, side="right") - 1
        ok = (a >= 0) & (a < ekc.size - 1) & (b >= 0) & (b < ngc)
        cidx = np.where(ok, a * ngc + b, -1)
        mr = g["rows"][:, :-1] @ kmc[g["idx"] // ngc] + g["rows"][:, -1]
        r0 = (mr / np.linalg.norm(mr)) @ g["rows"][:, :-1]
        sgn = 1.0 if r0 @ g["W"] >= 0 else -1.0
        with np.errstate(divide="ignore", invalid="ignore"):
            phic = np.where(g["W"] > 0, sgn * r0 / g["W"], 0.0)
        lut = np.zeros(ekc.size * ngc + 1)
        lut[g["idx"]] = phic
        phi = np.where(cidx >= 0, lut[np.where(cidx >= 0, cidx, ekc.size * ngc)], 0.0)
        cells.append(dict(W=W, kap=flatk[p["idx"]], gam=flatg[p["idx"]], phi=phi))
    kp, _ = xilin.kperp_grid(cos)
    chi = np.array([pop.pieces[s]["node"][1] for s in shells])
    dchi = np.array([pop.pieces[s]["node"][2] for s in shells])
    v0 = [np.abs(g0[s]["v0"]) for s in shells]
    rho = np.zeros(max(len(shells) - 1, 0))
    kpar, P = los_power(cos, kp)
    for i in range(rho.size):
        rj = los_rho(kpar, P, chi[i], dchi[i], chi[i + 1], dchi[i + 1])
        v1, v2 = v0[i][:kp.size], v0[i + 1][:kp.size]
        rho[i] = float(np.sum(v1 * v2 * rj) / np.sqrt(np.sum(v1 ** 2) * np.sum(v2 ** 2)))
    rbar = np.array([np.mean([rho[j] for j in (s - 1, s) if 0 <= j < rho.size]) if rho.size else 0.0
                     for s in range(len(cells))])
    psi = xilin.Psi2(pop.xi_groups)
    vg = 0.0
    for g in psi.xg:
        if g["m"] == 1:
            a_ = g["rows"] @ gmc[g["idx_q"][g["qinv"]]] + g["res"]
            vg += float(a_ @ a_)
    return dict(cells=cells, shells=shells, rho=rho, fvar=(1 - rbar) / (1 + rbar),
                target_kappa=psi.var(), target_gamma=vg, psi=psi)


def vp_rates(v, sgn, u, w):
    """vp_rates with the loading's sign (pos_closure.py:273-308): the rates of
    -phi are those of +phi on the mirrored (symmetric) nodes."""
    lam = closure.vp_rates(v, u, w)
    neg = np.asarray(sgn) < 0
    lam[neg] = lam[neg][:, ::-1]
    return lam


def rate_intervals(d, u):
    """rate_intervals (pos_closure_fused.py:18-49): D = rates - 1 (nstate, G)
    as c + a (u - anchor) + b on the active states lo..hi and c elsewhere;
    the reconstruction is checked to 2e-12 as the reference does."""
    d = np.asarray(d, float)
    rate = d + 1.0
    active = rate > 0
    count = active.sum(axis=0)
    if np.any(count == 0):
        raise ValueError("VP rate has no active state")
    lo = active.argmax(axis=0).astype(np.int64)
    hi = (len(u) - 1 - active[::-1].argmax(axis=0)).astype(np.int64)
    if np.any(hi - lo + 1 != count):
        raise ValueError("VP active states must form an interval")
    all_active = count == len(u)
    c = np.where(all_active, 0.0, -1.0)
    col = np.arange(d.shape[1])
    first = np.where(all_active, d[lo, col], rate[lo, col])
    last = np.where(all_active, d[hi, col], rate[hi, col])
    span = u[hi] - u[lo]
    a = np.divide(last - first, span, out=np.zeros_like(first), where=span != 0)
    b = np.where(all_active, 0.5 * (first + last - a * (u[lo] + u[hi])),
                 np.where(a >= 0, first, last))
    anchor = np.where(all_active, 0.0, np.where(a >= 0, u[lo], u[hi]))
    if np.any((~all_active) & (lo != 0) & (hi != len(u) - 1)):
        raise ValueError("clipped VP support must reach one end of the state grid")
    model = c[None, :] + np.where(active, (u[:, None] - anchor) * a + b, 0.0)
    scaled = float(np.max(np.abs(model - d) / np.maximum(1.0, np.abs(d))))
    if scaled > 2e-12:
        raise ValueError(f"rate reconstruction exceeds accuracy bound: {scaled}")
    return a, b, c, lo, hi


class Closure:
    """pos_closure.Closure with exact per-group vp rates (pos_closure.py:453-523)
    and the fused evaluation (pos_closure_fused.py:107-140)."""

    def __init__(self, pp, nstate=NSTATE, eta=None):
        self.pop = pp
        self.u, self.w = closure.gh_nodes(nstate)
        self.K = [closure.transition(self.u, self.w, float(r)) for r in pp["rho"]]
        self.shear_var = float(pp["target_gamma"])
        cells = pp["cells"]
        self.grp = []
        for cl in cells:
            ph, gidx = np.unique(cl["phi"], return_inverse=True)
            self.grp.append((ph, gidx.astype(np.int64),
                             np.bincount(gidx, cl["W"] * cl["kap"], minlength=ph.size)))
        self.eta = self.calibrate(pp["target_kappa"]) if eta is None else float(eta)
        kap = np.concatenate([cl["kap"] for cl in cells])
        gam = np.concatenate([cl["gam"] for cl in cells])
        uu, inv = np.unique(np.stack([kap, gam], 1), axis=0, return_inverse=True)
        inv = inv.ravel()
        self.ku, self.gu = uu[:, 0].copy(), uu[:, 1].copy()
        off = np.r_[0, np.cumsum([len(cl["W"]) for cl in cells])]
        cu, cw, gst, shg, ivs = [], [], [np.zeros(1, np.int64)], [0], []
        c_off = 0
        for s, cl in enumerate(cells):
            ph, gidx, _ = self.grp[s]
            D = vp_rates(self.eta * pp["fvar"][s] * ph ** 2, np.sign(ph), self.u, self.w) - 1.0
            uidx = inv[off[s]:off[s + 1]].astype(np.int64)
            o = np.argsort(gidx, kind="stable")
            gstart = np.searchsorted(gidx[o], np.arange(ph.size + 1)).astype(np.int64)
            cu.append(uidx[o])
            cw.append(np.ascontiguousarray(cl["W"][o]))
            gst.append(gstart[1:] + c_off)
            c_off += uidx.size
            shg.append(shg[-1] + ph.size)
            ivs.append(rate_intervals(D.T, self.u))
        self.cu = np.ascontiguousarray(np.concatenate(cu), dtype=np.int64)
        self.cw = np.ascontiguousarray(np.concatenate(cw))
        self.gstart = np.ascontiguousarray(np.concatenate(gst), dtype=np.int64)
        self.sh_g = np.array(shg, np.int64)
        self.aa, self.bb, self.cc, self.lo, self.hi = [
            np.ascontiguousarray(np.concatenate([v[j] for v in ivs])) for j in range(5)]
        self.lo = self.lo.astype(np.int64)
        self.hi = self.hi.astype(np.int64)
        self.Karr = (np.ascontiguousarray(np.array(self.K)) if self.K
                     else np.zeros((1, nstate, nstate)))
        self.n_cells = int(self.cu.size)
        self.n_groups = int(self.sh_g[-1])

    def variance(self, eta):
        """The model's kappa clustering variance (pos_closure.py:560-573)."""
        ms = [vp_rates(eta * f * ph ** 2, np.sign(ph), self.u, self.w).T @ wk
              for (ph, _, wk), f in zip(self.grp, self.pop["fvar"])]
        return closure.chain_variance(ms, self.w, self.K)

    def calibrate(self, target, lo=1e-3, hi=1e4):
        """pos_closure.py:575-591."""
        fun = lambda le: self.variance(np.exp(le)) - target
        a, b = -np.log(1.5), np.log(1.5)
        fa, fb = fun(a), fun(b)
        while fa * fb > 0:
            if fa > 0 and a > np.log(lo):
                b, fb = a, fa
                a -= np.log(1.5)
                fa = fun(a)
            elif fb < 0 and b < np.log(hi):
                a, fa = b, fb
                b += np.log(1.5)
                fb = fun(b)
            else:
                raise ValueError(f"positive closure: kappa clustering target {target:.3e} "
                                 f"outside [{lo}, {hi}]")
        return float(np.exp(brentq(fun, a, b, xtol=2e-12)))

    def lnR(self, kv, qv):
        """ln R on the tensor grid kv x qv (oracle_vp_lnr)."""
        kv = np.ascontiguousarray(np.atleast_1d(kv), dtype=float)
        qv = np.ascontiguousarray(np.atleast_1d(qv), dtype=float)
        ore = np.empty(kv.size * qv.size)
        oim = np.empty_like(ore)
        L = chain.lib()
        p64 = lambda a: a.ctypes.data_as(chain.ctypes.POINTER(chain.ctypes.c_int64))
        L.oracle_vp_lnr(len(self.sh_g) - 1, p64(self.sh_g), p64(self.gstart), p64(self.cu),
                        chain._p(self.cw), chain._p(self.aa), chain._p(self.bb), chain._p(self.cc),
                        p64(self.lo), p64(self.hi), self.u.size, chain._p(self.u),
                        chain._p(self.Karr), chain._p(self.w), self.ku.size, chain._p(self.ku),
                        chain._p(self.gu), kv.size, chain._p(kv), qv.size, chain._p(qv),
                        chain._p(ore), chain._p(oim))
        return (ore + 1j * oim).reshape(kv.size, qv.size)

    def qxi(self, kk, kg):
        """pos_closure.py:690-705: ln R with Im unwrapped, minus the shear term."""
        kk, kg = np.atleast_1d(np.asarray(kk, float)), np.atleast_1d(np.asarray(kg, float))
        v = self.lnR(kk, kg)
        im = v.imag.copy()
        im[:, 0] = np.unwrap(im[:, 0])
        im = np.unwrap(im, axis=1)
        return v.real + 1j * im - 0.25 * self.shear_var * kg[None, :] ** 2


def window_pd(k_k, k_g):
    """_window_pd (sgl_full.py:83-94): cubic B-spline in k_kappa x Wendland
    phi_{3,1} in k_gamma on the box."""
    K, Q = float(k_k[-1]), float(k_g[-1])
    x = np.abs(k_k) * 2.0 / K
    wk = np.where(x < 1, 2 / 3 - x ** 2 + x ** 3 / 2, np.where(x < 2, (2 - x) ** 3 / 6, 0.0)) / (2 / 3)
    r = np.clip(np.abs(k_g) / Q, 0.0, 1.0)
    return wk, (1 - r) ** 4 * (4 * r + 1)


def coarse_sizes(grid, sigma, n0=xilin.XI_COARSE):
    """The closure's coarse grid sizes in invert (sgl_full.py:226-238)."""
    kks, kgs = grid["k_k_max"] * sigma, grid["k_g_max"] * sigma
    n = int(min(XI_COARSE_G_MAX, max(n0, 1 + math.ceil(12.65 * math.sqrt(kgs)))))
    nk = int(min(XI_COARSE_K_MAX, max(n, 1 + math.ceil(66.7 * math.sqrt(kks)))))
    return n, nk


def invert_vp(lam_halo, clo, sigma, xi_out, tol=ref2d.PHI_TOL):
    """sgl_full.invert (sgl_full.py:159-375) with the positive closure: the
    reach on Lambda + qxi evaluated directly, the coarse grid of the closure
    sized by the box, the window only when the reach did not converge, the
    normalisation on 4000 points."""
    direct = lambda a, b: lam_halo(a, b) + clo.qxi(a, b)
    grid = ref2d.grid_plan(direct, sigma, tol)
    n, nk = coarse_sizes(grid, sigma)
    grid.update(xi_coarse=float(n), xi_coarse_k=float(nk))
    k_k = np.linspace(0.0, grid["k_k_max"], grid["nk_k"])
    k_g = np.linspace(0.0, grid["k_g_max"], grid["nk_g"])
    kap = np.linspace(grid["kap_min"], grid["kap_max"], grid["n_kap"])
    gam = np.linspace(0.0, grid["gam_max"], grid["n_gam"])
    use_win = max(grid["phi_at_k_cut"], grid["phi_at_g_cut"]) > tol
    grid["window"] = float(use_win)
    Phi = np.exp(lam_halo(k_k, k_g) + xilin.spline_on_coarse(clo.qxi, k_k, k_g, n, nk))
    if use_win:
        wk, wg = window_pd(k_k, k_g)
        Phi *= wk[:, None] * wg[None, :]
    P3 = ref2d.p3_joint(k_k, k_g, Phi, kap, gam)
    xw = np.linspace(grid["xi_min"], grid["xi_max"], N_XI_WIDE_VP)
    pw = ref2d.arc_integral(kap, gam, P3, grid["g_def"], xw)
    norm = np.trapezoid(pw, xw)
    ps = pw / norm
    mean = float(np.trapezoid(xw * ps, xw))
    var = float(np.trapezoid((xw - mean) ** 2 * ps, xw))
    skew = float(np.trapezoid((xw - mean) ** 3 * ps, xw) / var ** 1.5)
    x = np.exp(-0.5 * xw)
    md = float(np.trapezoid(x * ps, xw))
    sdl = float(sqrt(np.trapezoid((x - md) ** 2 * ps, xw)) / md)
    P = ref2d.arc_integral(kap, gam, P3, grid["g_def"], np.asarray(xi_out, float)) / norm
    return ref2d.Result(np.asarray(xi_out, float), P, sigma, mean, var, skew, sdl, float(norm),
                        grid)


def run(zs, cosmo: dict = None, xi_out=None):
    """pipeline.run(zs, "full", xi_closure="vp") (pipeline.py:114-274)."""
    cos = ref2d.Cosmology(**(cosmo or {}))
    pop = ref2d.build(cos, zs, ell=True, fil=True, sub=True, xi=True, pos=True)
    pp = population(pop, cos)
    clo = Closure(pp)
    sig = float(sqrt(pop.sigma() ** 2 + pp["target_kappa"]))
    xo = (np.concatenate([ref2d.XI_BODY, ref2d.XI_TAIL]) if xi_out is None
          else np.asarray(xi_out, float))
    res = invert_vp(lambda a, b: ref2d.lam(pop, a, b), clo, sig, xo)
    res.pop = pop
    res.closure = clo
    res.pos = pp
    return res
