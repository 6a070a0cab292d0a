"""FP64 CPU restatement of the xi_lin (Limber) two-point term (checker only).

TEST INFRASTRUCTURE: imported only by tests/ (through ``oracle.ref2d``), never
by the product.

The reference's default clustering for the "bias" ingredient
(pipeline.run clustering="xilin", xi_closure="gauss"):

1. ``xi_kperp_grid`` (sgl_full.py:2483-2511): 42 Gauss-Legendre nodes in ln k
   on [1e-4, 1e4] Mpc^-1, (2, 2, 12, 8, 6, 4, 4, 4) per decade, weights
   c_j = w_j k_j^2 P0(k_j) / (2 pi);
2. the deposit (Cells._xi_deposit / _xi_add_grouped, sgl_full.py:1081-1130,
   1397-1441; ``oracle_xi_deposit`` in chain.c): per lens-redshift shell and
   coarse (kappa, gamma) bin (341 x 110, sgl_full.py:457),
   Hx[bin, m, j] = sum beta w J_{2m}(k_j R) over the points in the bin, R the
   comoving distance of the point's lens centre from the beam (halos:
   tangential shear, m = 0..2; filaments: m = 0), and the exact first moments
   lx[0] = sum beta w kappa J0, lx[1] = sum beta w gamma J2; plus the
   analytic shear tail beyond each halo's outer radius (xi_add_shear_tail,
   sgl_full.py:1466-1476);
3. ``xi_end`` (sgl_full.py:1478-1523): per shell and m, A = [Hx; lres]
   diag(sqrt(c / dchi)) with lres the exact first moment minus the lattice
   centres' share, compressed by the eigendecomposition of A^T A (singular
   values above 1e-6 of the largest kept);
4. ``Lam`` (sgl_full.py:2801-2860): rows below 1e-7 of the largest row norm
   pruned, and ``_Qxi_gauss`` (sgl_full.py:2869-2951):
   Psi_2 = sum_groups sign_m / 2 sum_r G_r^2, G_r = sum_b rows_rb f_m(b) +
   res_r f_res (f_0 = e^{i kk kap_b} J0(kg gam_b) - 1, f_res = i kk;
   f_1 = e^{i kk kap_b} J1, f_res = kg / 2; f_2 = e^{i kk kap_b} J2;
   sign 1, -2, +2), direct on the reach probes, on the full grid through the
   101 x 101 coarse grid in s = sqrt(k / k_max) and scipy's not-a-knot
   bicubic spline (``_Q_interpolated``, sgl_full.py:2953-3022);
5. ``Lam.sigma`` with the clustering variance (sgl_full.py:3437-3452).

Pinned by the explorer's "+ clustering" ladder row and the stored
pdf_zs*.json fixtures (tests/test_oracle_xilin.py).
"""
from __future__ import annotations

import numpy as np
from numpy import log, pi, sqrt
from scipy import special as sp

from . import chain

XI_PER_DECADE = (2, 2, 12, 8, 6, 4, 4, 4)          # sgl_full.py:2480
XI_KMIN, XI_KMAX, XI_MMAX = 1e-4, 1e4, 2
XI_SVD_TOL = 1e-6                                   # sgl_full.py:2436
XI_PRUNE = 1e-7                                     # Lam xi_prune default
XI_COARSE = 101                                     # Lam xi_coarse default
COARSE_K = (1e-7, 30.0, 170)                        # sgl_full.py:457
COARSE_G = (1e-9, 30.0, 110)


def kperp_grid(cos):
    """xi_kperp_grid (sgl_full.py:2483-2511)."""
    nd = int(np.ceil(np.log10(XI_KMAX / XI_KMIN)))
    lk, wk = [], []
    for d in range(nd):
        x, w = np.polynomial.legendre.leggauss(XI_PER_DECADE[d])
        a = log(XI_KMIN) + d * log(10.0)
        b = min(a + log(10.0), log(XI_KMAX))
        lk.append(0.5 * (a + b) + 0.5 * (b - a) * x)
        wk.append(0.5 * (b - a) * w)
    kp, wl = np.exp(np.concatenate(lk)), np.concatenate(wk)
    return kp, wl * kp ** 2 * cos.P0(kp) / (2 * pi)


def coarse_edges():
    pos = np.geomspace(COARSE_K[0], COARSE_K[1], COARSE_K[2] + 1)
    ekc = np.concatenate([-pos[::-1], pos])
    egc = np.geomspace(COARSE_G[0], COARSE_G[1], COARSE_G[2] + 1)
    return ekc, egc


def coarse_centres():
    ekc, egc = coarse_edges()
    kmc = np.where(ekc[:-1] * ekc[1:] > 0, np.sign(ekc[1:]) * sqrt(np.abs(ekc[:-1] * ekc[1:])), 0.0)
    gmc = sqrt(egc[:-1] * egc[1:])
    return kmc, gmc


class Shell:
    """The deposit of one lens-redshift shell (xi_begin .. xi_end)."""

    def __init__(self, kp, c, dchi, pos=False):
        self.kp, self.c, self.dchi = kp, c, float(dchi)
        self.ekc, self.egc = coarse_edges()
        self.ngc = self.egc.size - 1
        self.Hx = np.zeros(((self.ekc.size - 1) * self.ngc, 3, kp.size))
        self.lx = np.zeros((2, kp.size))
        # positive closure: lens rate per coarse bin (Cells.xi_begin, sgl_full.py:1394-1395)
        self.Wb = np.zeros((self.ekc.size - 1) * self.ngc) if pos else None

    def deposit(self, K, G, W, R, beta, tang):
        K, G, W, R = (np.ascontiguousarray(a, dtype=float).ravel() for a in (K, G, W, R))
        G = np.clip(G, 1.0000001e-9, None)
        if self.Wb is not None:
            chain.lib().oracle_xi_rate(K.size, chain._p(K), chain._p(G), chain._p(W),
                                       chain._p(self.ekc), self.ekc.size, chain._p(self.egc),
                                       self.egc.size, chain._p(self.Wb))
        chain.lib().oracle_xi_deposit(K.size, chain._p(K), chain._p(G), chain._p(W), chain._p(R),
                                      float(beta), int(bool(tang)), self.kp.size,
                                      chain._p(self.kp), chain._p(self.ekc), self.ekc.size,
                                      chain._p(self.egc), self.egc.size, chain._p(self.Hx),
                                      chain._p(self.lx))

    def shear_tail(self, beta, Gam, Rm, n_area):
        """xi_add_shear_tail (sgl_full.py:1466-1476)."""
        a = self.kp * Rm
        self.lx[1] += beta * n_area * 2 * pi * Gam * sp.j1(a) / a

    def end(self, shell, compress=True):
        """xi_end (sgl_full.py:1478-1523): the compressed rows per m.
        compress=False keeps A^T itself (one row per k_perp node): the same
        quadratic form without the dropped modes below 1e-6 of the largest."""
        kmc, gmc = coarse_centres()
        ngc, nkp = self.ngc, self.kp.size
        sw = sqrt(self.c / self.dchi)
        out = []
        for m in range(3):
            H = self.Hx[:, m, :]
            idx = np.flatnonzero(np.any(H != 0, axis=1))
            if idx.size == 0:
                if m > 1 or not np.any(self.lx[m] != 0):
                    continue
                idx = np.array([0])
            Hm = H[idx]
            if m == 0:
                lres = self.lx[0] - (kmc[idx // ngc][:, None] * Hm).sum(axis=0)
            elif m == 1:
                lres = self.lx[1] - (gmc[idx % ngc][:, None] * Hm).sum(axis=0)
            else:
                lres = np.zeros(nkp)
            A = np.vstack([Hm, lres[None, :]]) * sw[None, :]
            if not compress:
                out.append(dict(shell=shell, m=m, idx=idx, rows=np.ascontiguousarray(A.T)))
                continue
            ev, V = np.linalg.eigh(A.T @ A)
            S = sqrt(np.maximum(ev[::-1], 0.0))
            Vd = V[:, ::-1]
            keep = S > XI_SVD_TOL * S[0] if S[0] > 0 else np.zeros(S.size, bool)
            if not keep.any():
                continue
            rows = (A @ Vd[:, keep]).T
            grp = dict(shell=shell, m=m, idx=idx, rows=rows)
            if m == 0 and self.Wb is not None:
                # positive closure: bin rates and the leading k_perp profile
                # (sgl_full.py:1520-1521)
                grp.update(W=self.Wb[idx].copy(), v0=Vd[:, 0].copy())
            out.append(grp)
        return out


def spline_on_coarse(fn, kk, kg, n, nk):
    """Lam._Q_interpolated (sgl_full.py:2953-3022) of fn: fn on the coarse grid
    k_kappa = kk_max s_k^2 (nk nodes) x k_gamma = kg_max s^2 (n nodes), s
    uniform on [0, 1], then scipy's not-a-knot cubic spline in s_k along
    k_kappa and in s along k_gamma, evaluated on the full grid."""
    from scipy.interpolate import BSpline, make_interp_spline
    kk, kg = np.asarray(kk, float), np.asarray(kg, float)
    s = np.linspace(0.0, 1.0, n)
    s_k = s if nk == n else np.linspace(0.0, 1.0, nk)
    Qc = fn(float(kk[-1]) * s_k ** 2, float(kg[-1]) * s ** 2)
    sk = sqrt(np.clip(kk / kk[-1], 0.0, 1.0))
    sg = sqrt(np.clip(kg / kg[-1], 0.0, 1.0))
    out = np.empty((kk.size, kg.size), complex)
    for part, setter in ((Qc.real, "real"), (Qc.imag, "imag")):
        sp0 = make_interp_spline(s_k, part, k=3, axis=0)
        sp1 = make_interp_spline(s, sp0.c, k=3, axis=1)
        val = BSpline.design_matrix(sk, sp0.t, 3) @ (BSpline.design_matrix(sg, sp1.t, 3) @ sp1.c).T
        setattr(out, setter, val)
    return out


class Psi2:
    """Lam's xi_lin term over the groups of every shell (sgl_full.py:2801-2860)."""

    def __init__(self, groups, prune=XI_PRUNE):
        kmc, gmc = coarse_centres()
        self.kmc, self.gmc = kmc, gmc
        ngc = gmc.size
        if prune:
            s0 = max(float(np.max(np.linalg.norm(g["rows"], axis=1))) for g in groups)
            groups = [{**g, "rows": g["rows"][np.linalg.norm(g["rows"], axis=1) > prune * s0]}
                      for g in groups]
            groups = [g for g in groups if g["rows"].shape[0]]
        self.xg = []
        w0, r0, s0_, stk = [], [], [], {0: ([], [], []), 1: ([], [], []), 2: ([], [], [])}
        for g in groups:
            up, pinv = np.unique(g["idx"] // ngc, return_inverse=True)
            uq, qinv = np.unique(g["idx"] % ngc, return_inverse=True)
            m = int(g["m"])
            rows, res = g["rows"][:, :-1], g["rows"][:, -1]
            srow = rows.sum(axis=1)
            self.xg.append(dict(m=m, sign=1.0 if m == 0 else 2.0 * (-1) ** m, pinv=pinv, qinv=qinv,
                                idx_p=up, idx_q=uq, rows=rows, res=res, sum_rows=srow,
                                kc=kmc[g["idx"] // ngc]))
            if m == 0:
                Wg = np.zeros((rows.shape[0], kmc.size))
                np.add.at(Wg.T, g["idx"] // ngc, rows.T)
                w0.append(Wg)
                r0.append(res)
                s0_.append(srow)
            Ug = np.zeros((rows.shape[0], gmc.size))
            np.add.at(Ug.T, g["idx"] % ngc, rows.T)
            stk[m][0].append(Ug)
            stk[m][1].append(res)
            stk[m][2].append(srow)
        self.W0 = np.vstack(w0)
        self.res0 = np.concatenate(r0)
        self.sum0 = np.concatenate(s0_)
        self.stack = {m: dict(U=np.vstack(v[0]), res=np.concatenate(v[1]), sum=np.concatenate(v[2]),
                              sign=1.0 if m == 0 else 2.0 * (-1) ** m)
                      for m, v in stk.items() if v[0]}

    def var(self):
        """The clustering variance of Lam.sigma (sgl_full.py:3440-3445)."""
        return float(np.sum((self.W0 @ self.kmc + self.res0) ** 2))

    def direct(self, kk, kg):
        """_Qxi_gauss (sgl_full.py:2869-2951)."""
        kk, kg = np.atleast_1d(np.asarray(kk, float)), np.atleast_1d(np.asarray(kg, float))
        if kg.size == 1 and kg[0] == 0.0:
            E = np.exp(1j * np.outer(kk, self.kmc))
            G = E @ self.W0.T - self.sum0[None, :] + 1j * kk[:, None] * self.res0[None, :]
            return (0.5 * np.sum(G * G, axis=1))[:, None]
        if kk.size == 1 and kk[0] == 0.0:
            out = np.zeros((1, kg.size), complex)
            xg = np.outer(self.gmc, kg)
            for m, d in self.stack.items():
                G = d["U"] @ (sp.j0(xg) if m == 0 else sp.j1(xg) if m == 1 else sp.jv(m, xg))
                if m == 0:
                    G -= d["sum"][:, None]
                elif m == 1:
                    G += 0.5 * kg[None, :] * d["res"][:, None]
                out[0] += 0.5 * d["sign"] * np.sum(G * G, axis=0)
            return out
        ph = np.outer(kk, self.kmc)
        Ecat = np.concatenate([np.cos(ph), np.sin(ph)], axis=0)
        xg = np.outer(self.gmc, kg)
        J = {0: sp.j0(xg), 1: sp.j1(xg), 2: sp.jv(2, xg)}
        out = np.zeros((kk.size, kg.size), complex)
        for g in self.xg:
            r, m = g["rows"].shape[0], g["m"]
            nP, nQ = g["idx_p"].size, g["idx_q"].size
            Hd = np.zeros((r, nP, nQ))
            Hd[:, g["pinv"], g["qinv"]] = g["rows"]
            T = (Hd.reshape(r * nP, nQ) @ J[m][g["idx_q"], :]).reshape(r, nP, kg.size)
            T2 = np.ascontiguousarray(T.transpose(1, 0, 2)).reshape(nP, r * kg.size)
            Gc = (Ecat[:, g["idx_p"]] @ T2).reshape(2, kk.size, r, kg.size)
            Gr, Gi = Gc[0].transpose(1, 0, 2), Gc[1].transpose(1, 0, 2)
            if m == 0:
                Gr = Gr - g["sum_rows"][:, None, None]
                Gi = Gi + kk[None, :, None] * g["res"][:, None, None]
            elif m == 1:
                Gr = Gr + 0.5 * kg[None, None, :] * g["res"][:, None, None]
            out += 0.5 * g["sign"] * (np.sum(Gr ** 2 - Gi ** 2, axis=0)
                                      + 2j * np.sum(Gr * Gi, axis=0))
        return out

    def coarse(self, kk_max, kg_max, n=XI_COARSE):
        """Q on the coarse grid k = k_max s^2, s = linspace(0, 1, n)."""
        s = np.linspace(0.0, 1.0, n)
        return s, self.direct(kk_max * s ** 2, kg_max * s ** 2)

    def interpolated(self, kk, kg, n=XI_COARSE):
        """_Q_interpolated (sgl_full.py:2953-3022): the coarse grid and scipy's
        not-a-knot bicubic spline in s = sqrt(k / k_max)."""
        return spline_on_coarse(self.direct, kk, kg, n, n)

    def __call__(self, kk, kg):
        """Lam._direct_call's xi term (sgl_full.py:3114-3128): interpolated on a
        full grid (more than 101^2 points, both axes), direct otherwise."""
        kk, kg = np.atleast_1d(np.asarray(kk, float)), np.atleast_1d(np.asarray(kg, float))
        if (kk.size * kg.size > XI_COARSE ** 2 and kk.size > 1 and kg.size > 1 and kk[-1] > 0
                and kg[-1] > 0):
            return self.interpolated(kk, kg)
        return self.direct(kk, kg)
