"""FP64 CPU restatement of the subhalo ingredient (checker only).

TEST INFRASTRUCTURE: imported only by tests/ (through ``oracle.ref2d``), never
by the product.

The reference's "sub" ingredient in the configuration ``pipeline.run`` builds
by default (pipeline.py:171-183: sector="exact", exact_kw method="moments",
clump_edge = the halo edge "r200", clump_mass "vir" (unused with r200),
clump_shape "circular"):

1. ``SubhaloModel.fs`` / ``host`` (subhalos.py:672-772): for every (shell, mass)
   host with M >= 1e9 Msun, the JvdB14 subhalo mass fraction f_s (formation
   redshift from delta_c(z_f) = delta_c(z) + w_f dsigma, Brent root), the
   clump mass function dN/dlnpsi = g psi^alpha e^{-beta psi^omega} on 28 log
   bins of psi = m / M_vir in [1e7 Msun / M_vir, 1], clumps as NFW halos of
   M200 = m truncated at their own r200 (reach D = r200), the clump number
   selected to the host's r200 sphere (``_frac_inside``) and projected over
   it (``_n3_proj_nb``: the Green+21 radial profile, a 256 x 160 trapezoid);
2. ``ExactSector._clump_cells`` (subhalos_exact.py:1553-1622,
   ``_clump_cells_kernel`` 1234-1304): clump cells (mass bin i, ray distance
   cell d of the 90-cell ``_dgrid``) with kappa_c, gamma_c, mass, and rate
   prefactor dN_i dA_d;
3. ``host_cells_moments`` (subhalos_exact.py:1647-1811): on 128 coarse host
   radii, the ring-averaged rates (``oracle_sub_ring``), the rare-mark cut,
   the bulk cumulants of K = sum kappa_c and S = sum m_c, the 3-node Gauss
   rules in K and in S | K (regression on K with a cancellation-free
   residual variance), the conditional shear regression, three log bins of
   rare strong clumps (two points each), i.e. 9 + 2 nbr points per coarse
   radius with K offset, removed mass, weight and the two shear terms; the
   host-mass table of the rebuilt host (17 log nodes on [M - max S, M]);
4. the emission (subhalos_exact.py:802-925, ``oracle_sub_emit``): every host
   cell within reach of a clump becomes 9 + 2 nbr points on the host rebuilt
   at M - S;
5. the decoration step of ``_truncated_host`` (sgl_full.py:2088-2127).

Pinned by the explorer notebook's "+ subhalos" ingredient-ladder row at
z_s = 1 (tests/test_oracle_ref2d.py) and by the properties in
tests/test_oracle_sub.py.
"""
from __future__ import annotations

import math

import numpy as np
from numpy import exp, log, pi, sqrt
from scipy.optimize import brentq
from scipy.special import gamma as Gfn
from scipy.special import gammaincc

from . import chain
from .cosmo import DC0

ALPHA, BETA, OMEGA, PSI_RES, PSI_MAX = -0.82, 50.0, 4.0, 1e-4, 1.0   # subhalos.py:353
X0_RVIR, BIAS_EXP = 0.86, 2.5                                       # subhalos.py:354
M_FLOOR = 1e7                                                       # subhalos.py:355
MHOST_MIN, NPSI, ND, NTH = 1e9, 28, 90, 48                          # subhalos.py:541
EPS_RARE, NBR, NR_COARSE, NM_TAB = 3e-3, 3, 128, 17                 # subhalos_exact.py:1534-1536
G3_TOL_T, G3_TOL_W = 1e-10, 1e-12                                   # subhalos_exact.py:70-71
MT_WFLOOR, KS_RIDGE = 1e-12, 1e-10                                  # subhalos_exact.py:73-74
MMIN = 1e7                                                          # engine mass grid node 0

S_EXP = (1 + ALPHA) / OMEGA                                         # subhalos.py:656-659
WF = sqrt(2 * log(0.815 * exp(-0.25) / 0.5 ** 0.707 + 1))
GDEN = (gammaincc(S_EXP, BETA * PSI_RES ** OMEGA) - gammaincc(S_EXP, BETA)) * Gfn(S_EXP)
TH = (np.arange(NTH) + 0.5) * pi / NTH
CTH = np.cos(TH)


def _mu(y):
    return log(1 + y) - y / (1 + y)


def n3_proj(xe, C, x0):
    """subhalos.py:430-456: line-of-sight projection of rho_NFW B(x) on a
    256 x 160 trapezoid; returns (Rh, n3)."""
    nR, nu = 256, 160
    Rh = (np.arange(nR) + 0.5) / nR * xe
    um = sqrt(np.maximum(xe * xe - Rh * Rh, 0.0))
    u = np.arange(nu) / (nu - 1)
    X = sqrt(Rh[:, None] ** 2 + (u[None, :] * um[:, None]) ** 2)
    q = X / x0
    B = 1.0 / sqrt(1.0 / (q * q * sqrt(q)) + 1.0)
    p = B / (X * (1.0 + C * X) ** 2)
    n3 = np.sum(0.5 * (p[:, 1:] + p[:, :-1]), axis=1) / (nu - 1) * um * 2.0
    return Rh, n3


def frac_inside(xa, xb, C, x0, n=4096):
    """subhalos.py:459-467: fraction of dN/dx ~ x B(x)/(1+cx)^2 on [0, xb] inside xa."""
    def mass(xm):
        x = (np.arange(n) + 0.5) * (xm / n)
        B = 1.0 / sqrt((x / x0) ** -BIAS_EXP + 1.0)
        return np.sum(x * B / (1.0 + C * x) ** 2) * (xm / n)
    return mass(xa) / mass(xb)


def std_moments(k1, k2, k3, k4, k5):
    """subhalos_exact.py:571-583: standardised raw moments tau_0..tau_5 of a
    law with cumulants k1..k5."""
    sd = sqrt(np.maximum(k2, 0.0))
    ss = np.where(sd > 0, sd, 1.0)
    g3, g4, g5 = k3 / ss ** 3, k4 / ss ** 4, k5 / ss ** 5
    tau = np.zeros(k1.shape + (6,))
    tau[..., 0] = 1.0
    tau[..., 2] = 1.0
    tau[..., 3] = g3
    tau[..., 4] = g4 + 3.0
    tau[..., 5] = g5 + 10.0 * g3
    return sd, tau


def gauss3(tau, lo):
    """subhalos_exact.py:587-660 / 663-724 with hi = +inf: the 3-node Gauss
    rule of a standardised law (monic cubic orthogonal to 1, t, t^2 from the
    closed-form 3 x 3 Hankel system, roots by trigonometric Cardano, weights
    by Lagrange), accepted when the roots are distinct by > 1e-7, the weights
    > -1e-12 and the roots >= lo within 1e-10 (1 + max|t|) (then clamped);
    else the 2-node rule exact to the third moment in the 3-node layout
    (low, high, high with weight 0) when inside; else one node at 0."""
    n = tau.shape[0]
    t = np.zeros((n, 3))
    pr = np.zeros((n, 3))
    pr[:, 0] = 1.0
    t3, t4, t5 = tau[:, 3], tau[:, 4], tau[:, 5]
    D = t4 - t3 ** 2 - 1.0
    ok3 = np.zeros(n, bool)
    with np.errstate(all="ignore"):
        good = np.isfinite(D) & (np.abs(D) > 1e-10)
        Ds = np.where(good, D, 1.0)
        al = (t3 * t4 + t3 - t5) / Ds
        be = -t4 - t3 * al
        ga = -t3 - al
        pp = be - al ** 2 / 3.0
        qq = 2.0 * al ** 3 / 27.0 - al * be / 3.0 + ga
        neg = good & (pp < 0)
        rp = sqrt(np.where(neg, -pp / 3.0, 1.0))
        arg = 3.0 * qq / (2.0 * np.where(neg, pp, -1.0) * rp)
        real = neg & (np.abs(arg) <= 1.0 + 1e-12)
        ph = np.arccos(np.clip(arg, -1.0, 1.0)) / 3.0
        r = np.sort(np.stack([2 * rp * np.cos(ph - 2 * pi * j / 3) for j in range(3)], axis=1)
                    - (al / 3.0)[:, None], axis=1)
        r0, r1, r2 = r[:, 0], r[:, 1], r[:, 2]
        dist = real & (r1 - r0 > 1e-7) & (r2 - r1 > 1e-7)
        w0 = (1.0 + r1 * r2) / ((r0 - r1) * (r0 - r2))
        w1 = (1.0 + r0 * r2) / ((r1 - r0) * (r1 - r2))
        w2 = (1.0 + r0 * r1) / ((r2 - r0) * (r2 - r1))
        tt = G3_TOL_T * (1.0 + np.maximum(np.abs(r0), np.abs(r2)))
        fine = dist & (w0 > -G3_TOL_W) & (w1 > -G3_TOL_W) & (w2 > -G3_TOL_W) & (r0 >= lo - tt)
    if fine.any():
        t[fine] = np.maximum(r[fine], lo[fine, None])
        pr[fine] = np.maximum(np.stack([w0, w1, w2], axis=1)[fine], 0.0)
        ok3 = fine
    r2m = ~ok3
    if r2m.any():
        g = t3[r2m]
        rt = sqrt(g * g + 4.0)
        a1, a2 = (g + rt) / 2, (g - rt) / 2
        tol = G3_TOL_T * (1.0 + np.maximum(np.abs(a1), np.abs(a2)))
        ok2 = a2 >= lo[r2m] - tol
        idx = np.nonzero(r2m)[0][ok2]
        t[idx, 0] = np.maximum(a2[ok2], lo[idx])
        t[idx, 1] = np.maximum(a1[ok2], lo[idx])
        t[idx, 2] = t[idx, 1]
        pr[idx, 0] = a1[ok2] / (a1[ok2] - a2[ok2])
        pr[idx, 1] = -a2[ok2] / (a1[ok2] - a2[ok2])
        pr[idx, 2] = 0.0
    return t, pr


def cut_for(nu, k, eps):
    """subhalos_exact.py:78-83: smallest kappa cut whose marks above it have
    total rate <= eps."""
    o = np.argsort(-k, kind="stable")
    cr = np.cumsum(nu[o])
    j = int(np.searchsorted(cr, eps, side="right"))
    return float(k[o[j]]) if j < k.size else 0.0


class SubhaloModel:
    """Hosts of one (cosmology, z_s) on the build's shell nodes zl."""

    def __init__(self, cos, zs, zl, Ml):
        self.cos, self.zs, self.zl, self.Ml = cos, float(zs), np.asarray(zl), np.asarray(Ml)
        self._cache = {}

    def fs(self, M, z):
        """subhalos.py:672-705 (JvdB14 f_s, Brent z_f)."""
        cos = self.cos
        dcz = DC0 / cos.Dg(z)
        ds2 = cos.sigmaM(0.5 * M, 0.0) ** 2 - cos.sigmaM(M, 0.0) ** 2
        if ds2 <= 0:
            return 0.0
        rhs = dcz + WF * sqrt(ds2)
        if DC0 / cos.Dg(30.0) < rhs:
            return 0.0
        zf = brentq(lambda zz: DC0 / cos.Dg(zz) - rhs, z, 30.0, xtol=1e-15, rtol=8.9e-16,
                    maxiter=200)
        zz = z + (np.arange(200) + 0.5) * (zf - z) / 200
        Ntau = np.sum(6.006 * sqrt(cos.Dvir(zz) / 178.0)
                      / (1 + zz)) * (zf - z) / 200
        if Ntau <= 0:
            return 0.0
        f = 0.3563 / Ntau ** 0.6 - 0.075
        return f if 0.0 < f < 0.95 else 0.0

    def host(self, jz, jM):
        """subhalos.py:707-772 with clump_edge "r200", or None."""
        key = (jz, jM)
        if key in self._cache:
            return self._cache[key]
        cos, z, M = self.cos, float(self.zl[jz]), float(self.Ml[jM])
        out = None
        f = self.fs(M, z) if M >= MHOST_MIN else 0.0
        if f > 0:
            C, rs, ks, _ = cos.nfw_params(M, z, self.zs)
            r200 = rs * C
            eta = cos.eta(C, z)
            Mvir = M * _mu(eta * C) / _mu(C)
            g = OMEGA * BETA ** S_EXP / GDEN * f
            psi_min = M_FLOOR / Mvir
            if psi_min < PSI_MAX:
                le = np.linspace(log(psi_min), 0.0, NPSI + 1)
                psi = exp(0.5 * (le[1:] + le[:-1]))
                dN = g * psi ** ALPHA * exp(-BETA * psi ** OMEGA) * np.diff(le)
                m = psi * Mvir
                Cc, rsc, ksc, _ = cos.nfw_params(m, z, self.zs)
                Cc, rsc, ksc = (np.asarray(a, float) for a in (Cc, rsc, ksc))
                xtc = Cc.copy()
                D = xtc * rsc
                x0 = X0_RVIR * eta
                dN = dN * frac_inside(1.0, eta, C, x0)
                Rh, n3 = n3_proj(1.0, C, x0)
                n2 = n3 / np.trapezoid(n3 * 2 * pi * Rh, Rh)
                out = dict(M=M, z=z, r200=r200, eta=eta, vr=1.0, f=f, dN=dN, m=m, rsc=rsc,
                           ksc=ksc, D=D, xtc=xtc, Rn=Rh * r200, n2=n2 / r200 ** 2, rs=rs, ks=ks,
                           C=C, Mvir=Mvir)
        self._cache[key] = out
        return out

    @staticmethod
    def dgrid(h):
        """subhalos.py:798-802."""
        Dm = h["D"].max()
        de = np.concatenate([[0.0], np.geomspace(1e-5 * Dm, Dm, ND)])
        dc = np.where(de[:-1] > 0, sqrt(de[:-1] * np.maximum(de[1:], 0.0)), 0.5 * de[1])
        return de, dc, pi * np.diff(de ** 2)

    def clump_cells(self, h):
        """subhalos_exact.py:1234-1304 (r200-truncated clumps)."""
        if "_cc" in h:
            return h["_cc"]
        _, dc, dA = self.dgrid(h)
        I, Dd, Kc, Gc = [], [], [], []
        for i in range(h["m"].size):
            sel = np.nonzero(dc < h["D"][i])[0]
            if not sel.size:
                continue
            k_, g_ = chain.trunc_kg(dc[sel] / h["rsc"][i], h["ksc"][i], h["xtc"][i])
            I.append(np.full(sel.size, i))
            Dd.append(sel)
            Kc.append(k_)
            Gc.append(g_)
        I, Dd = np.concatenate(I), np.concatenate(Dd)
        cc = dict(i=I, d=Dd, k=np.concatenate(Kc), g=np.concatenate(Gc), m=h["m"][I] / h["vr"],
                  pref=h["dN"][I] * dA[Dd], dc=dc)
        h["_cc"] = cc
        return cc

    def coarse(self, h, xref):
        """subhalos_exact.py:1655-1811: the per-coarse-radius point tables and the
        host-mass table, cached on the host."""
        if "_cm" in h:
            return h["_cm"]
        cc = self.clump_cells(h)
        rs, M = h["rs"], h["M"]
        xpos = xref[xref > 0]
        xg = np.geomspace(xpos.min(), xref.max(), NR_COARSE)
        lxg = log(xg)
        r = np.ascontiguousarray(xg * rs)
        dc = np.ascontiguousarray(cc["dc"])
        nd = dc.size
        ring = np.empty((r.size, nd))
        ecd = np.empty((r.size, nd))
        Rn, n2 = np.ascontiguousarray(h["Rn"]), np.ascontiguousarray(h["n2"])
        chain.lib().oracle_sub_ring(r.size, chain._p(r), nd, chain._p(dc), NTH, chain._p(CTH),
                                    Rn.size, chain._p(Rn), chain._p(n2), chain._p(ring),
                                    chain._p(ecd))
        k, m, g, pref, d, ib = cc["k"], cc["m"], cc["g"], cc["pref"], cc["d"], cc["i"]
        kstar = cut_for(pref * ring.max(axis=0)[d], k, EPS_RARE)
        bulk = k <= kstar
        ms = max(float(np.sum(pref * ring.mean(axis=0)[d] * m)), 1e-300)
        mh = m / ms
        cols = [pref * k ** a for a in range(1, 6)] + [pref * mh ** b for b in range(1, 6)] + \
            [pref * k * mh, pref * g ** 2, pref * k * g ** 2, pref * g, pref * k * g, pref * mh * g]
        Q = np.zeros((nd, len(cols)))
        for j, cv in enumerate(cols):
            Q[:, j] = np.bincount(d[bulk], cv[bulk], nd)
        CQ = ring @ Q
        C = CQ[:, :-3]
        RE = ring * ecd
        E1, KE1, SE1 = RE @ Q[:, -3], RE @ Q[:, -2], RE @ Q[:, -1]
        rare = ~bulk
        nbr = NBR if rare.any() else 0
        n = xg.size
        if nbr:
            kr, dr = k[rare], d[rare]
            be = np.geomspace(kr.min() * (1 - 1e-12), kr.max() * (1 + 1e-12), nbr + 1)
            bi = np.clip(np.searchsorted(be, kr, side="right") - 1, 0, nbr - 1)
            pr_ = pref[rare]
            R = ring[:, dr]                                          # (n, n_rare)
            oh = np.zeros((kr.size, nbr))
            oh[np.arange(kr.size), bi] = 1.0
            nub_g = (R * pr_) @ oh
            k1b_g = (R * (pr_ * kr)) @ oh
            m1b_g = (R * (pr_ * m[rare])) @ oh
            g2b_g = (R * (pr_ * g[rare] ** 2)) @ oh
            e1b_g = ((R * ecd[:, dr]) * (pr_ * g[rare])) @ oh
            kbg = k1b_g / np.where(nub_g > 0, nub_g, 1.0)
            v2b_g = (R * pr_ * (kr[None, :] - kbg[:, bi]) ** 2) @ oh
        c = [None] + [C[:, a] for a in range(5)]
        sS = [None] + [C[:, 5 + b] for b in range(5)]
        c11, E2b = C[:, 10], C[:, 11]
        sdK, tauK = std_moments(*c[1:])
        tK, pK = gauss3(tauK, (0.0 - c[1]) / np.where(sdK > 0, sdK, 1.0))
        Kn = c[1][:, None] + sdK[:, None] * tK
        sdS, tauS = std_moments(*sS[1:])
        tS, pS = gauss3(tauS, (0.0 - sS[1]) / np.where(sdS > 0, sdS, 1.0))
        c2s = np.where(c[2] > 0, c[2], 1.0)
        bK = np.where(c[2] > 0, c11 / c2s, 0.0)
        # det = c2 sS2 - c11^2 without cancellation (subhalos_exact.py:1486-1528)
        mbin = h["m"] / h["vr"] / ms
        nb = mbin.size
        P = np.zeros((nd, nb))
        Pa = np.zeros((nd, nb))
        np.add.at(P, (d[bulk], ib[bulk]), pref[bulk])
        np.add.at(Pa, (d[bulk], ib[bulk]), pref[bulk] * k[bulk])
        n_, a_ = ring @ P, ring @ Pa
        ab = np.where(n_ > 0, a_ / np.where(n_ > 0, n_, 1.0), 0.0)
        Wk = np.sum(ring[:, d[bulk]] * pref[bulk][None, :]
                    * (k[bulk][None, :] - ab[:, ib[bulk]]) ** 2, axis=1)
        X = mbin[None, None, :] * ab[:, :, None] - mbin[None, :, None] * ab[:, None, :]
        cross = 0.5 * np.einsum("ri,rj,rij->r", n_, n_, X ** 2)
        vperp = np.where(c[2] > 0, (cross + sS[2] * Wk) / c2s, sS[2])
        sres = sqrt(vperp)
        Sm = (sS[1][:, None, None] + bK[:, None, None] * (Kn[:, :, None] - c[1][:, None, None])
              + sres[:, None, None] * tS[:, None, :])
        Sm = np.maximum(Sm, 0.0) * ms
        pb = pK[:, :, None] * pS[:, None, :]
        dKn = np.repeat(Kn, 3, axis=1) - c[1][:, None]
        dSn = Sm.reshape(n, 9) / ms - sS[1][:, None]
        b1 = np.where(c[2] > 0, KE1 / c2s, 0.0)
        cov = SE1 - bK * KE1
        den = vperp + KS_RIDGE * sS[2]
        b2 = np.where(den > 0, cov / np.where(den > 0, den, 1.0), 0.0)
        E1n = E1[:, None] + b1[:, None] * dKn + b2[:, None] * (dSn - bK[:, None] * dKn)
        expl = b1 ** 2 * c[2] + b2 ** 2 * vperp
        G2n = np.repeat(np.maximum(E2b - expl, 0.0)[:, None], 9, axis=1)
        E2 = E2b + E1 ** 2
        if nbr:
            nur = nub_g.sum(axis=1)
            p0r = np.exp(-nur)
            sc = np.where(nur > 0, -np.expm1(-nur) / np.where(nur > 0, nur, 1.0), 0.0)
            nz = np.where(nub_g > 0, nub_g, 1.0)
            kb_ = k1b_g / nz
            skb = sqrt(v2b_g / nz)
            kb_ = np.concatenate([kb_ - skb, kb_ + skb], axis=1)
            mb_, g2b, e1b = (np.tile(a / nz, (1, 2)) for a in (m1b_g, g2b_g, e1b_g))
            nub = np.tile(nub_g / 2, (1, 2))
        else:
            p0r = np.ones(n)
        Kp_c = np.concatenate([np.repeat(Kn, 3, axis=1)] + ([c[1][:, None] + kb_] if nbr else []),
                              axis=1)
        Sp_c = np.concatenate([Sm.reshape(n, 9)] + ([(sS[1] * ms)[:, None] + mb_] if nbr else []),
                              axis=1)
        wp_c = np.concatenate([pb.reshape(n, 9) * p0r[:, None]] + ([nub * sc[:, None]] if nbr else []),
                              axis=1)
        A1_c = np.concatenate([E1n] + ([E1[:, None] + e1b] if nbr else []), axis=1)
        A2_c = np.concatenate([G2n + E1n ** 2] + ([E2[:, None] + g2b + 2 * E1[:, None] * e1b]
                                                   if nbr else []), axis=1)
        Mlo = max(M - float(np.max(np.where(wp_c > MT_WFLOOR, Sp_c, 0.0))), MMIN)
        Mt = np.unique(np.concatenate([np.geomspace(Mlo, M, NM_TAB), [M]]))
        cos, z = self.cos, h["z"]
        _, rs_t, ks_t, _ = cos.nfw_params(Mt, z, self.zs)
        cm = dict(xg=xg, lxg=lxg, Kp_c=Kp_c, Sp_c=Sp_c, wp_c=wp_c, A1_c=A1_c, A2_c=A2_c,
                  Mt=Mt, lMt=log(Mt), log_rs=log(np.asarray(rs_t, float)),
                  log_ks=log(np.asarray(ks_t, float)),
                  ept=np.asarray(cos.epsilon(Mt, z), float) * np.ones_like(Mt), nbr=nbr,
                  kstar=kstar, ms=ms, ring=ring, ecd=ecd, C=C, nq=Kp_c.shape[1])
        h["_cm"] = cm
        return cm

    def emit(self, h, cm, xt, xq, W, phq=None):
        """Points of decorated cells (oracle_sub_emit): (K, G, W), [n_cell x nq]."""
        xq = np.ascontiguousarray(xq, float)
        W = np.ascontiguousarray(W, float)
        nq, ng = cm["nq"], cm["xg"].size
        out = [np.empty(xq.size * nq) for _ in range(3)]
        arr = [np.ascontiguousarray(cm[k]) for k in ("Kp_c", "Sp_c", "wp_c", "A1_c", "A2_c")]
        tab = [np.ascontiguousarray(cm[k]) for k in ("lMt", "log_rs", "log_ks", "ept")]
        ph = None if phq is None else chain._p(np.ascontiguousarray(phq, float))
        lxg = np.ascontiguousarray(cm["lxg"])
        chain.lib().oracle_sub_emit(
            xq.size, chain._p(xq), ph, chain._p(W), h["rs"], xt, nq, ng, float(cm["xg"][0]),
            *[chain._p(a) for a in arr], chain._p(lxg), tab[0].size, *[chain._p(a) for a in tab],
            h["M"], MMIN, *[chain._p(a) for a in out])
        return out


def decorate(model, h, grid, ks, xt, eps, with_x=False):
    """The subhalo branch of _truncated_host (sgl_full.py:2088-2127): host cells
    within reach of a clump, x <= x_sub = (r_t + max D) / r_s (1 + 1e-3), are
    replaced by the decorated points; returns (K, G, W) in the reference's
    order (plain inner cells, decorated inner cells, plain tail, decorated tail),
    and with ``with_x`` also each point's host-cell radius x (a decorated cell's
    x repeated for each of its points, sgl_full.py:2079-2085)."""
    from .ref2d import kg_eps, trunc_kg
    rs = h["rs"]
    x_sub = (xt * rs + float(np.max(h["D"]))) / rs * 1.001
    xc, xq, phq, W, xo, Wo = (grid[k] for k in ("xc", "xq", "phq", "W", "xo", "Wo"))
    xref = np.concatenate([xc, xo])
    xref = xref[xref <= x_sub]
    cm = model.coarse(h, xref)
    parts = []

    def emitted(x, Wx, ph):
        K, G, Wd = model.emit(h, cm, xt, x, Wx, ph)
        nq, rem = divmod(K.size, max(x.size, 1))
        if rem or nq == 0:
            raise RuntimeError("decorated host: not nq points per cell")
        return K, G, Wd, np.repeat(x, nq)

    mi = xq <= x_sub
    if (~mi).any():
        if phq is not None:
            Kn, Gn = kg_eps(xq[~mi], phq[~mi], eps, ks, xt)
        else:
            Kn, Gn = trunc_kg(xq[~mi], ks, xt)
        parts.append((Kn, Gn, W[~mi], xq[~mi]))
    if mi.any():
        parts.append(emitted(xq[mi], W[mi], None if phq is None else phq[mi]))
    if xo.size:
        mo = xo <= x_sub
        if (~mo).any():
            xn = xo[~mo]
            parts.append((np.zeros(xn.size), ks * 4 * _mu(xt) / xn ** 2, Wo[~mo], xn))
        if mo.any():
            parts.append(emitted(xo[mo], Wo[mo], None))
    return tuple(np.concatenate([p[j] for p in parts]) for j in range(4 if with_x else 3))
