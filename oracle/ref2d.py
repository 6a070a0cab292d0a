"""FP64 CPU restatement of the reference's 2-D (k_kappa, k_gamma) chain (checker only).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product never imports it.

This is the chain ``pipeline.run`` executes for the ingredient sets without
subhalos and clustering (halos, ellipticity, filaments), in the default
configuration (edge="r200", kap_floor=1e-8, zint="midpoint", phi_tol=1e-8,
tail box, no window, n_xi_wide=500):

1. shells: the engine's 103-node log-z grid below z_s as shell edges plus 0
   and z_s, one midpoint node per shell (sgl_full.py:471-518);
2. rows per (shell, mass node 1..99 of the 100-node 1e7..1e17 Msun grid):
   Sheth-Tormen dn with the backward dlnM, Ludlow+16 NFW parameters, x_t = c
   (r200c), outer radius x_m where the monopole shear falls to kap_floor
   (sgl_full.py:1848-1884);
3. cells per row (``_truncated_host``, sgl_full.py:2025-2138): ~43 points per
   decade from 1e-6 x_t to x_ell, edges at the truncation, critical-curve
   refinement (``_refine_edges`` on the untruncated profile, or per
   orientation ``_refine_edges_eps_batch`` for elliptical halos), the
   truncated (pseudo-elliptical) NFW at the geometric cell centres, and the
   monopole shear tail out to x_m at 12 points per decade;
   filaments per row on 120 u bins x 13 orientation nodes (sgl_full.py:1984-2007);
4. binning into the 961 x 300 signed-log fine (kappa, gamma) lattice with
   weight and centroid sums (``Cells.add`` / ``_add_direct_kernel``,
   sgl_full.py:1314-1383, 971-1046);
5. Lambda(k_k, k_g) = sum_b w_b [e^{i k_k kap_b} J0(k_g gam_b) - 1 - i k_k kap_b]
   over the occupied bins at their centroids (``Lam``, sgl_full.py:2728-2771,
   3088-3131), evaluated in the direct form of ``_lam_worker_block`` /
   ``sgl_direct.direct_real`` (sgl_full.py:2682-2704, sgl_direct.py:40-51);
   the reference's separable evaluation agrees with it to 7e-12
   (sgl_full.py:3031-3035);
6. ``invert`` (sgl_full.py:159-375): width-relative grid rules
   (``sgl.vector_grid``, sgl.py:1229-1258), tail box, measured k-reach
   (``_reach``), Hankel + Fourier transforms (``sgl.P_vector_joint``,
   sgl.py:1127-1156), Euler-Maclaurin endpoint terms, the constant-mu arc
   integral on the bilinear interpolant of P3 and the normalisation on the
   500-point wide grid, with its mean, variance, skewness and sigma_DL.

Pins (tests/test_oracle_ref2d.py): the explorer notebook's ingredient-ladder
rows at z_s = 1 for halos, + ellipticity and + filaments (sigma_kappa,
clipped variance and skewness, sigma_DL/D_L, P(ln mu > 1); 5 significant
figures), the compound-Poisson cumulant identity (curvature of ln Phi at 0 =
sum w kappa^2), the Hankel/Fourier/arc inversion of a Gaussian Phi against the
closed-form density, and brute-force Lambda on a tiny table.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np
from numpy import exp, log, pi, sqrt
from scipy.special import j0 as _sj0

from . import chain
from .cosmo import CKMS, Cosmology

# ---------------------------------------------------------------- grids --
ZMIN_ENG, ZMAX_ENG, NZ_ENG = 0.01, 12.341169644129371, 103     # sgl_full.py:386
MMIN_ENG, MMAX_ENG, NM_ENG = 1e7, 1e17, 100                    # sgl_full.py:387
FINE_K = (1e-7, 30.0, 480)                                     # sgl_full.py:456
FINE_G = (1e-9, 30.0, 300)
NX = 260                   # build_population nx (sgl_full.py:1536)
NPHI = 12                  # orientation nodes of an elliptical halo (sgl_full.py:1536)
KAP_FLOOR = 1e-8           # pipeline.run default (pipeline.py:117)
XG = np.logspace(-6, 3.5, 900)   # sgl.py:754, the circular critical-curve scan
BISECT_ITERS = 45                # sgl.py:757
EDGE_OFF = np.logspace(-11, -1.2, 22)   # sgl_full.py:2264, 2362

# invert / vector_grid constants (sgl.py:1176-1186, sgl_full.py:35-37, pipeline.py:117)
SIG_REACH, PPP_K, PPP_G, BOX_REACH = 12.0, 18.0, 96.0, 6.0
PTS_PER_SIG_KAP, PTS_PER_SIG_GAM, XI_REACH = 72.0, 36.0, 24.0
TAIL_KAP_MAX, TAIL_GAM_MAX, TAIL_PPP_G = 0.99, 0.8, 12.0
PHI_TOL = 1e-8
N_K_INT = 2400
N_XI_WIDE = 500


def engine_grids():
    """sgl_full.py:471-474."""
    z = np.exp(np.linspace(log(ZMIN_ENG), log(ZMAX_ENG), NZ_ENG))
    M = np.exp(np.linspace(log(MMIN_ENG), log(MMAX_ENG), NM_ENG))
    return z, M


def z_shells(zs: float, zint: str = "midpoint"):
    """The lens shells of z_shells (sgl_full.py:480-518) as a list of
    (z_lo, z_hi, z_node, dz).  "midpoint": the engine nodes below z_s as
    edges plus 0 and z_s, one node at each shell's middle; "engine": the
    shells [z_{j-1}, z_j] for j >= 1 with z_j < z_s, one node at the upper
    edge with the backward difference as its width (the builder's rule for
    dz=None, sgl_full.py:1858-1859)."""
    zl, _ = engine_grids()
    if zint == "engine":
        return [(float(zl[j - 1]), float(zl[j]), float(zl[j]), float(zl[j] - zl[j - 1]))
                for j in range(1, zl.size) if zl[j] < zs]
    if zint.startswith("gauss"):
        # "gaussN": the midpoint shells at N Gauss-Legendre nodes each, one
        # entry per node (z_shells' tasks, sgl_full.py:503-518)
        t, w = np.polynomial.legendre.leggauss(int(zint[5:]))
        edges = np.concatenate([[0.0], zl[zl < zs], [float(zs)]])
        return [(float(a), float(b), float(a + 0.5 * (b - a) * (tk + 1.0)), float(0.5 * (b - a) * wk))
                for a, b in zip(edges[:-1], edges[1:]) if b > a for tk, wk in zip(t, w)]
    if zint != "midpoint":
        raise ValueError(f"zint {zint!r} is not restated here")
    edges = np.concatenate([[0.0], zl[zl < zs], [float(zs)]])
    out = []
    for a, b in zip(edges[:-1], edges[1:]):
        if b <= a:
            continue
        out.append((float(a), float(b), float(a + 0.5 * (b - a) * 1.0), float(0.5 * (b - a) * 2.0)))
    return out


def fine_edges():
    """Signed-log kappa edges and log gamma edges of the fine lattice (sgl_full.py:460-468)."""
    pos = np.geomspace(FINE_K[0], FINE_K[1], FINE_K[2] + 1)
    ek = np.concatenate([-pos[::-1], pos])
    eg = np.geomspace(FINE_G[0], FINE_G[1], FINE_G[2] + 1)
    return ek, eg


# ------------------------------------------------------- lens profiles --
def kappa_gamma(x, ks):
    """Untruncated NFW (sgl.py:722-740) through the C restatement of nfw_Fh."""
    x = np.ascontiguousarray(np.atleast_1d(np.asarray(x, float)))
    F, h = np.zeros_like(x), np.zeros_like(x)
    chain.lib().oracle_nfw_Fh(x.size, chain._p(x), chain._p(F), chain._p(h))
    k = 2 * ks * F
    return k, 4 * ks * h / x ** 2 - k


def trunc_kg(x, ks, xt):
    return chain.trunc_kg(np.atleast_1d(np.asarray(x, float)), ks, xt)


def kg_eps(x, phi, eps, ks, xt):
    return chain.kg_eps(np.atleast_1d(np.asarray(x, float)), phi, eps, ks, xt)


def refine_edges(xe, ks, xm):
    """sgl_full.py:2342-2372 (_refine_edges, one z_s): bisected roots of detA of
    the untruncated circular profile on XG, 22 offsets on each side."""
    if ks < 0.05:
        return xe
    k, g = kappa_gamma(XG, ks)
    sgn = np.sign((1 - k) ** 2 - g * g)
    idx = np.nonzero(sgn[1:] * sgn[:-1] < 0)[0]
    if not idx.size:
        return xe
    a, b = XG[idx].copy(), XG[idx + 1].copy()
    s0 = sgn[idx]
    for _ in range(BISECT_ITERS):
        m = 0.5 * (a + b)
        km, gm = kappa_gamma(m, ks)
        same = np.sign((1 - km) ** 2 - gm * gm) == s0
        a = np.where(same, m, a)
        b = np.where(same, b, m)
    xc = 0.5 * (a + b)
    add = np.concatenate([(xc[:, None] * (1.0 + EDGE_OFF[None, :])).ravel(),
                          (xc[:, None] * (1.0 - EDGE_OFF[None, :])).ravel()])
    add = add[(add > xe[1]) & (add < xm)]
    return np.unique(np.concatenate([xe, add])) if add.size else xe


def refine_edges_eps(xe, eps, ks, xm, phis, xt):
    """sgl_full.py:2279-2339 (_refine_edges_eps_batch_one): 50-point screen at
    phi = 0 and pi/2, then per orientation the sign changes of detA on 1200
    log points bisected 45 times.  Returns one edge array per phi (the same
    object ``xe`` where nothing was refined)."""
    nph = len(phis)
    if ks < 0.05:
        return [xe] * nph
    xs = np.geomspace(max(xe[1], 1e-7), xm, 50)
    k0, g0 = kg_eps(xs, 0.0, eps, ks, xt)
    k1, g1 = kg_eps(xs, 0.5 * pi, eps, ks, xt)
    if np.all((1 - k0) ** 2 - g0 ** 2 > 0.02) and np.all((1 - k1) ** 2 - g1 ** 2 > 0.02):
        return [xe] * nph
    xb = np.geomspace(max(xe[1], 1e-7), xm, 1200)
    nb = xb.size
    k, g = kg_eps(np.tile(xb, nph), np.repeat(phis, nb), eps, ks, xt)
    d = (1 - k) ** 2 - g * g
    sg = np.where(d > 0, 1, np.where(d < 0, -1, 0)).reshape(nph, nb)
    A, Bv, S0, PH, OWN = [], [], [], [], []
    for j in range(nph):
        idx = np.nonzero(sg[j, 1:] * sg[j, :-1] < 0)[0]
        if idx.size:
            A.append(xb[idx]); Bv.append(xb[idx + 1]); S0.append(sg[j, idx])
            PH.append(np.full(idx.size, phis[j])); OWN.append(np.full(idx.size, j))
    if not A:
        return [xe] * nph
    a, b, s0 = np.concatenate(A), np.concatenate(Bv), np.concatenate(S0)
    ph, own = np.concatenate(PH), np.concatenate(OWN)
    for _ in range(BISECT_ITERS):
        m = 0.5 * (a + b)
        km, gm = kg_eps(m, ph, eps, ks, xt)
        dm = (1 - km) ** 2 - gm * gm
        same = np.where(dm > 0, 1, np.where(dm < 0, -1, 0)) == s0
        a, b = np.where(same, m, a), np.where(same, b, m)
    xc = 0.5 * (a + b)
    out = []
    for j in range(nph):
        xcj = xc[own == j]
        if not xcj.size:
            out.append(xe)
            continue
        add = np.concatenate([(xcj[:, None] * (1 + EDGE_OFF)).ravel(),
                              (xcj[:, None] * (1 - EDGE_OFF)).ravel()])
        add = add[(add > xe[1]) & (add < xm)]
        out.append(np.unique(np.concatenate([xe, add])))
    return out


def _centres(e):
    return np.where(e[:-1] > 0, sqrt(e[:-1] * np.maximum(e[1:], 0.0)), 0.5 * e[1])


def host_grid(ks, xt, xm, Nbar, eps=0.0, phis=None):
    """The cell grid of one (z, M) row of truncated halos (_truncated_host,
    sgl_full.py:2025-2087): the circular inner grid with the truncation edges
    and critical-curve refinement (centres xc), the evaluated cells (xq,
    orientation phq or None, weight W: per orientation on the
    ``_refine_edges_eps_batch`` grids for elliptical halos) and the monopole
    tail out to x_m at 12 points per decade (xo, Wo)."""
    x_ell = min(xt / sqrt(1 - eps), xm) if eps > 0 else min(xt, xm)
    x_in = 1e-6 * xt
    nx_eff = NX if ks > 0.05 else max(80, NX // 3)
    n_in = max(nx_eff, int(np.ceil(nx_eff / 6.0 * np.log10(x_ell / x_in))))
    xe = np.concatenate([[0.0], np.geomspace(x_in, x_ell, n_in)])
    xe = np.unique(np.concatenate([xe, [xt * (1 - 1e-9), xt, xt * (1 + 1e-9)]]))
    xe = xe[xe <= x_ell]
    xe = refine_edges(xe, ks, x_ell)
    xc = _centres(xe)
    if phis is not None:
        eps_edges = refine_edges_eps(xe, eps, ks, x_ell, phis, xt)
        xq = np.concatenate([_centres(e) for e in eps_edges])
        phq = np.concatenate([np.full(e.size - 1, p) for e, p in zip(eps_edges, phis)])
        W = np.concatenate([Nbar * (e[1:] ** 2 - e[:-1] ** 2) / xm ** 2 / len(phis)
                            for e in eps_edges])
    else:
        xq, phq = xc, None
        W = Nbar * (xe[1:] ** 2 - xe[:-1] ** 2) / xm ** 2
    if xm > x_ell * (1 + 1e-12):
        n_out = max(8, int(np.ceil(12 * np.log10(xm / x_ell))))
        eo = np.geomspace(x_ell, xm, n_out + 1)
        xo = sqrt(eo[:-1] * eo[1:])
        Wo = Nbar * (eo[1:] ** 2 - eo[:-1] ** 2) / xm ** 2
    else:
        xo = Wo = np.zeros(0)
    return dict(xc=xc, xq=xq, phq=phq, W=W, xo=xo, Wo=Wo)


def host_cells(ks, xt, xm, Nbar, eps=0.0, phis=None, grid=None):
    """(K, G, W) of one (z, M) row of truncated halos without subhalos
    (_truncated_host, sgl_full.py:2025-2138): the (pseudo-elliptical)
    truncated NFW at the cell centres, circular monopole shear on the tail."""
    g = grid if grid is not None else host_grid(ks, xt, xm, Nbar, eps, phis)
    if g["phq"] is not None:
        K, G = kg_eps(g["xq"], g["phq"], eps, ks, xt)
    else:
        K, G = trunc_kg(g["xq"], ks, xt)
    xo = g["xo"]
    Go = ks * 4 * (log(1 + xt) - xt / (1 + xt)) / xo ** 2
    return (np.concatenate([K, np.zeros_like(xo)]), np.concatenate([G, Go]),
            np.concatenate([g["W"], g["Wo"]]))


def kappa_gamma_cyl(u, kappa0):
    """filaments.py:54-66."""
    u = np.abs(np.asarray(u, float))
    s = sqrt(1.0 + u * u)
    kap = np.where(u <= 1.0, kappa0 * 2.0 * pi / s, 0.0)
    return kap, kappa0 * 4.0 * pi / (s + 1.0) - kap


# --------------------------------------------------------- population --
@dataclass
class Population:
    """Occupied fine bins (np.nonzero order) with weight and centroids, plus
    the builder's bookkeeping (sgl_full.py:1535-1808, 2742-2755)."""
    w: np.ndarray
    k: np.ndarray
    g: np.ndarray
    N_expl: float = 0.0
    N_fil: float = 0.0
    kbar: float = 0.0
    n_points: int = 0
    n_decorated: int = 0
    dropped: float = 0.0
    rows: list = field(default_factory=list)

    def sigma(self):
        """Lam.sigma without clustering (sgl_full.py:3437-3452)."""
        return float(sqrt(np.sum(self.w * self.k ** 2)))


def rows(cos: Cosmology, zs: float, ell=False, fil=False, edge="r200", shells=None,
         zint="midpoint"):
    """Per-(shell, mass) parameters of _build_shell_into (sgl_full.py:1848-1884,
    1984-1991) as dicts, shell-major (only the listed shells when given).
    edge "r200": x_t = c; "rvir": x_t = r_vir / r_s (sgl_full.py:1874-1880)."""
    _, Ml = engine_grids()
    dlnM = np.diff(log(Ml))
    out = []
    for i, (_, _, z, dz) in enumerate(z_shells(zs, zint)):
        if shells is not None and i not in shells:
            continue
        wz = CKMS / cos.Hz(z) * dz * pi * (1 + z) ** 2
        D = cos.Dg(z)
        scr = cos.Sigma_cr(z, zs)
        for jM in range(1, NM_ENG):
            M = Ml[jM]
            dn = cos.dndlnM(M, z) * dlnM[jM - 1]
            C, rs, ks, _ = cos.nfw_params(M, z, zs)
            xt = C if edge == "r200" else cos.eta(C, z) * C
            g_xt = ks * 4 * (log(1 + xt) - xt / (1 + xt)) / xt ** 2
            xm = xt * sqrt(g_xt / KAP_FLOOR) if g_xt > KAP_FLOOR else xt
            r = dict(shell=i, jM=jM, z=z, dz=dz, wz=wz, D=D, scr=scr, M=M, dn=dn, C=C, rs=rs,
                     ks=ks, xt=xt, xm=xm, beta=D * cos.halo_bias(M, z),
                     Nbar=wz * (xm * rs) ** 2 * dn,
                     eps=float(cos.epsilon(M, z)) if ell else 0.0)
            if fil:
                rsF = 1.0 * (M / 1e14) ** (1 / 3)
                LF = 20.0 * rsF
                r.update(rsF=rsF, LF=LF, k0base=rsF * 14.4 * cos.rhoc0 * LF / scr,
                         NF=wz * rsF ** 2 * chain.dndlnM_fil(cos, M, z) * dlnM[jM - 1],
                         beta_fil=D * chain.fil_bias(cos, M, z))
            out.append(r)
    return out


def _accumulate(cos, zs, ell, fil, edge, sub, shells=None, keep_points=False, xi=False,
                pos=False, zint="midpoint"):
    """The rows of the given shells binned into the fine lattice: the sums
    (w, w kappa, w gamma) per bin and the builder's bookkeeping.  With ``xi``
    also the xi_lin deposit of every row (oracle/xilin.py), one Limber node of
    width dchi = c dz / H(z) per shell (sgl_full.py:1673-1683), closed into
    compressed row groups at the shell's end.  With ``pos`` (the positive
    closure, xi['pos'] = True) also every shell's own fine-lattice sums and
    its node geometry (z, chi, dchi) (the pos_pieces of build_population,
    sgl_full.py:1693-1699, 1800-1802) and the lens rate per coarse bin."""
    ek, eg = fine_edges()
    nkf, ngf = ek.size - 1, eg.size - 1
    acc = dict(H0=np.zeros(nkf * ngf), H0k=np.zeros(nkf * ngf), H0g=np.zeros(nkf * ngf),
               N_expl=0.0, N_fil=0.0, kbar=0.0, n_points=0, n_decorated=0, drop=np.zeros(1),
               pts=[], model=None, xi_groups=[], pieces={})
    cur = dict(shell=None)

    def close_piece():
        if not pos or cur["shell"] is None:
            return
        h0, h0k, h0g = cur["H0"], cur["H0k"], cur["H0g"]
        idx = np.flatnonzero((h0 != 0) | (h0k != 0) | (h0g != 0))
        r = cur["row"]
        acc["pieces"][cur["shell"]] = dict(
            idx=idx, h0=h0[idx], h0k=h0k[idx], h0g=h0g[idx],
            node=(float(r["z"]), float(cos.chi(r["z"])), float(CKMS / cos.Hz(r["z"]) * r["dz"])))
    if xi:
        from . import xilin
        kp, cj = xilin.kperp_grid(cos)
    shell_xi = None
    phis = (np.arange(NPHI) + 0.5) * (0.5 * pi / NPHI)
    fph, fwt = chain.fil_phi_rule()
    ue = chain.fil_u_edges()
    uc = np.where(ue[:-1] > 0, sqrt(ue[:-1] * ue[1:]), 0.5 * ue[1])
    ufa = (ue[1:] ** 2 - ue[:-1] ** 2)

    def add(K, G, W):
        K, G, W = (np.ascontiguousarray(a, dtype=float).ravel() for a in (K, G, W))
        chain.lib().oracle_bin_fine(K.size, chain._p(K), chain._p(G), chain._p(W), chain._p(ek),
                                    ek.size, chain._p(eg), eg.size, chain._p(acc["H0"]),
                                    chain._p(acc["H0k"]), chain._p(acc["H0g"]),
                                    chain._p(acc["drop"]))
        if pos:
            chain.lib().oracle_bin_fine(K.size, chain._p(K), chain._p(G), chain._p(W),
                                        chain._p(ek), ek.size, chain._p(eg), eg.size,
                                        chain._p(cur["H0"]), chain._p(cur["H0k"]),
                                        chain._p(cur["H0g"]), chain._p(cur["drop"]))
        acc["n_points"] += K.size
        acc["kbar"] += float(np.sum(W * K))
        if keep_points:
            acc["pts"].append((K, np.clip(G, 1.0000001 * FINE_G[0], None), W))

    model = None
    if sub:
        if zint != "midpoint":
            raise ValueError("subhalos are restated on the midpoint shells only")
        from . import sub as _sub
        zl, Ml = np.array([sh[2] for sh in z_shells(zs)]), engine_grids()[1]
        model = acc["model"] = _sub.SubhaloModel(cos, zs, zl, Ml)
    for r in rows(cos, zs, ell=ell, fil=fil, edge=edge, shells=shells, zint=zint):
        if pos and cur["shell"] != r["shell"]:
            close_piece()
            cur.update(shell=r["shell"], row=r, H0=np.zeros(nkf * ngf), H0k=np.zeros(nkf * ngf),
                       H0g=np.zeros(nkf * ngf), drop=np.zeros(1))
        if xi and (shell_xi is None or shell_xi[0] != r["shell"]):
            if shell_xi is not None:
                acc["xi_groups"] += shell_xi[1].end(shell_xi[0], compress=xi != "dense")
            shell_xi = (r["shell"], xilin.Shell(kp, cj, CKMS / cos.Hz(r["z"]) * r["dz"], pos=pos))
        acc["N_expl"] += r["Nbar"]
        grid = host_grid(r["ks"], r["xt"], r["xm"], r["Nbar"], r["eps"], phis if ell else None)
        h = model.host(r["shell"], r["jM"]) if model is not None else None
        if h is not None:
            K, G, W, X = _sub.decorate(model, h, grid, r["ks"], r["xt"], r["eps"], with_x=True)
            acc["n_decorated"] += 1
        else:
            K, G, W = host_cells(r["ks"], r["xt"], r["xm"], r["Nbar"], r["eps"], grid=grid)
            X = np.concatenate([grid["xq"], grid["xo"]])
        add(K, G, W)
        if xi:
            # halos: lens distance R = x r_s (1 + z), tangential shear, and the
            # analytic shear tail beyond x_m (sgl_full.py:2128-2136)
            a = r["rs"] * (1 + r["z"])
            shell_xi[1].deposit(K, G, W, X * a, r["beta"], True)
            Rm = r["xm"] * a
            shell_xi[1].shear_tail(r["beta"], r["ks"] * 4 * (log(1 + r["xt"]) - r["xt"] / (1 + r["xt"]))
                                   * a ** 2, Rm, r["Nbar"] / (pi * Rm ** 2))
        if fil:
            acc["N_fil"] += r["NF"]
            k0 = r["k0base"] / np.maximum(2 * r["rsF"], np.abs(np.cos(fph)) * r["LF"])
            K, G = kappa_gamma_cyl(uc[:, None], k0[None, :])
            W = r["NF"] * ufa[:, None] * fwt[None, :] * np.ones_like(K)
            add(K, G, W)
            if xi:
                # filaments: R = u r_sF (1 + z), axis not tangential (m = 0),
                # sgl_full.py:2002-2004
                Rf = uc[:, None] * (r["rsF"] * (1 + r["z"])) * np.ones_like(K)
                shell_xi[1].deposit(K, G, W, Rf, r["beta_fil"], False)
    if shell_xi is not None:
        acc["xi_groups"] += shell_xi[1].end(shell_xi[0], compress=xi != "dense")
    close_piece()
    return acc


def _worker(args):
    cos_kw, zs, ell, fil, edge, sub, shells, xi, pos, zint = args
    acc = _accumulate(Cosmology(**cos_kw), zs, ell, fil, edge, sub, shells, xi=xi, pos=pos,
                      zint=zint)
    acc.pop("model")
    acc.pop("pts")
    return acc


def build(cos: Cosmology, zs: float, ell=False, fil=False, keep_points=False, edge="r200",
          sub=False, workers=None, xi=False, pos=False, zint="midpoint"):
    """build_population(..., kap_floor=1e-8, trunc="r200", zint="midpoint"
    or "engine")
    without clustering: every row's cells (decorated by subhalos on hosts
    that carry them when ``sub``, oracle/sub.py) binned into the fine lattice
    with weight and centroid sums.  With any ingredient beyond plain halos the
    shells are shared out over ``workers`` processes (default: up to 8), each
    binning its shells, and the partial sums are added in worker order."""
    n_sh = len(z_shells(zs, zint))
    if workers is None:
        workers = min(8, os.cpu_count() or 1) if (sub or xi or ell or fil) else 1
    workers = max(1, min(int(workers), n_sh))
    if workers == 1 or keep_points:
        accs = [_accumulate(cos, zs, ell, fil, edge, sub, None, keep_points, xi=xi, pos=pos,
                            zint=zint)]
    else:
        import multiprocessing as mp
        jobs = [(cos.kw, zs, ell, fil, edge, sub, set(range(w, n_sh, workers)), xi, pos, zint)
                for w in range(workers)]
        # fresh interpreters (no inherited OpenMP or device state), one
        # OpenMP thread each
        old = os.environ.get("OMP_NUM_THREADS")
        os.environ["OMP_NUM_THREADS"] = "1"
        try:
            with mp.get_context("spawn").Pool(workers) as pool:
                accs = pool.map(_worker, jobs)
        finally:
            if old is None:
                os.environ.pop("OMP_NUM_THREADS", None)
            else:
                os.environ["OMP_NUM_THREADS"] = old
    H0, H0k, H0g = (sum(a[k] for a in accs) for k in ("H0", "H0k", "H0g"))
    pop = Population(np.zeros(0), np.zeros(0), np.zeros(0))
    for k in ("N_expl", "N_fil", "kbar", "n_points", "n_decorated"):
        setattr(pop, k, sum(a[k] for a in accs))
    pop.dropped = float(sum(a["drop"][0] for a in accs))
    pop.sub_model = accs[0].get("model")
    nz = np.nonzero(H0)[0]
    pop.w = H0[nz]
    pop.k = H0k[nz] / pop.w
    pop.g = H0g[nz] / pop.w
    pop.bins = nz
    # xi_lin row groups in shell order, m ascending within a shell
    pop.xi_groups = sorted((g for a in accs for g in a["xi_groups"]),
                           key=lambda g: (g["shell"], g["m"]))
    if pos:
        pop.pieces = {k: v for a in accs for k, v in a["pieces"].items()}
        pop.H0, pop.H0k, pop.H0g = H0, H0k, H0g
    if keep_points:
        pop.points = accs[0]["pts"]
    return pop


# ------------------------------------------------------------- Lambda --
def j0m1(x):
    """J0(x) - 1 without cancellation (sgl_full.py:2576-2582)."""
    x = np.asarray(x, float)
    out = _sj0(x) - 1.0
    s = x < 1e-3
    out[s] = -x[s] ** 2 / 4 * (1 - x[s] ** 2 / 16)
    return out


def lam(pop: Population, kk, kg, kblock=256, gblock=512):
    """Lambda(kk, kg) = sum_b w_b [e^{i kk k_b} J0(kg g_b) - 1 - i kk k_b],
    split as the reference's separable form does (sgl_full.py:3040-3086):
    sum_b w_b (e^{ix} - 1 - ix) exact per bin (compensated cos - 1 and the
    sin - x series below |x| < 1e-2, sgl_full.py:2598-2611) plus the dense
    product sum_b w_b e^{i kk k_b} (J0 - 1) (sgl_direct.py:40-51)."""
    from concurrent.futures import ThreadPoolExecutor
    kk = np.atleast_1d(np.asarray(kk, float))
    kg = np.atleast_1d(np.asarray(kg, float))
    nk = kk.size
    out = np.empty((nk, kg.size), complex)
    w, k0, g0 = pop.w, pop.k, pop.g
    packed = np.empty((2 * nk, w.size))
    a1 = np.empty(nk, complex)

    def kblock_rows(lo):
        hi = min(lo + kblock, nk)
        ph = np.outer(kk[lo:hi], k0)
        hs = np.sin(0.5 * ph)
        cm1 = -2.0 * hs * hs
        sn = np.sin(ph)
        p2 = ph * ph
        smp = np.where(np.abs(ph) < 1e-2, -ph * p2 / 6.0 * (1.0 - p2 / 20.0 * (1.0 - p2 / 42.0)),
                       sn - ph)
        a1[lo:hi] = (cm1 @ w) + 1j * (smp @ w)
        packed[lo:hi] = (cm1 + 1.0) * w
        packed[nk + lo:nk + hi] = sn * w

    def gblock_cols(glo):
        ghi = min(glo + gblock, kg.size)
        v = packed @ j0m1(np.outer(g0, kg[glo:ghi]))
        out[:, glo:ghi] = v[:nk] + 1j * v[nk:] + a1[:, None]

    # blocks are independent (each writes its own rows / columns), so the
    # result does not depend on how they are spread over threads (numpy's
    # elementwise kernels release the GIL); one BLAS thread per block thread
    from threadpoolctl import threadpool_limits
    with threadpool_limits(limits=1, user_api="blas"), \
            ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        list(ex.map(kblock_rows, range(0, nk, kblock)))
        list(ex.map(gblock_cols, range(0, kg.size, gblock)))
    return out


# -------------------------------------------------------------- invert --
def vector_grid(sigma):
    """sgl.py:1229-1258."""
    sig_xi = 2.0 * sigma
    xi_max = XI_REACH * sig_xi
    xi_min = -BOX_REACH * sig_xi
    kap_max = 1.0 - exp(-0.5 * xi_max)
    kap_min = -BOX_REACH * sigma
    gam_max = BOX_REACH * sigma
    k_max = SIG_REACH / sigma
    kap_ext = max(abs(kap_min), kap_max)
    dk_k = 2.0 * pi / (PPP_K * kap_ext)
    dk_g = 2.0 * pi / (PPP_G * gam_max)
    return dict(nk_k=int(np.ceil(k_max / dk_k)) + 1, nk_g=int(np.ceil(k_max / dk_g)) + 1,
                k_k_max=k_max, k_g_max=k_max,
                n_kap=int(np.ceil((kap_max - kap_min) / (sigma / PTS_PER_SIG_KAP))) + 1,
                n_gam=int(np.ceil(gam_max / (sigma / PTS_PER_SIG_GAM))) + 1,
                kap_min=kap_min, kap_max=kap_max, gam_max=gam_max,
                xi_min=xi_min, xi_max=xi_max, sigma=sigma)


def reach(lamfn, k0, axis, tol=PHI_TOL, max_mult=64.0):
    """sgl_full.py:38-47 (_reach)."""
    mults = 1.25 ** np.arange(0, int(np.ceil(log(max_mult) / log(1.25))) + 1)
    ks = k0 * mults
    zero = np.zeros(1)
    L = lamfn(ks, zero)[:, 0] if axis == "k" else lamfn(zero, ks)[0, :]
    phi = np.abs(np.exp(L))
    ok = np.nonzero(phi <= tol)[0]
    i = int(ok[0]) if len(ok) else int(np.argmin(phi))
    return float(ks[i]), float(phi[i])


def grid_plan(lamfn, sigma, tol=PHI_TOL):
    """The k and (kappa, gamma) grids of invert with the tail box
    (sgl_full.py:190-262)."""
    grid = vector_grid(sigma)
    g0 = grid["gam_max"]
    kap_max = max(grid["kap_max"], TAIL_KAP_MAX)
    gam_max = max(g0, TAIL_GAM_MAX)
    dk0 = (2 * pi / (PPP_K * max(abs(grid["kap_min"]), grid["kap_max"])),
           2 * pi / (PPP_G * grid["gam_max"]))
    dk0 = (dk0[0], min(dk0[1], 2 * pi / (TAIL_PPP_G * gam_max)))
    if kap_max != grid["kap_max"] or gam_max != grid["gam_max"]:
        grid["kap_max"], grid["gam_max"] = float(kap_max), float(gam_max)
        grid["n_kap"] = int(np.ceil((grid["kap_max"] - grid["kap_min"])
                                    / (sigma / PTS_PER_SIG_KAP))) + 1
        grid["n_gam"] = int(np.ceil(grid["gam_max"] / (sigma / PTS_PER_SIG_GAM))) + 1
    grid["k_k_max"], phi_k = reach(lamfn, grid["k_k_max"], "k", tol)
    grid["k_g_max"], phi_g = reach(lamfn, grid["k_g_max"], "g", tol)
    grid["nk_k"] = int(np.ceil(grid["k_k_max"] / dk0[0])) + 1
    grid["nk_g"] = int(np.ceil(grid["k_g_max"] / dk0[1])) + 1
    grid.update(phi_at_k_cut=phi_k, phi_at_g_cut=phi_g, g_def=min(g0, grid["gam_max"]))
    return grid


def p3_joint(k_k, k_g, Phi, kap, gam):
    """sgl.P_vector_joint (sgl.py:1127-1156) plus the Euler-Maclaurin
    endpoint terms of the k_gamma trapezoid (sgl_full.py:277-298)."""
    dk_k, dk_g = k_k[1] - k_k[0], k_g[1] - k_g[0]
    w_kg = np.full(k_g.size, dk_g)
    w_kg[0] *= 0.5
    w_kg[-1] *= 0.5
    M = (k_g * w_kg)[:, None] * _sj0(np.outer(gam, k_g)).T
    F = Phi @ M
    w_kk = np.full(k_k.size, dk_k)
    w_kk[0] *= 0.5
    w_kk[-1] *= 0.5
    ph = np.outer(kap, k_k)
    Fw = F * w_kk[:, None]
    P3 = (1.0 / (2.0 * pi ** 2)) * (np.cos(ph) @ np.ascontiguousarray(Fw.real)
                                    + np.sin(ph) @ np.ascontiguousarray(Fw.imag))
    Phi0, Phi1 = Phi[:, 0], Phi[:, 1]
    Phi2 = (Phi1 - Phi0) / dk_g ** 2
    c = 1.0 / (2.0 * pi ** 2)
    va = (dk_g ** 2 / 12.0) * Phi0 - (dk_g ** 4 / 120.0) * Phi2
    vb = (dk_g ** 4 / 480.0) * Phi0
    C, S = np.cos(ph) * w_kk[None, :], np.sin(ph) * w_kk[None, :]
    a = c * (C @ va.real + S @ va.imag)
    b = c * (C @ vb.real + S @ vb.imag)
    return P3 + a[:, None] + b[:, None] * gam[None, :] ** 2


def bilinear(kap, gam, P3, pk, pg):
    """Linear RegularGridInterpolator on (kap, gam), fill 0 outside."""
    nk, ng = kap.size, gam.size
    hk = (kap[-1] - kap[0]) / (nk - 1)
    hg = (gam[-1] - gam[0]) / (ng - 1)
    ik = np.clip(np.floor((pk - kap[0]) / hk).astype(np.int64), 0, nk - 2)
    ig = np.clip(np.floor((pg - gam[0]) / hg).astype(np.int64), 0, ng - 2)
    ik = np.where((ik < nk - 2) & (pk >= kap[np.minimum(ik + 1, nk - 1)]), ik + 1, ik)
    ik = np.where((ik > 0) & (pk < kap[ik]), ik - 1, ik)
    ig = np.where((ig < ng - 2) & (pg >= gam[np.minimum(ig + 1, ng - 1)]), ig + 1, ig)
    ig = np.where((ig > 0) & (pg < gam[ig]), ig - 1, ig)
    tk = (pk - kap[ik]) / (kap[ik + 1] - kap[ik])
    tg = (pg - gam[ig]) / (gam[ig + 1] - gam[ig])
    v = (P3[ik, ig] * (1 - tk) * (1 - tg) + P3[ik, ig + 1] * (1 - tk) * tg
         + P3[ik + 1, ig] * tk * (1 - tg) + P3[ik + 1, ig + 1] * tk * tg)
    out = (pk < kap[0]) | (pk > kap[-1]) | (pg < gam[0]) | (pg > gam[-1])
    return np.where(out, 0.0, v)


def arc_integral(kap, gam, P3, g_def, xi_grid):
    """P_i of invert (sgl_full.py:299-360): pi/mu^2 times the integral of P3
    along the constant-mu arc, the default segment (gamma <= g_def) over
    N_K_INT kappa points and the tail-box segment parametrised by gamma at
    half the gamma step."""
    gmax = gam[-1]
    dgam = gam[1] - gam[0]
    out = np.zeros(len(xi_grid))
    for i, xi in enumerate(xi_grid):
        mu = exp(xi)
        rm = 1.0 / sqrt(mu)
        hi = 1.0 - rm
        lo = max(kap[0], 1.0 - sqrt(1.0 / mu + g_def ** 2))
        if hi <= lo:
            continue
        ke = np.linspace(lo, hi, N_K_INT)
        ge = sqrt(np.clip((hi - ke) * (1.0 - ke + rm), 0.0, None))
        out[i] = pi / mu ** 2 * np.trapezoid(bilinear(kap, gam, P3, ke, ge), ke)
        if gmax > g_def and lo > kap[0]:
            gb = min(gmax, sqrt(max((hi - kap[0]) * (1.0 - kap[0] + rm), 0.0)))
            if gb > g_def:
                n = int(np.ceil((gb - g_def) / (0.5 * dgam))) + 1
                gx = np.linspace(g_def, gb, n)
                r = np.sqrt(1.0 / mu + gx ** 2)
                kx = np.clip(1.0 - r, kap[0], kap[-1])
                out[i] += pi / mu ** 2 * np.trapezoid(bilinear(kap, gam, P3, kx, gx) * gx / r, gx)
    return out


@dataclass
class Result:
    xi: np.ndarray
    P: np.ndarray
    sigma: float
    mean: float
    var: float
    skew: float
    sigmaDL: float
    norm: float
    grid: dict
    pop: Population = None
    P3: np.ndarray = None
    kap: np.ndarray = None
    gam: np.ndarray = None


def invert(lamfn, sigma, xi_out, tol=PHI_TOL, keep=False):
    """sgl_full.invert (sgl_full.py:159-375) for the Gaussian closure, no
    window, non-blocked (every grid here is below INVERT_BLOCK_PTS)."""
    grid = grid_plan(lamfn, sigma, tol)
    k_k = np.linspace(0.0, grid["k_k_max"], grid["nk_k"])
    k_g = np.linspace(0.0, grid["k_g_max"], grid["nk_g"])
    kap = np.linspace(grid["kap_min"], grid["kap_max"], grid["n_kap"])
    gam = np.linspace(0.0, grid["gam_max"], grid["n_gam"])
    Phi = np.exp(lamfn(k_k, k_g))
    P3 = p3_joint(k_k, k_g, Phi, kap, gam)
    xw = np.linspace(grid["xi_min"], grid["xi_max"], N_XI_WIDE)
    pw = arc_integral(kap, gam, P3, grid["g_def"], xw)
    norm = np.trapezoid(pw, xw)
    ps = pw / norm
    mean = float(np.trapezoid(xw * ps, xw))
    var = float(np.trapezoid((xw - mean) ** 2 * ps, xw))
    skew = float(np.trapezoid((xw - mean) ** 3 * ps, xw) / var ** 1.5)
    x = exp(-0.5 * xw)
    md = float(np.trapezoid(x * ps, xw))
    sdl = float(sqrt(np.trapezoid((x - md) ** 2 * ps, xw)) / md)
    P = arc_integral(kap, gam, P3, grid["g_def"], np.asarray(xi_out, float)) / norm
    r = Result(np.asarray(xi_out, float), P, sigma, mean, var, skew, sdl, float(norm), grid)
    if keep:
        r.P3, r.kap, r.gam = P3, kap, gam
    return r


def clipped_moments(xi, P, clip=1.0):
    """pipeline.py:277-284."""
    m = np.abs(xi) <= clip
    x, p = xi[m], P[m]
    p = p / np.trapezoid(p, x)
    mu = np.trapezoid(x * p, x)
    v = np.trapezoid((x - mu) ** 2 * p, x)
    return dict(mean=float(mu), var=float(v), skew=float(np.trapezoid((x - mu) ** 3 * p, x) / v ** 1.5))


XI_BODY = np.linspace(-1.0, 1.0, 1601)    # cli.py / explorer body grid
XI_TAIL = np.linspace(1.005, 9.0, 1600)   # and tail grid


def limber_lss_var(cos: Cosmology, zs, lmin=1.0, lmax=1e7, nchi=600):
    """limber_lss_var (sgl_full.py:2541-2572): the Limber variance of the
    Gaussian large-scale-structure convergence of clustering="limber",
      v = int dchi W^2 Dg^2 int_{lmin/f_K}^{lmax/f_K} dk k P0(k) / (2 pi),
      W = 3/2 Om H0^2 (1 + z) f_K(chi) f_K(chi_s - chi) / f_K(chi_s),
    on the reference's nodes: chi(z) on 400 uniform z and z(chi) by linear
    interpolation, 600 uniform chi in [1e-4, 1 - 1e-4] chi_s, the k integral
    a cumulative trapezoid on 40000 geometric k in [1e-7, 1e5] read by
    linear interpolation, the chi integral a trapezoid."""
    chis = cos.chi(zs)
    zt = np.linspace(0.0, zs, 400)
    ct = np.array([cos.chi(z) for z in zt])
    chi = np.linspace(chis * 1e-4, chis * (1 - 1e-4), nchi)
    z = np.interp(chi, ct, zt)
    H0 = 100.0 * cos.h / CKMS
    if cos.Ok == 0.0:
        transverse = chi
        W = 1.5 * cos.Om * H0 ** 2 * (1 + z) * chi * (chis - chi) / chis
    else:
        transverse = cos.f_K(chi)
        W = 1.5 * cos.Om * H0 ** 2 * (1 + z) * transverse * cos.f_K(chis - chi) / cos.f_K(chis)
    D = np.array([cos.Dg(zz) for zz in z])
    Kg = np.geomspace(1e-7, 1e5, 40000)
    f = Kg * cos.P0(Kg) / (2 * pi)
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(Kg))])
    I = np.interp(lmax / transverse, Kg, cum) - np.interp(lmin / transverse, Kg, cum)
    return float(np.trapezoid(W ** 2 * D ** 2 * I, chi))


def lss_term(v, kk, kg):
    """The limber term of ln Phi, -v/2 (k_kappa^2 + k_gamma^2 / 2)
    (sgl_full.py:3104-3105)."""
    kk = np.atleast_1d(np.asarray(kk, float))
    kg = np.atleast_1d(np.asarray(kg, float))
    return -0.5 * v * (kk[:, None] ** 2 + 0.5 * kg[None, :] ** 2)


def run(zs, cosmo: dict = None, ell=False, fil=False, xi_out=None, keep=False, edge="r200",
        sub=False, bias=False, lss=False, zint="midpoint"):
    """pipeline.run (pipeline.py:114-275) with the Gaussian pair closure.
    ``bias``: the xi_lin clustering term (oracle/xilin.py) in Lambda and in
    sigma_kappa (sgl_full.py:3094-3128, 3437-3446); bias="dense" keeps every
    shell's rows uncompressed and unpruned (one per k_perp node).
    ``lss``: clustering="limber" with the bias ingredient instead
    (pipeline.py:195-202): no xi_lin deposit, the Gaussian term of
    limber_lss_var in Lambda and in sigma_kappa (sgl_full.py:3104-3105, 3439).
    ``zint``: the lens shells, "midpoint" (the default) or "engine".
    xi_out None: the body and tail grids of the CLI (pipeline.py:211-215)."""
    if lss and bias:
        raise ValueError("the limber term replaces the xi_lin term")
    cos = Cosmology(**(cosmo or {}))
    pop = build(cos, zs, ell=ell, fil=fil, edge=edge, sub=sub, xi=bias, zint=zint)
    xo = np.concatenate([XI_BODY, XI_TAIL]) if xi_out is None else np.asarray(xi_out, float)
    if lss:
        v = pop.lss_var = limber_lss_var(cos, zs)
        sig = float(sqrt(float(np.sum(pop.w * pop.k ** 2)) + v))
        res = invert(lambda a, b: lam(pop, a, b) + lss_term(v, a, b), sig, xo, keep=keep)
    elif bias:
        from . import xilin
        psi = pop.psi2 = xilin.Psi2(pop.xi_groups, prune=0 if bias == "dense" else xilin.XI_PRUNE)
        sig = float(sqrt(pop.sigma() ** 2 + psi.var()))
        res = invert(lambda a, b: lam(pop, a, b) + psi(a, b), sig, xo, keep=keep)
    else:
        sig = pop.sigma()
        res = invert(lambda a, b: lam(pop, a, b), sig, xo, keep=keep)
    res.pop = pop
    return res


def ladder_row(res: Result):
    """The explorer's ladder columns (build_explorer.py:164-176) from a run
    on the body + tail grids."""
    nb = XI_BODY.size
    cm = clipped_moments(res.xi[:nb], res.P[:nb])
    m = res.xi >= 1.0
    return dict(sigma_kappa=res.sigma, clipped_var=cm["var"], clipped_skew=cm["skew"],
                sigmaDL=res.sigmaDL, p_tail1=float(np.trapezoid(res.P[m], res.xi[m])))
