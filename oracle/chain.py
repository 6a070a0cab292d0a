"""FP64 CPU oracle of the kernel-shape magnification-PDF chain (checker only).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  Never on the product path.

The chain (one cosmology, one source redshift):

1. slices: ``n_z`` slices uniform in comoving distance on [0, chi(z_s)],
   node at the slice midpoint (the kernel shape of the lens-z quadrature
   that sgl_full.z_shells, sgl_full.py:480-518, performs on the reference grid);
2. rows per (slice, mass): Sheth-Tormen dn/dlnM, Ludlow+16 NFW parameters,
   truncation at r200 (x_t = c, sgl_full.py:1876-1880), r_vir = eta r200
   (subhalos.py:472-481), linear bias beta = Dg b_h with the bias growth Dg
   (sgl_full.py:521-543,1861,1871),
   lens weight dchi (1+z)^2 dn (sgl_full.py:1860);
3. cells: ``n_b`` log impact-parameter bins over [u_lo, u_hi] r_vir, each
   cell at the geometric centre of its bin (sgl_full.py:1905), weight = row
   weight x annulus area, kappa/gamma of the truncated NFW, jump
   xi = -ln detA and the source-plane tilt w e^{-xi} (sgl.py:808-815,949-951);
   with ``filaments``, a second population per (slice, mass): cylinder
   filaments on 120 u bins x 13 orientation nodes (sgl_full.py:1984-2007);
4. Lambda(k) compensated, a_i(k) per slice (oracle/chain.c);
5. Psi_2(k) = 1/2 sum_ij C_ij a_i(k) a_j(k), C_ij the slice covariance of the
   smoothed 1-D linear density field (sgl_full.py:2514-2538 restated on these
   slices; the quadratic form of Lam.Q, sgl_full.py:3353-3397, nord = 1);
6. P(xi) = (1/pi) Re int_0^inf e^{-ik xi + Lambda + Psi_2} dk (sgl.py:994-1012),
   phi linear between the k nodes and phi(0) = 1 (Filon weights);
   with phi multiplied by the reference's positive-definite cubic B-spline
   window of support [0, k_max] (sgl_full.py:83-94, k_kappa factor): the
   transform of p keeps unit mass exactly (phi(0) = 1), so p is not
   renormalised;
7. diagnostics on the output grid in FP64 (trapezoid, as api.py:75-79 and
   sgl_full.py:362-371): grid mass, mean, variance, skewness, sigma_DL/D_L,
   <mu>, P(mu > 2), negative mass.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np
from scipy.optimize import brentq

from .cosmo import Cosmology

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        from .build import LIB, build
        if not os.path.exists(LIB):
            build()
        _LIB = ctypes.CDLL(LIB)
        d = ctypes.POINTER(ctypes.c_double)
        i64 = ctypes.c_int64
        p64 = ctypes.POINTER(ctypes.c_int64)
        _LIB.oracle_cells.argtypes = [i64, d, i64, d, d, d, d, d]
        _LIB.oracle_cells_fil.argtypes = [i64, d, i64, d, i64, d, d, d, d, d, d]
        _LIB.oracle_lnphi.argtypes = [i64, p64, d, d, d, i64, d, d, d]
        _LIB.oracle_invert.argtypes = [i64, d, i64, d, d, d]
        _LIB.oracle_filon_weights.argtypes = [i64, d, i64, d, d]
        _LIB.oracle_trunc_kg.argtypes = [i64, d, ctypes.c_double, ctypes.c_double, d, d]
        _LIB.oracle_nfw_Fh.argtypes = [i64, d, d, d]
        _LIB.oracle_kg_eps.argtypes = [i64, d, d, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_double, d, d]
        _LIB.oracle_cells_ell.argtypes = [i64, d, d, i64, d, i64, d, d, d, d]
        _LIB.oracle_sub_ring.argtypes = [i64, d, i64, d, i64, d, i64, d, d, d, d]
        f64 = ctypes.c_double
        _LIB.oracle_bin_fine.argtypes = [i64, d, d, d, d, i64, d, i64, d, d, d, d]
        _LIB.oracle_xi_deposit.argtypes = [i64, d, d, d, d, ctypes.c_double, i64, i64, d, d, i64,
                                           d, i64, d, d]
        _LIB.oracle_sub_emit.argtypes = [i64, d, d, d, f64, f64, i64, i64, f64, d, d, d, d, d, d,
                                         i64, d, d, d, d, f64, f64, d, d, d]
        _LIB.oracle_xi_rate.argtypes = [i64, d, d, d, d, i64, d, i64, d]
        _LIB.oracle_vp_lnr.argtypes = [i64, p64, p64, p64, d, d, d, d, p64, p64, i64, d, d, d,
                                       i64, d, d, i64, d, i64, d, d, d]
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


@dataclass
class Shape:
    """Kernel-shape grids.  Defaults are the 64 x 256 x 512 / 8192-k /
    2048-node shape; tests shrink them."""
    n_z: int = 64
    n_m: int = 256
    n_b: int = 512
    m_lo_h: float = 1e8          # Msun/h
    m_hi_h: float = 1e16
    u_lo: float = 1e-4           # impact parameter / r_vir
    u_hi: float = 10.0
    n_k: int = 8192
    k_lo: float = 1e-2
    k_hi: float = 1e6
    n_out: int = 2048
    out_lo: float = -0.5
    out_hi: float = 3.0
    clustering: bool = True
    window: bool = True
    filaments: bool = False
    edge: str = "r200"           # halo truncation: "r200" (x_t = c) or "rvir" (x_t = eta c)
    ell: bool = False            # halo ellipticity: N_PHI orientation nodes per cell

    def k_nodes(self):
        """k_q: 0 followed by n_k log-spaced nodes."""
        return np.concatenate([[0.0], np.logspace(math.log10(self.k_lo), math.log10(self.k_hi), self.n_k)])

    def out_grid(self):
        return np.linspace(self.out_lo, self.out_hi, self.n_out)

    def u_edges(self):
        return np.logspace(math.log10(self.u_lo), math.log10(self.u_hi), self.n_b + 1)

    def m_nodes(self, h):
        """Mass bins log-uniform over [m_lo_h, m_hi_h] Msun/h; node = bin
        centre in ln M, converted to Msun (M = M_h / h)."""
        e = np.linspace(math.log(self.m_lo_h), math.log(self.m_hi_h), self.n_m + 1)
        return np.exp(0.5 * (e[1:] + e[:-1])) / h, float(e[1] - e[0])


def slice_nodes(cos: Cosmology, zs: float, n_z: int):
    """Uniform-in-chi slices: returns (z_node, chi_node, dchi, chi_edges)."""
    chis = cos.chi(zs)
    edges = np.linspace(0.0, chis, n_z + 1)
    mid = 0.5 * (edges[1:] + edges[:-1])
    z = np.array([brentq(lambda zz: cos.chi(zz) - c, 0.0, zs, xtol=1e-14, rtol=1e-14) for c in mid])
    return z, mid, chis / n_z, edges


def p1d(cos: Cosmology, kq: np.ndarray, Rs: float = 20.0):
    """The 1-D power of the top-hat-smoothed linear field along the line of
    sight, (1/2 pi) int_{k}^{inf} K P0(K) W^2(K Rs) dK, as BiasField1D forms it
    (sgl_full.py:2514-2530): a running trapezoid on 20000 geometric nodes from
    k_1 / 2 to 200 / Rs, interpolated at kq."""
    Kg = np.geomspace(kq[0] * 0.5, 200.0 / Rs, 20000)
    x = Kg * Rs
    W = np.where(x < 1e-2, 1 - x ** 2 / 10 * (1 - x ** 2 / 28), 3 * (np.sin(x) - x * np.cos(x)) / x ** 3)
    f = Kg * cos.P0(Kg) * W ** 2
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(Kg))])
    return (cum[-1] - np.interp(kq, Kg, cum)) / (2 * np.pi)


def slice_covariance(cos: Cosmology, chi_edges: np.ndarray, chis: float, Rs: float = 20.0):
    """sgl_full.py:2514-2538 (BiasField1D, bias_window=1) on the given
    comoving segments: exact mode sum over k_q = 2 pi q / L, L = 1.05 chi_s."""
    a, b = chi_edges[:-1], chi_edges[1:]
    L = 1.05 * chis
    Nmax = max(4, int(np.floor(L / Rs)))
    kq = 2 * np.pi * np.arange(1, Nmax + 1) / L
    P1D = p1d(cos, kq, Rs)
    Lh = b - a
    rw = np.sqrt(2 * P1D / L)
    cq = rw[:, None] * (np.sin(kq[:, None] * b) - np.sin(kq[:, None] * a)) / (kq[:, None] * Lh)
    sq = rw[:, None] * (np.cos(kq[:, None] * a) - np.cos(kq[:, None] * b)) / (kq[:, None] * Lh)
    return cq.T @ cq + sq.T @ sq


def rows(cos: Cosmology, zs: float, shape: Shape):
    """Per-(slice, mass) rows [n_z, n_m, 6] = (k_s, r_s, x_t, r_vir, w_row, beta)."""
    if shape.edge not in ("r200", "rvir"):
        raise ValueError(f"edge must be 'r200' or 'rvir', not {shape.edge!r}")
    z, chi_mid, dchi, chi_edges = slice_nodes(cos, zs, shape.n_z)
    M, dlnM = shape.m_nodes(cos.h)
    R = np.zeros((shape.n_z, shape.n_m, 6))
    for i, zi in enumerate(z):
        D = cos.Dg(zi)            # the bias growth (sgl_full.py:1861)
        wz = dchi * (1 + zi) ** 2
        for j, Mj in enumerate(M):
            C, rs, ks, fC = cos.nfw_params(Mj, zi, zs)
            dn = cos.dndlnM(Mj, zi) * dlnM
            rvir = cos.eta(C, zi) * C * rs
            beta = D * cos.halo_bias(Mj, zi)
            # truncation radius (sgl_full.py:1874-1880): r200 (x_t = c) or
            # r_vir (x_t = r_vir / r_s)
            xt = C if shape.edge == "r200" else rvir / rs
            R[i, j] = (ks, rs, xt, rvir, wz * dn, beta)
    # halo ellipticity eps(M, z) (Allgood+06 with M*(z), sgl_full.py:551-579)
    eps = np.array([cos.epsilon(M, zi) for zi in z]) if shape.ell else np.zeros(R.shape[:2])
    return dict(rows=R, z=z, chi=chi_mid, dchi=dchi, chi_edges=chi_edges, M=M, dlnM=dlnM,
                chis=cos.chi(zs), eps=eps)


# ----------------------------------------------------------- filaments --
FIL_U = 120      # u bins of a filament row, log-uniform from 1e-5 to 1 (sgl_full.py:1993)
FIL_RC = 0.1     # 2 r_s / L: the kink of the orientation rule (sgl_full.py:1828-1845)
P_FIL, Q_FIL, A_FIL = 0.0, 0.7, 0.5   # flat barrier of the filament mass function (filaments.py:39-50)


def fil_phi_rule(n: int = 12, rc: float = FIL_RC):
    """sgl_full.py:1828-1845: n Gauss-Legendre nodes on [0, acos rc] (where
    k0 depends on phi) plus one node carrying the clamped measure; weights
    are fractions of the quarter turn and sum to 1."""
    pc = math.acos(rc)
    x, w = np.polynomial.legendre.leggauss(n)
    ph = 0.5 * pc * (x + 1)
    wt = 0.5 * pc * w / (0.5 * math.pi)
    return (np.concatenate([ph, [0.5 * (pc + 0.5 * math.pi)]]),
            np.concatenate([wt, [(0.5 * math.pi - pc) / (0.5 * math.pi)]]))


def fil_u_edges():
    """[0, geomspace(1e-5, 1, FIL_U)] (sgl_full.py:1993): FIL_U bins, the
    first a disc of radius 1e-5."""
    return np.concatenate([[0.0], np.geomspace(1e-5, 1.0, FIL_U)])


def dndlnM_fil(cos: Cosmology, M, z):
    """filaments.py:43-50: the Sheth-Tormen form with (p, q, A) = (0, 0.7, 1/2)."""
    from .cosmo import DC
    nu = (DC / cos.sigmaM(M, z)) ** 2
    qn = Q_FIL * nu
    mult = A_FIL * 2.0 * math.sqrt(qn / (2 * math.pi)) * math.exp(-qn / 2)
    return (cos.rhom / M) * mult * (-2 * cos.dlnsig_dlnM(M))


def fil_bias(cos: Cosmology, M, z):
    """sgl_full.py:546-548: flat-barrier peak-background bias (q = 0.7)."""
    from .cosmo import DC0
    qn2 = 0.7 * (DC0 / cos.Dg(z) / cos.sigmaM(M, 0.0)) ** 2
    return 1 + (qn2 - 1) / DC0


def fil_rows(cos: Cosmology, zs: float, shape: Shape):
    """Per-(slice, mass) filament rows [n_z, n_m, 4] = (k0base, r_s, w_row, beta):
    r_s = (M / 1e14 Msun)^(1/3) Mpc, L = 20 r_s, k0base = r_s 14.4 rho_c0 L /
    Sigma_cr (k0 = k0base / max(2 r_s, |cos phi| L)), w_row = dchi (1+z)^2
    dn_fil, beta = D b_fil (sgl_full.py:1984-2002, filaments.py:69-77)."""
    z, chi_mid, dchi, chi_edges = slice_nodes(cos, zs, shape.n_z)
    M, dlnM = shape.m_nodes(cos.h)
    F = np.zeros((shape.n_z, shape.n_m, 4))
    for i, zi in enumerate(z):
        D = cos.Dg(zi)            # the bias growth (sgl_full.py:1861, 2001)
        wz = dchi * (1 + zi) ** 2
        scr = cos.Sigma_cr(zi, zs)
        for j, Mj in enumerate(M):
            rs = (Mj / 1e14) ** (1.0 / 3.0)
            k0base = rs * 14.4 * cos.rhoc0 * (20.0 * rs) / scr
            F[i, j] = (k0base, rs, wz * dndlnM_fil(cos, Mj, zi) * dlnM, D * fil_bias(cos, Mj, zi))
    return F


def cells_fil(F: np.ndarray):
    """Filament cell arrays [n_z, n_m, FIL_U * 13]: xi, ws, kap, gam (cell
    order: u bin, then orientation node)."""
    L = lib()
    n_rows = F.shape[0] * F.shape[1]
    ff = np.ascontiguousarray(F.reshape(n_rows, 4))
    ue = fil_u_edges()
    ph, wp = fil_phi_rule()
    n = FIL_U * ph.size
    out = [np.zeros(n_rows * n) for _ in range(4)]
    L.oracle_cells_fil(n_rows, _p(ff), FIL_U, _p(ue), ph.size, _p(np.ascontiguousarray(ph)),
                       _p(np.ascontiguousarray(wp)), *[_p(o) for o in out])
    return [o.reshape(F.shape[0], F.shape[1], n) for o in out]


N_PHI = 12       # orientation nodes of an elliptical halo cell (sgl_full.py:1535, 1659)


def cells(R: np.ndarray, shape: Shape, eps: np.ndarray = None):
    """Cell arrays [n_z, n_m, n_b] (or [n_z, n_m, n_b * N_PHI] with eps, the
    rows' ellipticities, as for shape.ell): xi, ws (tilted weight, 0 =
    dropped), kap, gam."""
    L = lib()
    n_rows = R.shape[0] * R.shape[1]
    rr = np.ascontiguousarray(R.reshape(n_rows, 6))
    ue = shape.u_edges()
    if eps is None:
        out = [np.zeros(n_rows * shape.n_b) for _ in range(4)]
        L.oracle_cells(n_rows, _p(rr), shape.n_b, _p(ue), *[_p(o) for o in out])
        return [o.reshape(R.shape[0], R.shape[1], shape.n_b) for o in out]
    ee = np.ascontiguousarray(np.asarray(eps, dtype=np.float64).reshape(n_rows))
    out = [np.zeros(n_rows * shape.n_b * N_PHI) for _ in range(4)]
    L.oracle_cells_ell(n_rows, _p(rr), _p(ee), shape.n_b, _p(ue), N_PHI, *[_p(o) for o in out])
    return [o.reshape(R.shape[0], R.shape[1], shape.n_b * N_PHI) for o in out]


def trunc_kg(x, ks, xt):
    """Truncated-NFW kappa, gamma at x (the C restatement, sgl_full.py:660-669)."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    k, g = np.zeros_like(x), np.zeros_like(x)
    lib().oracle_trunc_kg(x.size, _p(x), float(ks), float(xt), _p(k), _p(g))
    return k, g


def kg_eps(x, phi, eps, ks, xt):
    """Pseudo-elliptical truncated-NFW kappa, gamma at (x, phi) (sgl_full.py:582-613)."""
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    phi = np.ascontiguousarray(np.broadcast_to(np.asarray(phi, dtype=np.float64), x.shape))
    k, g = np.zeros_like(x), np.zeros_like(x)
    lib().oracle_kg_eps(x.size, _p(x), _p(phi), float(eps), float(ks), float(xt), _p(k), _p(g))
    return k, g


def lnphi(xi, ws, beta, k):
    """Lambda[n_k] (complex) and a[n_z, n_k] (complex) over the given k."""
    L = lib()
    n_z = xi.shape[0]
    per = xi[0].size
    xs = np.ascontiguousarray(xi.reshape(-1))
    w = np.ascontiguousarray(ws.reshape(-1))
    bw = np.ascontiguousarray((ws * beta[:, :, None]).reshape(-1))
    off = (np.arange(n_z + 1) * per).astype(np.int64)
    kk = np.ascontiguousarray(k, dtype=float)
    lam = np.zeros(2 * kk.size)
    amp = np.zeros(n_z * 2 * kk.size)
    L.oracle_lnphi(n_z, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), _p(xs), _p(w), _p(bw),
                   kk.size, _p(kk), _p(lam), _p(amp))
    lam = lam[0::2] + 1j * lam[1::2]
    amp = amp.reshape(n_z, kk.size, 2)
    return lam, amp[..., 0] + 1j * amp[..., 1]


def quadform(C, a):
    """Psi_2[n_k] = 1/2 sum_ij C_ij a_i a_j (complex square, no conjugate)."""
    return 0.5 * np.einsum("ik,ij,jk->k", a, C, a)


def invert(t, kq, phi):
    """p(t) for phi on nodes kq (phi[0] at k = 0)."""
    L = lib()
    tt = np.ascontiguousarray(t, dtype=float)
    kk = np.ascontiguousarray(kq, dtype=float)
    ph = np.ascontiguousarray(np.stack([phi.real, phi.imag], -1).reshape(-1))
    p = np.zeros(tt.size)
    L.oracle_invert(tt.size, _p(tt), kk.size, _p(kk), _p(ph), _p(p))
    return p


def filon_weights(t, kq):
    L = lib()
    tt = np.ascontiguousarray(t, dtype=float)
    kk = np.ascontiguousarray(kq, dtype=float)
    W = np.zeros((tt.size, 2 * kk.size))
    L.oracle_filon_weights(tt.size, _p(tt), kk.size, _p(kk), _p(W))
    return W


def window(kq):
    """Cubic B-spline in k of support [0, K = kq[-1]], 1 at k = 0
    (sgl_full.py:83-94, the k_kappa factor of _window_pd)."""
    K = float(kq[-1])
    x = np.abs(kq) * 2.0 / K
    return np.where(x < 1, 2 / 3 - x ** 2 + x ** 3 / 2, np.where(x < 2, (2 - x) ** 3 / 6, 0.0)) / (2 / 3)


def moments(xw, pw):
    """Diagnostics of p on its grid: sgl_full.py:362-371 (mean, var, skew,
    sigma_DL with the grid mass as norm) plus <mu>, P(mu > 2) and the
    negative mass of api.py:75."""
    norm = float(np.trapezoid(pw, xw))
    ps = pw / norm
    mean = float(np.trapezoid(xw * ps, xw))
    var = float(np.trapezoid((xw - mean) ** 2 * ps, xw))
    skew = float(np.trapezoid((xw - mean) ** 3 * ps, xw) / var ** 1.5)
    x = np.exp(-0.5 * xw)
    md = float(np.trapezoid(x * ps, xw))
    sdl = float(np.sqrt(np.trapezoid((x - md) ** 2 * ps, xw)) / md)
    mu_mean = float(np.trapezoid(np.exp(xw) * ps, xw))
    tail = float(tail_mass(xw, ps, math.log(2.0)))
    neg = float(np.trapezoid(np.maximum(-pw, 0.0), xw))
    return dict(norm=norm, mean=mean, var=var, skew=skew, sigmaDL=sdl, mu_mean=mu_mean,
                p_mu_gt_2=tail, negative_mass=neg)


def tail_mass(x, p, x0):
    """Trapezoid of p over [x0, x_end] with linear interpolation at x0."""
    i = int(np.searchsorted(x, x0, side="right"))
    if i >= x.size:
        return 0.0
    if i == 0:
        return float(np.trapezoid(p, x))
    f = (x0 - x[i - 1]) / (x[i] - x[i - 1])
    p0 = p[i - 1] + f * (p[i] - p[i - 1])
    return float(0.5 * (p0 + p[i]) * (x[i] - x0) + np.trapezoid(p[i:], x[i:]))


@dataclass
class Result:
    x: np.ndarray
    p: np.ndarray
    info: dict = field(default_factory=dict)


def vp_inputs(xi, ws, beta):
    """Closure inputs of the cells [n_z, n_m, n_cells]: Y_sg(k) needs the cells,
    wk_sg = sum w xi per (slice, mass row)."""
    return np.sum(ws * xi, axis=2)


def pdf(zs: float, shape: Shape = None, cosmo: dict = None, keep: bool = False,
        closure: str = "gauss") -> Result:
    """End-to-end FP64 oracle PDF on the output grid; closure "gauss" (Psi_2)
    or "vp" (the positive closure of oracle/closure.py)."""
    shape = shape or Shape()
    cos = Cosmology(**(cosmo or {}))
    T = rows(cos, zs, shape)
    R = T["rows"]
    xi, ws, kap, gam = cells(R, shape, T["eps"] if shape.ell else None)
    kq = shape.k_nodes()
    lam, amp = lnphi(xi, ws, R[:, :, 5], kq[1:])
    if shape.filaments:
        F = fil_rows(cos, zs, shape)
        fxi, fws, _, _ = cells_fil(F)
        lam_f, amp_f = lnphi(fxi, fws, F[:, :, 3], kq[1:])
        lam, amp = lam + lam_f, amp + amp_f
    C = slice_covariance(cos, T["chi_edges"], T["chis"])
    if closure == "vp":
        from .closure import VP
        # groups: the halo mass rows, then (with filaments) the filament rows
        parts = [(xi, ws, R[:, :, 5])]
        if shape.filaments:
            parts.append((fxi, fws, F[:, :, 3]))
        n_z = R.shape[0]
        beta = np.concatenate([p[2] for p in parts], axis=1)
        wk = np.concatenate([vp_inputs(*p) for p in parts], axis=1)
        # Y_sg(k): every row as its own "slice" of the amplitude sum, beta = 1
        Y = []
        for x_, w_, b_ in parts:
            n_g = b_.shape[1]
            _, Yr = lnphi(x_.reshape(n_z * n_g, 1, -1), w_.reshape(n_z * n_g, 1, -1),
                          np.ones((n_z * n_g, 1)), kq[1:])
            Y.append(Yr.reshape(n_z, n_g, -1))
        vp = VP(beta, wk, C)
        psi2 = vp.lnR(np.concatenate(Y, axis=1))
    else:
        psi2 = quadform(C, amp) if shape.clustering else np.zeros_like(lam)
    phi = np.concatenate([[1.0 + 0j], np.exp(lam + psi2)])
    if shape.window:
        phi = phi * window(kq)
    xo = shape.out_grid()
    po = invert(xo, kq, phi)
    mom = moments(xo, po)
    info = dict(moments=mom, n_cells=int(np.count_nonzero(ws)), sum_w=float(ws.sum()),
                var_xi=float(np.sum(ws * xi ** 2)))
    if shape.filaments:
        info.update(n_cells_fil=int(np.count_nonzero(fws)), sum_w_fil=float(fws.sum()),
                    var_xi_fil=float(np.sum(fws * fxi ** 2)))
    if closure == "vp":
        info.update(eta=vp.eta, vp_target=vp.target)
    if keep:
        info.update(cos=cos, tables=T, xi=xi, ws=ws, kap=kap, gam=gam, lam=lam, amp=amp, C=C,
                    psi2=psi2, phi=phi, kq=kq)
    return Result(xo, po, info)
