"""FP64 CPU restatement of the projected NFW lens (checker only).

TEST INFRASTRUCTURE: imported only by tests/, smoke() and bench.py's
cpu_baseline leg.

* ``nfw_Fh``: Wright & Brainerd F(x), h(x) with the reference's series and
  cancellation-free branches (sgl.py:618-720).
* ``trunc_corr``: the sharply truncated NFW correction (I1, I2) of
  sgl_full.py:612-816 (``_tc_point``: small-a expansion, closed form with the
  exact arccosh/arccos arguments, a linear bridge over |x-1| < 1e-5, GL
  fallback inside the bridge).
* ``nfw_trunc_kg``: (kappa, gamma) of the truncated profile, sgl_full.py:660-669.

Pinned in tests/test_oracle_nfw.py against mpmath quadrature of the 3-D
density (untruncated and truncated line-of-sight integrals) -- the reference's
own accuracy claims are <= 8.9e-16 (sgl.py:633-637) and <= 1e-10
(sgl_full.py:624-626).
"""
from __future__ import annotations

import math

import numpy as np

NFW_W_SERIES = 0.25          # sgl.py:676
NFW_N_SERIES = 28            # sgl.py:677
NFW_X_H = 0.9                # sgl.py:678
TC_A_SMALL = 1e-3            # sgl_full.py:631
TC_BAND = 1e-5               # sgl_full.py:632
_GLX, _GLW = np.polynomial.legendre.leggauss(24)        # sgl_full.py:634-635
_GLX, _GLW = 0.5 * (_GLX + 1.0), 0.5 * _GLW


def nfw_Fh(x: float):
    """(F, h) at one point; sgl.py:694-720 branch for branch."""
    w = (1.0 - x) * (1.0 + x)
    if abs(w) < NFW_W_SERIES:
        T = 0.0
        for n in range(NFW_N_SERIES - 1, -1, -1):
            T = T * w + 1.0 / (2 * n + 3)
        return T, math.log(0.5 * x) + (1.0 + w * T)
    if x < 1.0:
        e = (1.0 - x) / x
        s = math.sqrt(w)
        A = math.log1p(e + math.sqrt(e * (e + 2.0)))
        F = (A / s - 1.0) / w
        if x < NFW_X_H:
            h = (math.log(2.0 / x) * x * x / (1.0 + s) + math.log1p(-x * x / (2.0 * (1.0 + s)))) / s
        else:
            h = math.log(0.5 * x) + A / s
        return F, h
    v = -w
    sv = math.sqrt(v)
    B = math.atan(sv)
    return (1.0 - B / sv) / v, math.log(0.5 * x) + B / sv


def kappa_gamma(x, ks):
    """Untruncated NFW (kappa, gamma), sgl.py:722-740."""
    x = np.atleast_1d(np.asarray(x, float))
    k = np.empty_like(x)
    g = np.empty_like(x)
    for i, xi in enumerate(x):
        F, h = nfw_Fh(float(xi))
        k[i] = 2.0 * ks * F
        g[i] = 4.0 * ks * h / (xi * xi) - k[i]
    return k, g


def _tc_closed_raw(x, c):
    """sgl_full.py:740-762: I1 = F - F_t, I2 = (h - m_t)/x^2, 0 < x < c, x != 1."""
    q = math.sqrt((c - x) * (c + x))
    d = (x - 1.0) * (x - c) / (x * (1.0 + c))
    L = math.log((c + q) / (2.0 * (1.0 + c))) - (q - c) / (1.0 + c)
    if x < 1.0:
        u = (1.0 - x) * (1.0 + x)
        e = (1.0 - x) / x
        A0 = math.log1p(e + math.sqrt(e * (e + 2.0)))
        At = math.log1p(d + math.sqrt(d * (d + 2.0)))
        su = math.sqrt(u)
        return ((q / (1.0 + c) - 1.0) / u + (A0 - At) / (u * su), (L + (A0 - At) / su) / (x * x))
    v = (x - 1.0) * (x + 1.0)
    B0 = 2.0 * math.asin(math.sqrt((x - 1.0) / (2.0 * x)))
    Bt = 2.0 * math.asin(math.sqrt(max(-d, 0.0) / 2.0))
    sv = math.sqrt(v)
    return ((1.0 - q / (1.0 + c)) / v - (B0 - Bt) / (v * sv), (L + (B0 - Bt) / sv) / (x * x))


def _tc_gl(a, xt):
    """sgl_full.py:764-777: 24-node Gauss-Legendre fallback."""
    s1 = s2 = 0.0
    for v, wv in zip(_GLX, _GLW):
        s = 1.0 - v * v
        d = max(1.0 - a * a * s * s, 1e-300)
        s1 += wv * 2.0 * v * s / ((s + xt) ** 2 * math.sqrt(d))
        e = max(1.0 - a * a * v * v, 0.0)
        s2 += wv * v / ((v + xt) ** 2 * (1.0 + math.sqrt(e)))
    return s1, s2


def trunc_corr(x, xt):
    """sgl_full.py:779-798 (_tc_point)."""
    a = x / xt
    if a >= 1.0:
        return 0.0, 0.0
    if a < TC_A_SMALL:
        i0 = math.log((1.0 + xt) / xt) - 1.0 / (1.0 + xt)
        u1 = 1.0 + xt
        jj = (0.5 * (u1 * u1 - xt * xt) - 3.0 * xt * (u1 - xt)
              + 3.0 * xt * xt * math.log(u1 / xt) + xt ** 3 * (1.0 / u1 - 1.0 / xt))
        return i0 + 0.5 * a * a * jj, 0.5 * i0 + 0.125 * a * a * jj
    if abs(x - 1.0) < TC_BAND:
        if xt <= 1.0 + 2.0 * TC_BAND:
            return _tc_gl(a, xt)
        a1, b1 = _tc_closed_raw(1.0 - TC_BAND, xt)
        a2, b2 = _tc_closed_raw(1.0 + TC_BAND, xt)
        f = (x - (1.0 - TC_BAND)) / (2.0 * TC_BAND)
        return a1 + f * (a2 - a1), b1 + f * (b2 - b1)
    return _tc_closed_raw(x, xt)


def nfw_trunc_kg_scalar(x, ks, xt):
    """sgl_full.py:660-669 at one point."""
    F, h = nfw_Fh(x)
    ku = 2.0 * F
    gu = 4.0 * h / (x * x) - ku
    if x < xt:
        I1, I2 = trunc_corr(x, xt)
        k = ku - 2.0 * I1
        kb = ku + gu - 4.0 * I2
    else:
        mut = math.log(1.0 + xt) - xt / (1.0 + xt)
        k = 0.0
        kb = 4.0 * mut / max(x, 1e-300) ** 2
    return ks * k, ks * (kb - k)


def nfw_trunc_kg(x, ks, xt):
    x = np.atleast_1d(np.asarray(x, float))
    out = np.array([nfw_trunc_kg_scalar(float(v), ks, xt) for v in x])
    return out[:, 0], out[:, 1]
