"""FP64 CPU restatement of the reference's cosmology sector (checker only).

TEST INFRASTRUCTURE.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package; the product path
(``sgl_analytic_amd``) never does.

Restates ``sgl.Cosmology`` on the configuration ``pipeline.make_cosmology``
builds (sgl_analytic/pipeline.py:36-72): window="smoothk", transfer="eh98",
anchor="tophat", conc_model=16, with growth_mode "ode" (default) or "legacy"
and radiation "engine" (default) or None.  The same scipy routines the
reference calls (quad, solve_ivp DOP853, CubicSpline) are used, so this is the
reference algorithm, not an approximation of it.

Pins (tests/test_oracle_cosmo.py): fixture ``dark_energy.Omega_r`` and
``primordial_growth`` (tests/reference/pdf_zs1.json); walkthrough cell 3
(sigma8 top-hat, D(1), D(3)) and cell 7 (c, r_s, kappa_s, r_vir/r_s); closed
forms (constant-w growth vs 2F1, EdS growth D=a, flat distances vs quadrature).
"""
from __future__ import annotations

import numpy as np
from numpy import exp, log, pi, sqrt
from scipy.integrate import quad, solve_ivp
from scipy.interpolate import CubicSpline
from scipy.special import gamma as Gamma

CKMS = 299792.458            # km/s              (sgl.py:56)
RHOC0_UNIT = 2.77536627e11   # Msun/Mpc^3/h^2    (sgl.py:57)
GNEWT = 4.30091e-9           # Mpc (km/s)^2/Msun (sgl.py:58)
TCMB = 2.7255                # sgl.py:52-53 DEFAULTS
DC = 1.686                   # sgl.py:52-53 DEFAULTS (delta_c of Sheth-Tormen)
DC0 = 3.0 / 5.0 * (1.5 * pi) ** (2.0 / 3.0)   # sgl_full.py:389
P_ST, Q_ST = 0.3, 0.8                          # sgl.py:555
A_ST = 1.0 / (1 + 2 ** -0.3 * Gamma(0.5 - 0.3) / sqrt(pi))   # sgl.py:556

PLANCK = dict(h=0.674, Om=0.315, s8=0.811, Ob=0.0493, ns=0.965, zeq=3402.0)  # pipeline.py:20


class Cosmology:
    """sgl.Cosmology restricted to the make_cosmology configuration."""

    def __init__(self, h=0.674, Om=0.315, s8=0.811, Ob=0.0493, ns=0.965,
                 zeq=3402.0, Ok=0.0, w0=-1.0, wa=0.0, radiation="engine", growth_mode="ode"):
        # sgl.py:68-167 (defaults, radiation closure, DE domain checks, growth
        # mode; "auto" is resolved by the caller)
        if growth_mode not in ("ode", "legacy"):
            raise ValueError("growth_mode must be 'ode' or 'legacy'")
        if growth_mode == "legacy" and not (w0 == -1.0 and wa == 0.0):
            raise ValueError("legacy growth is only valid for LambdaCDM")
        self.growth_mode = growth_mode
        self.kw = dict(h=h, Om=Om, s8=s8, Ob=Ob, ns=ns, zeq=zeq, Ok=Ok, w0=w0, wa=wa,
                       radiation=radiation, growth_mode=growth_mode)
        self.h, self.Om, self.s8, self.Ob, self.ns = h, Om, s8, Ob, ns
        self.zeq, self.Ok, self.w0, self.wa = float(zeq), float(Ok), float(w0), float(wa)
        self.Or = self.Om / (1.0 + self.zeq) if radiation == "engine" else 0.0   # sgl.py:100
        self.OL = 1.0 - self.Om - self.Ok - self.Or    # sgl.py:101
        if not (-2.0 <= self.w0 <= 1.0):
            raise ValueError("dark energy requires -2 <= w0 <= 1")
        if self.w0 + self.wa >= 0.0:
            raise ValueError("dark energy requires w0 + wa < 0")
        if self.de_matter_ratio(1100.0) > 1e-2:
            raise ValueError("dark energy exceeds 1e-2 of matter at z = 1100")
        e_grid = self.E2(1.0 / np.geomspace(1e-6, 1.0, 4096) - 1.0)
        if not np.all(e_grid > 0):
            raise ValueError("E(z)^2 <= 0 in 0 <= z <= 1e6")
        self.H0 = 100.0 * h
        self.rhoc0 = RHOC0_UNIT * h * h
        self.rhom = self.Om * self.rhoc0
        self._chi_cache = {}
        if growth_mode == "ode":
            self._build_growth_ode()
        else:
            self._build_growth_legacy()
        self._build_power()
        self._build_sigma()

    # ---------------------------------------------------------- background --
    def de_matter_ratio(self, z):                     # sgl.py:176-183
        z = np.asarray(z, float)
        return (self.OL / self.Om * (1 + z) ** (3 * (self.w0 + self.wa))
                * np.exp(-3 * self.wa * z / (1 + z)))

    def E2(self, z):                                  # sgl.py:185-197
        zp1 = 1 + np.asarray(z, float)
        if self.wa == 0.0:
            de = self.OL if self.w0 == -1.0 else self.OL * zp1 ** (3 * (1 + self.w0))
        else:
            de = (self.OL * zp1 ** (3 * (1 + self.w0 + self.wa))
                  * exp(-3 * self.wa * (1 - 1 / zp1)))
        return self.Om * zp1 ** 3 + self.Or * zp1 ** 4 + self.Ok * zp1 ** 2 + de

    def E(self, z):
        return sqrt(self.E2(z))

    def Hz(self, z):
        return self.H0 * self.E(z)

    def rhocrit(self, z):                             # sgl.py:202-203
        return self.rhoc0 * self.E2(z)

    def chi(self, z):                                 # sgl.py:205-217 (adaptive quad)
        key = float(z)
        v = self._chi_cache.get(key)
        if v is None:
            v = quad(lambda zz: CKMS / self.H0 / self.E(zz), 0, key)[0]
            self._chi_cache[key] = v
        return v

    def f_K(self, chi):                               # sgl.py:219-238
        if self.Ok == 0.0:
            return chi
        radial = np.asarray(chi, float)
        u = radial * self.H0 / CKMS
        t = self.Ok * u ** 2
        x = sqrt(abs(self.Ok)) * u
        with np.errstate(invalid="ignore", divide="ignore"):
            ratio = np.sinh(x) / x if self.Ok > 0 else np.sin(x) / x
        ratio = np.where(abs(t) < 1e-5, 1 + t / 6 + t ** 2 / 120 + t ** 3 / 5040, ratio)
        out = radial * ratio
        return float(out) if out.ndim == 0 else out

    def DA(self, z):                                  # sgl.py:240-244
        return self.f_K(self.chi(z)) / (1 + z)

    def DA_ls(self, zl, zs):                          # sgl.py:249-250
        return self.f_K(self.chi(zs) - self.chi(zl)) / (1 + zs)

    def Sigma_cr(self, zl, zs):                       # sgl.py:252-254
        return CKMS ** 2 * self.DA(zs) / (4 * pi * GNEWT * self.DA(zl) * self.DA_ls(zl, zs))

    # -------------------------------------------------------------- growth --
    def _build_growth_ode(self):                      # sgl.py:310-347
        w, wa, Or = self.w0, self.wa, self.Or
        x0 = log(1e-6)

        def rhs(x, y):
            matter = self.Om * exp(-3 * x)
            a = exp(x)
            de = self.OL * exp(-3 * (1 + w + wa) * x + 3 * wa * (a - 1))
            wz = w + wa * (1 - a)
            curvature = self.Ok * exp(-2 * x)
            rad = Or * exp(-4 * x)
            total = matter + rad + curvature + de
            dh = -0.5 * (3 * matter + 4 * rad + 2 * curvature + 3 * (1 + wz) * de) / total
            return (y[1], -(2 + dh) * y[1] + 1.5 * matter / total * y[0])

        d0 = exp(x0) + (2.0 / 3.0) * Or / self.Om
        sol = solve_ivp(rhs, (x0, 0.0), (d0, exp(x0)), method="DOP853",
                        rtol=2e-10, atol=1e-14, dense_output=True)
        if not sol.success:
            raise RuntimeError("growth integration failed: " + sol.message)
        self._growth_solution = sol
        self._D0 = float(sol.y[0, -1])

    def _build_growth_legacy(self):                   # sgl.py:274-306 (Heath integral)
        OLg = 1.0 - self.Om - self.Ok       # the integral never carries radiation

        def Ea(a):
            a = np.asarray(a, float)
            if self.Ok == 0.0:
                return sqrt(self.Om * a ** -3 + OLg)
            return sqrt(self.Om * a ** -3 + self.Ok * a ** -2 + OLg)

        ag = np.geomspace(1e-6, 1.0, 4096)
        seg = np.empty(ag.size)
        seg[0] = quad(lambda ap: 1.0 / (ap * Ea(ap)) ** 3, 0.0, ag[0])[0]
        for i in range(1, ag.size):
            seg[i] = quad(lambda ap: 1.0 / (ap * Ea(ap)) ** 3, ag[i - 1], ag[i])[0]
        lnI = CubicSpline(log(ag), log(np.cumsum(seg)))
        self._Dun = lambda a: Ea(a) * exp(lnI(log(np.asarray(a, float))))
        self._D0 = float(self._Dun(1.0))

    def primordial_growth_amplitude(self):            # sgl.py:349-352
        return self._D0 if self.growth_mode == "ode" else self._D0 * (2.5 * self.Om)

    def D(self, z):                                   # sgl.py:354-375
        a = 1.0 / (1 + np.asarray(z, float))
        if self.growth_mode == "ode":
            out = self._growth_solution.sol(log(a))[0] / self._D0
        else:
            out = self._Dun(a) / self._D0
        return float(out) if np.ndim(z) == 0 else out

    def Dg(self, z):                                  # sgl_full.py:521-538
        """The growth in the bias: D on the ODE growth, otherwise the
        Carroll-Press-Turner fit with E^2 including radiation when present."""
        if self.growth_mode == "ode":
            return self.D(z)

        def nn(zz):
            a3 = self.Om * (1 + zz) ** 3
            E2 = a3 + self.OL
            if self.Ok != 0.0:
                E2 += self.Ok * (1 + zz) ** 2
            if self.Or != 0.0:
                E2 += self.Or * (1 + zz) ** 4
            Omz, OLz = a3 / E2, self.OL / E2
            return 2.5 * Omz / (Omz ** (4 / 7) - OLz + (1 + Omz / 2) * (1 + OLz / 70)) / (1 + zz)
        return nn(z) / nn(0.0)

    # ------------------------------------------------- transfer, sigma(M) --
    def T(self, k):                                   # sgl.py:391-445 (_T_eh98)
        h, Om, Ob = self.h, self.Om, self.Ob
        Oc = Om - Ob
        wm, wb = Om * h ** 2, Ob * h ** 2
        zeq = self.zeq
        keq = (self.H0 / CKMS) * sqrt(2.0 * Om) * sqrt(1.0 + zeq)
        ksilk = 0.0016 * wb ** 0.52 * wm ** 0.73 * (1 + (10.4 * wm) ** -0.95)
        a1 = (46.9 * wm) ** 0.67 * (1 + (32.1 * wm) ** -0.532)
        a2 = (12.0 * wm) ** 0.424 * (1 + (45.0 * wm) ** -0.582)
        b1 = 0.944 / (1 + (458 * wm) ** -0.708)
        b2 = (0.395 * wm) ** -0.026
        ac = a1 ** (-Ob / Om) * a2 ** (-(Ob / Om) ** 3)
        bc = 1.0 / (1 + b1 * ((Oc / Om) ** b2 - 1))
        s = 44.5 * log(9.83 / wm) / sqrt(1 + 10 * wb ** 0.75)
        b3 = 0.313 * wm ** -0.419 * (1 + 0.607 * wm ** 0.674)
        b4 = 0.238 * wm ** 0.223
        zd = 1291 * wm ** 0.251 / (1 + 0.659 * wm ** 0.828) * (1 + b3 * wb ** b4)
        Rd = 31.5 * wb * (TCMB / 2.7) ** -4 / (zd / 1000.0)

        def g2(y):
            return y * (-6 * sqrt(1 + y) + (2.0 + 3.0 * y) * log((sqrt(1 + y) + 1) / (sqrt(1 + y) - 1)))

        ab = 2.07 * keq * s * (1 + Rd) ** -0.75 * g2((1 + zeq) / (1 + zd))
        bb = 0.5 + Ob / Om + (3.0 - 2.0 * Ob / Om) * sqrt(1 + (17.2 * wm) ** 2)
        bnode = 8.41 * wm ** 0.435
        fk = 1.0 / (1 + (k * s / 5.4) ** 4)

        def To1(kk, a_c, b_c):
            qq = kk / (13.41 * keq)
            C1 = 14.2 / a_c + 386.0 / (1 + 69.9 * qq ** 1.08)
            num = log(np.e + 1.8 * b_c * qq)
            return num / (num + C1 * qq * qq)

        TC = fk * To1(k, 1.0, bc) + (1 - fk) * To1(k, ac, bc)
        s3 = s / (1 + (bnode / (k * s)) ** 3) ** (1.0 / 3.0)
        x = k * s3
        jo = np.where(x < 1e-8, 1.0 - x * x / 6.0, np.sin(x) / np.where(x == 0, 1, x))
        TB = ((To1(k, 1.0, 1.0) / (1 + (k * s / 5.2) ** 2)
               + ab / (1 + (bb / (k * s)) ** 3) * exp(-(k / ksilk) ** 1.4)) * jo)
        return Ob / Om * TB + Oc / Om * TC

    def _build_power(self):                           # sgl.py:458-460, 474-476
        self._kg = np.logspace(-5, 3, 4000)
        self._Pshape = self._kg ** self.ns * self.T(self._kg) ** 2

    def _sigmaR_shape(self, R, window="smoothk"):    # sgl.py:510-515
        x = self._kg * R
        if window == "tophat":
            Wk = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
        else:
            Wk = 1.0 / (1.0 + (0.43 * x) ** 6)
        return sqrt(np.trapezoid(self._kg ** 2 * self._Pshape * Wk ** 2, self._kg) / (2 * pi * pi))

    def _build_sigma(self):                           # sgl.py:517-538 (anchor="tophat")
        self._Anorm = self.s8 / self._sigmaR_shape(8.0 / self.h, window="tophat")
        Mg = np.logspace(4, 17, 500)
        lnsig = np.array([log(self._Anorm * self._sigmaR_shape(self.RofM(M))) for M in Mg])
        self._spl = CubicSpline(log(Mg), lnsig)

    def sigma_tophat8(self):                          # sgl.py:543-546
        return self._Anorm * self._sigmaR_shape(8.0 / self.h, window="tophat")

    def RofM(self, M):                                # sgl.py:548-549
        return (3 * M / (4 * pi * self.rhom)) ** (1.0 / 3.0)

    def sigmaM(self, M, z=0.0):                       # sgl.py:551-552
        return exp(self._spl(log(M))) * self.D(z)

    def dlnsig_dlnM(self, M):                         # sgl.py:554-555
        return self._spl(log(M), 1)

    def dndlnM(self, M, z):                           # sgl.py:558-567 Sheth-Tormen
        nu = (DC / self.sigmaM(M, z)) ** 2
        qn = Q_ST * nu
        mult = A_ST * (1 + qn ** -P_ST) * sqrt(qn / (2 * pi)) * exp(-qn / 2)
        return (self.rhom / M) * mult * (-2 * self.dlnsig_dlnM(M))

    # ------------------------------------------------------------------ NFW --
    def conc(self, M, z):                             # sgl.py:570-606, conc_model=16
        if z > 7.0:
            return self.conc(M, 7.0)
        a = 1.0 + z
        poly = 4.135 - 0.564 * a - 0.210 * a ** 2 + 0.0557 * a ** 3 - 0.00348 * a ** 4
        q = DC / (self.sigmaM(M, 0.0) * poly)
        c0 = 3.395 * a ** -0.215
        beta = 0.307 * a ** 0.540
        g1 = 0.628 * a ** -0.047
        g2 = 0.317 * a ** -0.893
        return c0 * q ** -g1 * (1.0 + q ** (1.0 / beta)) ** (-beta * (g2 - g1))

    def nfw_params(self, M, zl, zs):                  # sgl.py:608-615
        C = self.conc(M, zl)
        r200 = (3 * M / (4 * pi * 200 * self.rhocrit(zl))) ** (1.0 / 3.0)
        rs = r200 / C
        fC = log(1 + C) - C / (1 + C)
        rhos = (200.0 / 3.0) * self.rhocrit(zl) * C ** 3 / fC
        ks = rhos * rs / self.Sigma_cr(zl, zs)
        return C, rs, ks, fC

    # ------------------------------------------------------- bias, virial --
    def halo_bias(self, M, z):                        # sgl_full.py:541-543
        qn2 = 0.8 * (DC0 / self.Dg(z) / self.sigmaM(M, 0.0)) ** 2
        return 1 + (qn2 - 1) / DC0 + 0.6 / (DC0 * (1 + qn2 ** 0.3))

    def Mstar(self, z):                               # sgl_full.py:551-567
        """sigma(M*, 0) Dg(z) = delta_c0: log-interpolated root on the
        reference's table M = logspace(0, 17, 1200) Msun of the sigma spline
        (extrapolated below its 1e4 node as the reference does), cut to the
        largest suffix on which sigma strictly decreases."""
        tab = getattr(self, "_mstar_tab", None)
        if tab is None:
            Mg = np.logspace(0, 17, 1200)
            s = np.array([self.sigmaM(m, 0.0) for m in Mg])
            up = np.nonzero(np.diff(s) >= 0)[0]
            i0 = int(up[-1]) + 1 if up.size else 0
            tab = self._mstar_tab = (Mg[i0:], s[i0:])
        Mg, s = tab
        return float(np.exp(np.interp(-DC0 / self.Dg(z), -s, log(Mg))))

    def epsilon(self, M, z):                          # sgl_full.py:570-579 (Allgood+06)
        s = 0.54 * (np.asarray(M, float) / self.Mstar(z)) ** -0.05
        return np.maximum(0.0, (1 - s) / (1 + s))

    def Dvir(self, z):                                # subhalos.py:381-385 (BN98 at Om(z))
        d = self.Om * (1 + z) ** 3 / self.E2(z) - 1.0
        return 18 * pi ** 2 + 82 * d - 39 * d * d

    def eta(self, c, z):                              # subhalos.py:472-481 (r_vir / r200)
        mu = lambda y: log(1 + y) - y / (1 + y)
        target = 200.0 * c ** 3 / mu(c) / self.Dvir(z)
        lo, hi = c, 3 * c
        for _ in range(80):
            m = 0.5 * (lo + hi)
            if m ** 3 / mu(m) < target:
                lo = m
            else:
                hi = m
        return 0.5 * (lo + hi) / c

    def P0(self, k):                                  # sgl_full.py:2428-2432
        return self._Anorm ** 2 * k ** self.ns * self.T(k) ** 2
