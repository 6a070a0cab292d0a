// Device-side FP64 physics shared by the table, cell and exponent kernels.
//
// Cosmology: the configuration sgl_analytic/pipeline.py:36-72 builds
// (smooth-k sigma(M) window, Eisenstein & Hu 1998 transfer, top-hat sigma8
// anchor, Ludlow+16 concentrations, growth ODE, radiation Omega_r = Om/(1+zeq)).
// Lens: projected NFW of Wright & Brainerd with the series / cancellation-free
// branches of sgl.py:694-720, truncated at x_t with the closed form of
// sgl_full.py:740-798.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

namespace sgl {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kCkms = 299792.458;          // km/s
constexpr double kRhoc0Unit = 2.77536627e11;  // Msun / Mpc^3 / h^2
constexpr double kGnewt = 4.30091e-9;         // Mpc (km/s)^2 / Msun
constexpr double kTcmb = 2.7255;
constexpr double kDeltaC = 1.686;             // Sheth-Tormen delta_c
constexpr double kPst = 0.3, kQst = 0.8;

// Cosmological parameters of one instance, as laid out in HBM (16 doubles).
struct Cosmo {
  double h, Om, s8, Ob, ns, zeq, Ok, w0, wa, zs;
  double n_z;            // slice count of this instance (stored as double)
  double rad_none;       // SGL_REC_RADIATION: 1 = no radiation
  double growth_legacy;  // SGL_REC_GROWTH: 1 = Heath-integral D, CPT growth in the bias
  double zint;           // SGL_REC_ZINT: 0 midpoint, 1 engine, 2 gauss (reference grid only)
  double zint_n;         // SGL_REC_ZINT_N: Gauss-Legendre nodes per shell with zint 2
  double pad[1];
};
static_assert(offsetof(Cosmo, rad_none) == SGL_REC_RADIATION * sizeof(double), "record slot");
static_assert(offsetof(Cosmo, growth_legacy) == SGL_REC_GROWTH * sizeof(double), "record slot");
static_assert(offsetof(Cosmo, zint) == SGL_REC_ZINT * sizeof(double), "record slot");
static_assert(offsetof(Cosmo, zint_n) == SGL_REC_ZINT_N * sizeof(double), "record slot");

// Derived per-instance scalars (16 doubles), written by the background kernel.
struct Derived {
  double Or, OL, H0, rhoc0, rhom, chis, DAs, Anorm;
  double D0, dchi, n_z, pad[5];
};

__device__ __forceinline__ double e2_of_z(const Cosmo& c, double Or, double OL, double z) {
  // E(z)^2 for CPL dark energy, curvature and radiation (sgl.py:185-197)
  double zp1 = 1.0 + z;
  double de;
  if (c.wa == 0.0)
    de = (c.w0 == -1.0) ? OL : OL * pow(zp1, 3.0 * (1.0 + c.w0));
  else
    de = OL * pow(zp1, 3.0 * (1.0 + c.w0 + c.wa)) * exp(-3.0 * c.wa * (1.0 - 1.0 / zp1));
  double z2 = zp1 * zp1;
  return c.Om * z2 * zp1 + Or * z2 * z2 + c.Ok * z2 + de;
}

// Transverse comoving distance f_K(chi) (sgl.py:219-238).
__device__ __forceinline__ double f_k(double Ok, double H0, double chi) {
  if (Ok == 0.0) return chi;
  double u = chi * H0 / kCkms;
  double t = Ok * u * u;
  if (fabs(t) < 1e-5) return chi * (1.0 + t / 6.0 + t * t / 120.0 + t * t * t / 5040.0);
  double x = sqrt(fabs(Ok)) * u;
  return chi * ((Ok > 0.0) ? sinh(x) / x : sin(x) / x);
}

// r_vir / r200 = eta solving m^3 / mu(m) = 200 c^3 / (mu(c) Delta_vir) on
// [c, 3c] (subhalos.py:409-441): safeguarded Newton on g(m) = m^3 - T mu(m),
// T = 200 c^3 / (mu(c) Delta_vir) (one log1p and two divisions a step).
// mu_c = mu(c) and cr = cbrt(200 / Delta_vir) come from the caller; Newton
// starts at m = c cr (the root if mu(m) = mu(c), a few per cent off).
__device__ inline double eta_rvir(double c, double Dvir, double mu_c, double cr) {
  const double T = 200.0 * c * c * c / (mu_c * Dvir);
  double lo = c, hi = 3.0 * c, m = fmin(fmax(c * cr, lo), hi);
  for (int it = 0; it < 80; ++it) {
    const double iop = 1.0 / (1.0 + m);
    const double mu = log1p(m) - m * iop;
    const double g = m * m * m - T * mu;
    const double gp = 3.0 * m * m - T * m * iop * iop;
    const double step = g / gp;
    // Newton is quadratic here: a step below 1e-13 m leaves an error far below
    // 1 ulp.  The test comes before the bracket update, so that a point where
    // g rounds to 0 (the step lands on the new bracket end) ends the solve
    // instead of falling back to bisection for the last bits.
    if (fabs(step) <= 1e-13 * m) { m -= step; break; }
    if (g < 0.0) lo = m; else hi = m;
    double mn = m - step;
    if (!(lo < mn && mn < hi)) mn = 0.5 * (lo + hi);
    if (hi - lo <= 4e-16 * m) { m = mn; break; }
    m = mn;
  }
  return m / c;
}

// Carroll-Press-Turner growth g(z) / (1 + z) with E^2 including radiation when
// the background has it (sgl_full.py:521-538); Dg(z) = cpt(z) / cpt(0).
__device__ inline double cpt_growth(double z, double Om, double Ok, double Or, double OL) {
  const double zp1 = 1.0 + z;
  const double m3 = Om * zp1 * zp1 * zp1;
  const double E2 = m3 + OL + Ok * zp1 * zp1 + Or * zp1 * zp1 * zp1 * zp1;
  const double Omz = m3 / E2, OLz = OL / E2;
  return 2.5 * Omz / (pow(Omz, 4.0 / 7.0) - OLz + (1.0 + 0.5 * Omz) * (1.0 + OLz / 70.0)) / zp1;
}

// Eisenstein & Hu 1998 CDM+baryon transfer function T(k), k in 1/Mpc
// (the formula set of sgl.py:391-445).  The k-independent constants are
// computed once per instance into EH98Const.
struct EH98Const {
  double keq, ksilk, ac, bc, s, ab, bb, bnode, fb, fc;
  double lq0, lksilk;   // ln(13.41 k_eq), ln k_silk
};

__device__ inline EH98Const eh98_const(const Cosmo& c) {
  EH98Const k;
  double h = c.h, Om = c.Om, Ob = c.Ob, Oc = Om - Ob;
  double wm = Om * h * h, wb = Ob * h * h;
  k.keq = (100.0 * h / kCkms) * sqrt(2.0 * Om) * sqrt(1.0 + c.zeq);
  k.ksilk = 0.0016 * pow(wb, 0.52) * pow(wm, 0.73) * (1.0 + pow(10.4 * wm, -0.95));
  double a1 = pow(46.9 * wm, 0.67) * (1.0 + pow(32.1 * wm, -0.532));
  double a2 = pow(12.0 * wm, 0.424) * (1.0 + pow(45.0 * wm, -0.582));
  double b1 = 0.944 / (1.0 + pow(458.0 * wm, -0.708));
  double b2 = pow(0.395 * wm, -0.026);
  double fbm = Ob / Om;
  k.ac = pow(a1, -fbm) * pow(a2, -fbm * fbm * fbm);
  k.bc = 1.0 / (1.0 + b1 * (pow(Oc / Om, b2) - 1.0));
  k.s = 44.5 * log(9.83 / wm) / sqrt(1.0 + 10.0 * pow(wb, 0.75));
  double b3 = 0.313 * pow(wm, -0.419) * (1.0 + 0.607 * pow(wm, 0.674));
  double b4 = 0.238 * pow(wm, 0.223);
  double zd = 1291.0 * pow(wm, 0.251) / (1.0 + 0.659 * pow(wm, 0.828)) * (1.0 + b3 * pow(wb, b4));
  double T27 = kTcmb / 2.7;
  double Rd = 31.5 * wb / (T27 * T27 * T27 * T27) / (zd / 1000.0);
  double y = (1.0 + c.zeq) / (1.0 + zd);
  double sy = sqrt(1.0 + y);
  double g2 = y * (-6.0 * sy + (2.0 + 3.0 * y) * log((sy + 1.0) / (sy - 1.0)));
  k.ab = 2.07 * k.keq * k.s * pow(1.0 + Rd, -0.75) * g2;
  k.bb = 0.5 + fbm + (3.0 - 2.0 * fbm) * sqrt(1.0 + (17.2 * wm) * (17.2 * wm));
  k.bnode = 8.41 * pow(wm, 0.435);
  k.fb = fbm;
  k.fc = Oc / Om;
  k.lq0 = log(13.41 * k.keq);
  k.lksilk = log(k.ksilk);
  return k;
}

// T(k) of Eisenstein & Hu 1998 (sgl.py:391-445, _T_eh98) with lk = ln k given by
// the caller: the powers q^1.08 and (k / k_silk)^1.4 are exp of multiples
// of ln q and ln(k / k_silk), q^1.08 and ln(e + 1.8 b_c q) are shared by the
// three T_0 terms (relative difference from separate pow / log calls
// ~1e-15).
__device__ inline double eh98_T(const EH98Const& e, double k, double lk) {
  const double q = k / (13.41 * e.keq);
  const double pq = exp(1.08 * (lk - e.lq0));               // q^1.08
  const double c386 = 386.0 / (1.0 + 69.9 * pq);
  const double q2 = q * q;
  const double nbc = log(2.718281828459045 + 1.8 * e.bc * q);
  const double n1 = log(2.718281828459045 + 1.8 * q);
  const double ks = k * e.s;
  const double t54 = ks / 5.4;
  const double fk = 1.0 / (1.0 + t54 * t54 * t54 * t54);
  const double t0_1bc = nbc / (nbc + (14.2 + c386) * q2);          // T0(q, 1, b_c)
  const double t0_acbc = nbc / (nbc + (14.2 / e.ac + c386) * q2);  // T0(q, a_c, b_c)
  const double t0_11 = n1 / (n1 + (14.2 + c386) * q2);             // T0(q, 1, 1)
  const double TC = fk * t0_1bc + (1.0 - fk) * t0_acbc;
  const double bn = e.bnode / ks;
  const double s3 = e.s / cbrt(1.0 + bn * bn * bn);
  const double x = k * s3;
  const double jo = (x < 1e-8) ? 1.0 - x * x / 6.0 : sin(x) / x;
  const double t52 = ks / 5.2;
  const double bbk = e.bb / ks;
  const double TB = (t0_11 / (1.0 + t52 * t52) +
                     e.ab / (1.0 + bbk * bbk * bbk) * exp(-exp(1.4 * (lk - e.lksilk)))) * jo;
  return e.fb * TB + e.fc * TC;
}

// ---------------------------------------------------------------- NFW ----
constexpr double kNfwWSeries = 0.25;
constexpr int kNfwNSeries = 28;
constexpr double kNfwXH = 0.9;
constexpr double kTcASmall = 1e-3;
constexpr double kTcBand = 1e-5;

__device__ inline void nfw_fh(double x, double& F, double& h) {
  double w = (1.0 - x) * (1.0 + x);
  if (fabs(w) < kNfwWSeries) {
    double T = 0.0;
#pragma unroll
    for (int n = kNfwNSeries - 1; n >= 0; --n) T = T * w + 1.0 / (2 * n + 3);
    F = T;
    h = log(0.5 * x) + (1.0 + w * T);
    return;
  }
  if (x < 1.0) {
    double e = (1.0 - x) / x, s = sqrt(w);
    double A = log1p(e + sqrt(e * (e + 2.0)));
    F = (A / s - 1.0) / w;
    h = (x < kNfwXH) ? (log(2.0 / x) * x * x / (1.0 + s) + log1p(-x * x / (2.0 * (1.0 + s)))) / s
                     : log(0.5 * x) + A / s;
    return;
  }
  double v = -w, sv = sqrt(v), B = atan(sv);
  F = (1.0 - B / sv) / v;
  h = log(0.5 * x) + B / sv;
}

__device__ inline void tc_closed_raw(double x, double c, double& I1, double& I2) {
  double q = sqrt((c - x) * (c + x));
  double d = (x - 1.0) * (x - c) / (x * (1.0 + c));
  double L = log((c + q) / (2.0 * (1.0 + c))) - (q - c) / (1.0 + c);
  if (x < 1.0) {
    double u = (1.0 - x) * (1.0 + x), e = (1.0 - x) / x;
    double A0 = log1p(e + sqrt(e * (e + 2.0)));
    double At = log1p(d + sqrt(d * (d + 2.0)));
    double su = sqrt(u);
    I1 = (q / (1.0 + c) - 1.0) / u + (A0 - At) / (u * su);
    I2 = (L + (A0 - At) / su) / (x * x);
    return;
  }
  double v = (x - 1.0) * (x + 1.0);
  double B0 = 2.0 * asin(sqrt((x - 1.0) / (2.0 * x)));
  double md = (-d > 0.0) ? -d : 0.0;
  double Bt = 2.0 * asin(sqrt(md / 2.0));
  double sv = sqrt(v);
  I1 = (1.0 - q / (1.0 + c)) / v - (B0 - Bt) / (v * sv);
  I2 = (L + (B0 - Bt) / sv) / (x * x);
}

// 24-node Gauss-Legendre rule on [0, 1] (nodes of P_24 mapped by (z+1)/2).
static __constant__ double kGl24X[24] = {
    0.0024063900014893447, 0.012635722014345263, 0.0308627239986336, 0.05679223649779952,
    0.08999900701304853, 0.12993790421072282, 0.17595317403151223, 0.22728926430558022,
    0.28310324618697746, 0.3424786601519183, 0.40444056626319186, 0.4679715535686972,
    0.5320284464313028, 0.5955594337368082, 0.6575213398480817, 0.7168967538130225,
    0.7727107356944198, 0.8240468259684878, 0.8700620957892772, 0.9100009929869515,
    0.9432077635022005, 0.9691372760013663, 0.9873642779856547, 0.9975936099985107};
static __constant__ double kGl24W[24] = {
    0.0061706148999935454, 0.014265694314466872, 0.022138719408709776, 0.02964929245771837,
    0.036673240705540205, 0.043095080765976644, 0.04880932605205703, 0.0537221350579828,
    0.05775283402686281, 0.06083523646390171, 0.06291872817341415, 0.06396909767337611,
    0.06396909767337611, 0.06291872817341415, 0.06083523646390171, 0.05775283402686281,
    0.0537221350579828, 0.04880932605205703, 0.043095080765976644, 0.036673240705540205,
    0.02964929245771837, 0.022138719408709776, 0.014265694314466872, 0.0061706148999935454};

__device__ inline void tc_gl(double a, double xt, double& s1o, double& s2o) {
  double s1 = 0.0, s2 = 0.0;
  for (int j = 0; j < 24; ++j) {
    double v = kGl24X[j], s = 1.0 - v * v;
    double d = fmax(1.0 - a * a * s * s, 1e-300);
    s1 += kGl24W[j] * 2.0 * v * s / ((s + xt) * (s + xt) * sqrt(d));
    double e = fmax(1.0 - a * a * v * v, 0.0);
    s2 += kGl24W[j] * v / ((v + xt) * (v + xt) * (1.0 + sqrt(e)));
  }
  s1o = s1;
  s2o = s2;
}

__device__ inline void tc_point(double x, double xt, double& I1, double& I2) {
  double a = x / xt;
  if (a >= 1.0) { I1 = 0.0; I2 = 0.0; return; }
  if (a < kTcASmall) {
    double i0 = log((1.0 + xt) / xt) - 1.0 / (1.0 + xt), u1 = 1.0 + xt;
    double jj = 0.5 * (u1 * u1 - xt * xt) - 3.0 * xt * (u1 - xt) + 3.0 * xt * xt * log(u1 / xt) +
                xt * xt * xt * (1.0 / u1 - 1.0 / xt);
    I1 = i0 + 0.5 * a * a * jj;
    I2 = 0.5 * i0 + 0.125 * a * a * jj;
    return;
  }
  if (fabs(x - 1.0) < kTcBand) {
    if (xt <= 1.0 + 2.0 * kTcBand) { tc_gl(a, xt, I1, I2); return; }
    double a1, b1, a2, b2;
    tc_closed_raw(1.0 - kTcBand, xt, a1, b1);
    tc_closed_raw(1.0 + kTcBand, xt, a2, b2);
    double f = (x - (1.0 - kTcBand)) / (2.0 * kTcBand);
    I1 = a1 + f * (a2 - a1);
    I2 = b1 + f * (b2 - b1);
    return;
  }
  tc_closed_raw(x, xt, I1, I2);
}

// (kappa, gamma) of an NFW halo truncated at x_t (amplitude ks).
__device__ inline void trunc_kg(double x, double ks, double xt, double& kap, double& gam) {
  if (x < xt) {
    double F, h;
    nfw_fh(x, F, h);
    double ku = 2.0 * F, gu = 4.0 * h / (x * x) - ku;
    double I1, I2;
    tc_point(x, xt, I1, I2);
    double k = ku - 2.0 * I1, kb = ku + gu - 4.0 * I2;
    kap = ks * k;
    gam = ks * (kb - k);
  } else {
    double mut = log1p(xt) - xt / (1.0 + xt);
    kap = 0.0;
    gam = ks * 4.0 * mut / (x * x);
  }
}

// ------------------------------------------------------------ filaments --
// Orientation rule of the filament population (sgl_full.py:1828-1845): 12
// Gauss-Legendre nodes on [0, acos 0.1] plus one node carrying the clamped
// measure; weights are fractions of the quarter turn.  The amplitude factor
// kf = r_s / max(2 r_s, |cos phi| 20 r_s) turns the row's k0base into k0.
static __constant__ double kGl12X[12] = {-0.9815606342467192, -0.9041172563704748, -0.7699026741943047, -0.5873179542866175, -0.3678314989981802, -0.1252334085114689, 0.1252334085114689, 0.3678314989981802, 0.5873179542866175, 0.7699026741943047, 0.9041172563704748, 0.9815606342467192};
static __constant__ double kGl12W[12] = {0.04717533638651202, 0.10693932599531888, 0.1600783285433461, 0.20316742672306565, 0.23349253653835464, 0.2491470458134027, 0.2491470458134027, 0.23349253653835464, 0.20316742672306565, 0.1600783285433461, 0.10693932599531888, 0.04717533638651202};

__device__ inline void fil_orientation(int n, double& kf, double& wphi) {
  const double pc = acos(0.1), half_pi = 0.5 * kPi;
  double ph;
  if (n < 12) {
    ph = 0.5 * pc * (kGl12X[n] + 1.0);
    wphi = 0.5 * pc * kGl12W[n] / half_pi;
  } else {
    ph = 0.5 * (pc + half_pi);
    wphi = (half_pi - pc) / half_pi;
  }
  kf = 1.0 / fmax(2.0, fabs(cos(ph)) * 20.0);
}

// u bin l of a filament row (edges 0, geomspace(1e-5, 1, 120)): centre
// (geometric, u/2 for the central disc) and u1^2 - u0^2.
__device__ inline void fil_ubin(int l, double& uc, double& du2) {
  auto edge = [](int m) { return pow(10.0, -5.0 + 5.0 * m / 119.0); };
  const double u1 = edge(l), u0 = (l > 0) ? edge(l - 1) : 0.0;
  uc = (l > 0) ? sqrt(u0 * u1) : 0.5 * u1;
  du2 = u1 * u1 - u0 * u0;
}

// Per-row constants of a truncated NFW lens row (one (slice, mass) pair).
// Probabilists' Gauss-Hermite rule with 48 nodes (numpy.polynomial.hermite_e.
// hermegauss(48)), weights normalised to sum 1: the latent states of the
// positive clustering closure.
static __constant__ double kGhU[48] = {
-12.693012315439578, -11.753178461616617, -10.973300958513551, -10.275741581733518,
    -9.63088568694906, -9.023480280417651, -8.444371322523448, -7.887517316500365,
    -7.348660565908913, -6.824651320745979, -6.313069914905499, -5.811999987702531,
    -5.319884620419086, -4.8354308858789565, -4.357544108790291, -3.8852811172483914,
    -3.4178160486847764, -2.9544146888801195, -2.494414743040023, -2.037210303259033,
    -1.582239319674964, -1.1289732315096481, -0.6769081422063474, -0.22555707298016336,
    0.22555707298016336, 0.6769081422063474, 1.1289732315096481, 1.582239319674964,
    2.037210303259033, 2.494414743040023, 2.9544146888801195, 3.4178160486847764,
    3.8852811172483914, 4.357544108790291, 4.8354308858789565, 5.319884620419086,
    5.811999987702531, 6.313069914905499, 6.824651320745979, 7.348660565908913,
    7.887517316500365, 8.444371322523448, 9.023480280417651, 9.63088568694906,
    10.275741581733518, 10.973300958513551, 11.753178461616617, 12.693012315439578};
static __constant__ double kGhW[48] = {
    4.477155473875909e-36, 3.3764561431353977e-31, 2.0790589714186624e-27, 3.139476644799188e-24,
    1.7988549162372895e-21, 4.925465109698984e-19, 7.419993598064828e-17, 6.7566772746571655e-15,
    3.975805958471501e-13, 1.5883609812463425e-11, 4.474287153436362e-10, 9.154055746313818e-09,
    1.3927916895599565e-07, 1.6063936955426115e-06, 1.4266092324246835e-05, 9.881804917611342e-05,
    0.0005395865846271887, 0.0023430821117604002, 0.008149696855351924, 0.022838212100091373,
    0.05180518354956028, 0.09546340056142932, 0.14328245699317615, 0.17546354181186605,
    0.17546354181186605, 0.14328245699317615, 0.09546340056142932, 0.05180518354956028,
    0.022838212100091373, 0.008149696855351924, 0.0023430821117604002, 0.0005395865846271887,
    9.881804917611342e-05, 1.4266092324246835e-05, 1.6063936955426115e-06, 1.3927916895599565e-07,
    9.154055746313818e-09, 4.474287153436362e-10, 1.5883609812463425e-11, 3.975805958471501e-13,
    6.7566772746571655e-15, 7.419993598064828e-17, 4.925465109698984e-19, 1.7988549162372895e-21,
    3.139476644799188e-24, 2.0790589714186624e-27, 3.3764561431353977e-31, 4.477155473875909e-36};

struct RowConst {
  double ks, s, xt, ic1, mut4ks, warea, beta, i0, jj, valid, eps;
};

__device__ inline RowConst row_const(const double* R) {
  // R: k_s, r_s, x_t, r_vir, w_row, beta
  RowConst c;
  c.ks = R[0];
  c.xt = R[2];
  c.s = R[3] / R[1];
  c.ic1 = 1.0 / (1.0 + c.xt);
  c.mut4ks = 4.0 * c.ks * (log1p(c.xt) - c.xt * c.ic1);
  c.warea = R[4] * kPi * R[3] * R[3];
  c.beta = R[5];
  const double xt = c.xt, u1 = 1.0 + xt;
  c.i0 = log(u1 / xt) - 1.0 / u1;
  c.jj = 0.5 * (u1 * u1 - xt * xt) - 3.0 * xt * (u1 - xt) + 3.0 * xt * xt * log(u1 / xt) +
         xt * xt * xt * (1.0 / u1 - 1.0 / xt);
  c.valid = (R[4] > 0.0) ? 1.0 : 0.0;
  c.eps = R[6];
  return c;
}


// (kappa, gamma) of the truncated NFW row at x, the same closed forms as
// trunc_kg with the shared subexpressions (arccosh(1/x), log x, 1/x, per-row
// constants) evaluated once; the rare |x - 1| < 1e-5 band takes trunc_kg.
__device__ inline void trunc_kg_row(double x, const RowConst& rc, double& kap, double& gam) {
  const double xt = rc.xt;
  if (x >= xt) {
    kap = 0.0;
    gam = rc.mut4ks / (x * x);
    return;
  }
  if (fabs(x - 1.0) < kTcBand) {
    trunc_kg(x, rc.ks, xt, kap, gam);
    return;
  }
  constexpr double kLn2 = 0.6931471805599453;
  const double a = x / xt;
  const double rx = 1.0 / x, rx2 = rx * rx;
  const double lx = log(x);
  double F, h, I1, I2;
  const double w = (1.0 - x) * (1.0 + x);
  const bool series = fabs(w) < kNfwWSeries;
  double T = 0.0;
  if (series) {
#pragma unroll
    for (int n = kNfwNSeries - 1; n >= 0; --n) T = T * w + 1.0 / (2 * n + 3);
    F = T;
    h = lx - kLn2 + (1.0 + w * T);
  }
  const bool small = a < kTcASmall;
  if (small) {
    I1 = rc.i0 + 0.5 * a * a * rc.jj;
    I2 = 0.5 * rc.i0 + 0.125 * a * a * rc.jj;
  }
  if (x < 1.0) {
    const double e = (1.0 - x) * rx, su = sqrt(w), isu = 1.0 / su, iu = 1.0 / w;
    const double A0 = log1p(e + sqrt(e * (e + 2.0)));
    if (!series) {
      F = (A0 * isu - 1.0) * iu;
      h = (x < kNfwXH) ? ((kLn2 - lx) * x * x / (1.0 + su) + log1p(-x * x / (2.0 * (1.0 + su)))) * isu
                       : lx - kLn2 + A0 * isu;
    }
    if (!small) {
      const double c = xt, q = sqrt((c - x) * (c + x));
      const double d = (x - 1.0) * (x - c) * rx * rc.ic1;
      const double L = log((c + q) * rc.ic1 * 0.5) - (q - c) * rc.ic1;
      const double At = log1p(d + sqrt(d * (d + 2.0)));
      I1 = (q * rc.ic1 - 1.0) * iu + (A0 - At) * iu * isu;
      I2 = (L + (A0 - At) * isu) * rx2;
    }
  } else {
    const double v = -w, sv = sqrt(v), isv = 1.0 / sv, iv = 1.0 / v;
    const double B0 = 2.0 * asin(sqrt((x - 1.0) * 0.5 * rx));
    if (!series) {
      F = (1.0 - B0 * isv) * iv;
      h = lx - kLn2 + B0 * isv;
    }
    if (!small) {
      const double c = xt, q = sqrt((c - x) * (c + x));
      const double d = (x - 1.0) * (x - c) * rx * rc.ic1;
      const double L = log((c + q) * rc.ic1 * 0.5) - (q - c) * rc.ic1;
      const double Bt = 2.0 * asin(sqrt(fmax(-d, 0.0) * 0.5));
      I1 = (1.0 - q * rc.ic1) * iv - (B0 - Bt) * iv * isv;
      I2 = (L + (B0 - Bt) * isv) * rx2;
    }
  }
  const double ku = 2.0 * F - 2.0 * I1;
  const double kb = 4.0 * h * rx2 - 4.0 * I2;
  kap = rc.ks * ku;
  gam = rc.ks * (kb - ku);
}

// Halo ellipticity: orientation nodes per cell (sgl_full.py:1535, 1659).
constexpr int kEllPhi = 12;

// cos(2 phi_q) of orientation node q, phi_q = (q + 1/2) (pi/2) / kEllPhi
// (sgl_full.py:1659).
static __constant__ double kEllCos2[kEllPhi] = {
    0.9914448613738104, 0.9238795325112867, 0.7933533402912352, 0.6087614290087207,
    0.38268343236508984, 0.1305261922200517, -0.1305261922200516, -0.3826834323650895,
    -0.6087614290087207, -0.793353340291235, -0.9238795325112867, -0.9914448613738104};
__device__ __forceinline__ double ell_cos2(int q) { return kEllCos2[q]; }

// Pseudo-elliptical truncated NFW of a row (lensing.cpp:56 kappagammaNFWeps as
// sgl_full.py:582-613 transcribes it): the circular profile at the elliptical
// radius x_e = x sqrt(1 - eps cos 2phi), c2 = (cos 2phi - eps) / (1 - eps cos 2phi)
// (= (x1^2 - x2^2) / x_e^2 of the reference's x1, x2),
//   kappa = k0 + eps c2 g0,  gamma^2 = g0^2 + 2 eps c2 g0 k0 + eps^2 (k0^2 - (c2 g0)^2),
// g0 = k_s below x_e = 1e-4; beyond the truncation kappa = 0 and gamma is the
// monopole shear at the circular x.
__device__ inline void kg_eps_row(double x, double cos2, const RowConst& rc, double& kap,
                                  double& gam) {
  const double eps = rc.eps, f2 = 1.0 - eps * cos2;
  const double xe = fmax(x * sqrt(f2), 1e-12);
  if (xe >= rc.xt) {
    kap = 0.0;
    gam = rc.mut4ks / (x * x);
    return;
  }
  const double c2 = (cos2 - eps) / f2;
  double k0, g0;
  trunc_kg_row(xe, rc, k0, g0);
  if (xe < 1e-4) g0 = rc.ks;
  kap = k0 + eps * c2 * g0;
  gam = sqrt(fmax(g0 * g0 + 2.0 * eps * c2 * g0 * k0 + eps * eps * (k0 * k0 - (c2 * g0) * (c2 * g0)),
                  0.0));
}

// Per-row record of the batched pass (80 B, built once per row by k_rowf):
// FP64 weight and bias, FP32 profile constants with log s and 1/s (so that
// log x and 1/x of a cell are one add and one multiply), and the FP32
// coefficients s2[n] of the small-x series of the truncation term I2 (see
// trunc_kg_row_f32).  warea = 0 marks an empty row.
constexpr int kI2Terms = 8;
struct RowF {
  double warea, beta;
  float ks, xt, ixt, ic1, mut4ks, s, log_s, inv_s;
  float s2[kI2Terms];
};
static_assert(sizeof(RowF) == 80, "RowF is one 80-byte record");

// Hardware FP32 reciprocal, reciprocal square root, square root (1 ulp) and
// natural log (v_log_f32 is log2; positive normal arguments only).
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float log_hw(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
__device__ __forceinline__ float rsq_hw(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }
// log1p(y), y > -1, from the hardware log: with u = fl(1 + y),
// log(u) y / (u - 1) cancels the rounding error of 1 + y (Goldberg 1991,
// "What every computer scientist should know about floating-point
// arithmetic", Theorem 4), leaving a few ulp.
__device__ __forceinline__ float log1p_hw(float y) {
  const float u = 1.0f + y, d = u - 1.0f;
  return (d == 0.0f) ? y : log_hw(u) * (y * rcp_hw(d));
}

// FP32 truncated-NFW profile of the batched path, in a form without the
// cancellation of the closed forms at x = 1, so no cell needs FP64.  With
// c = x_t, u = 1 - x^2 and
//   P = sqrt(2 / (x (1 + x))),  Q = sqrt((c + x) / (1 + c)),
//   R = sqrt((c - x) / (1 + c)),  z = (x - 1)(c - x) / (2 x (1 + c)),
//   G(z) = 2F1(1/2, 1/2; 3/2; z) = asinh(sqrt(-z)) / sqrt(-z) (z < 0),
//                                  asin(sqrt(z)) / sqrt(z)    (z > 0),
// the arccosh / arccos terms of the closed forms (trunc_kg_row) are
// sqrt(|u|) P G(z_e) and sqrt(|u|) P R G(z) on both sides of x = 1 (the
// truncation term's arccosh argument (x^2 + c) / (x (1 + c)) is 1 - 2z; the
// projected NFW truncated at c in this form: Takada & Jain 2003, MNRAS 340,
// 580), and the untruncated terms cancel exactly:
//   kappa / ks = -2 R (Q - P G(z)) / u,
//   kbar / ks  = 4 (ln(x/2) - L + P R G(z)) / x^2,   gamma = kbar - kappa,
// L = ln((c + q) / (2 (1 + c))) - (q - c) / (1 + c), q = (1 + c) R Q.
// Near x = 1 (|x - 1| < kProfT1) Q - P G vanishes like t = x - 1: it is
// summed from (Q - 1)/t, (P - 1)/t and (G - 1)/t, each formed without a
// difference, and u = -t (1 + x) cancels t exactly.  G - 1 for |z| <
// kProfZs is the hypergeometric series.  For x < kProfXs and x / c <
// kProfAs the bracket of kbar cancels (it is O(x^2)); there
//   kbar / ks = 4 h(x) / x^2 - 4 I2,  I2 = sum_n s2[n] (x/c)^(2n)
// with h the untruncated mass in the small-x form and s2 from k_rowf.
// Relative error of kappa and gamma < 1e-5 over x < c (hardware rcp / rsq /
// sqrt / log), against the FP64 closed forms.
constexpr float kProfT1 = 0.1f, kProfZs = 0.1f, kProfXs = 0.5f, kProfAs = 0.35f;

// (G(z) - 1) / z = sum_{n >= 1} g_n z^(n-1), g_{n+1} = g_n (n + 1/2)^2 / ((n + 1)(n + 3/2))
__device__ __forceinline__ float hyp_gm1z(float z) {
  float s = 12155.0f / 1245184.0f;
  s = s * z + 6435.0f / 557056.0f;
  s = s * z + 143.0f / 10240.0f;
  s = s * z + 231.0f / 13312.0f;
  s = s * z + 63.0f / 2816.0f;
  s = s * z + 35.0f / 1152.0f;
  s = s * z + 5.0f / 112.0f;
  s = s * z + 3.0f / 40.0f;
  s = s * z + 1.0f / 6.0f;
  return s;
}

__device__ inline void trunc_kg_row_f32(float x, float lx, float rx, const RowF& rc, float& kap,
                                        float& gam) {
  constexpr float kLn2f = 0.69314718f;
  const float c = rc.xt, ic1 = rc.ic1;
  const float rx2 = rx * rx;
  if (x >= c) {
    kap = 0.0f;
    gam = rc.mut4ks * rx2;
    return;
  }
  const float t = x - 1.0f, opx = 1.0f + x;
  const float P = rsq_hw(0.5f * x * opx);
  const float R2 = (c - x) * ic1;
  const float Q = sqrt_hw((c + x) * ic1), R = sqrt_hw(R2);
  const float z = 0.5f * t * R2 * rx;
  float G, gm1z = 0.0f;
  if (fabsf(z) < kProfZs) {
    gm1z = hyp_gm1z(z);
    G = 1.0f + z * gm1z;
  } else if (z < 0.0f) {
    const float is = rsq_hw(-z), s = -z * is;
    G = log_hw(s + sqrt_hw(1.0f - z)) * is;
  } else {
    const float is = rsq_hw(z), s = z * is;
    G = asinf(s) * is;
  }
  float ku;
  if (fabsf(t) < kProfT1) {
    const float ro = rcp_hw(opx);
    const float qm1t = ic1 * rcp_hw(Q + 1.0f);
    const float pm1t = -(2.0f + x) * rx * ro * rcp_hw(P + 1.0f);
    const float gm1t = gm1z * 0.5f * R2 * rx;
    ku = 2.0f * R * (qm1t - pm1t - P * gm1t) * ro;
  } else {
    ku = -2.0f * R * (Q - P * G) * rcp_hw(-t * opx);
  }
  float kb;
  const float a = x * rc.ixt;
  if (x < kProfXs && a < kProfAs) {
    const float w = -t * opx, isu = rsq_hw(w), su = w * isu;
    const float i1s = rcp_hw(1.0f + su);
    const float h = ((kLn2f - lx) * x * x * i1s + log1p_hw(-0.5f * x * x * i1s)) * isu;
    const float y = a * a;
    float I2 = rc.s2[kI2Terms - 1];
#pragma unroll
    for (int n = kI2Terms - 2; n >= 0; --n) I2 = I2 * y + rc.s2[n];
    kb = 4.0f * h * rx2 - 4.0f * I2;
  } else {
    const float q = R * Q * (1.0f + c);
    const float L = log_hw((c + q) * ic1 * 0.5f) - (q - c) * ic1;
    kb = 4.0f * (lx - kLn2f - L + P * R * G) * rx2;
  }
  kap = rc.ks * ku;
  gam = rc.ks * (kb - ku);
}

// Row record of a truncated NFW row (FP64 constants in, FP32 record out):
// s2[n] = b_n M_{2n+1}(c), b_n the coefficients of (1 - sqrt(1 - y)) / y =
// sum b_n y^n and M_m(c) = int_0^1 v^m / (v + c)^2 dv, so that the
// truncation term I2 = int_0^1 v / (v + c)^2 (1 - sqrt(1 - a^2 v^2)) /
// (a^2 v^2) dv is sum_n s2[n] a^(2n).  The moments take the 12-point
// Gauss-Legendre rule (kGl12X/W mapped to [0, 1]): against 40-digit quadrature its relative error for
// m <= 15 is below 2.3e-8 at c = 0.2, 1e-12 at c = 0.5 and 3e-15 for c >= 1,
// under the FP32 rounding of s2 throughout.
__device__ inline RowF make_rowf(const RowConst& c) {
  RowF f;
  f.warea = (c.valid != 0.0) ? c.warea : 0.0;
  f.beta = c.beta;
  f.ks = (float)c.ks;
  f.xt = (float)c.xt;
  f.ixt = (float)(1.0 / c.xt);
  f.ic1 = (float)c.ic1;
  f.mut4ks = (float)c.mut4ks;
  f.s = (float)c.s;
  f.log_s = (float)log(c.s);
  f.inv_s = (float)(1.0 / c.s);
  double m[kI2Terms];
  for (int n = 0; n < kI2Terms; ++n) m[n] = 0.0;
  for (int j = 0; j < 12; ++j) {
    const double v = 0.5 * (kGl12X[j] + 1.0), v2 = v * v;
    const double g = 0.5 * kGl12W[j] * v / ((v + c.xt) * (v + c.xt));
    double p = g;
    for (int n = 0; n < kI2Terms; ++n) { m[n] += p; p *= v2; }
  }
  double bn = 0.5;
  for (int n = 0; n < kI2Terms; ++n) {
    f.s2[n] = (float)(bn * m[n]);
    bn *= (2.0 * n + 1.0) / (2.0 * n + 4.0);
  }
  return f;
}

// Jump of one cell: xi = -ln detA and the tilted weight w e^{-xi} = w detA;
// dropped (weight 0) when detA <= 0 or xi > 30.
constexpr double kXiMax = 30.0;
// Batched-path variant: detA - 1 = kappa (kappa - 2) - gamma^2 and detA
// formed in FP64 (no cancellation), the log on the FP32 hardware path: a
// short log1p series for |detA - 1| < 1e-3, else log of detA rounded to FP32
// (relative precision kept down to detA ~ 1e-13, i.e. over the whole
// xi <= 30 range).  Error ~1e-7 in xi; the bin index and moments need no
// more.
__device__ __forceinline__ void cell_jump_f32log(double kap, double gam, double w, double& xi,
                                                 double& ws) {
  double dm1 = kap * (kap - 2.0) - gam * gam;
  double detA = 1.0 + dm1;
  if (!(detA > 0.0)) { xi = 0.0; ws = 0.0; return; }
  float fd = (float)dm1;
  float l1p = (fabsf(fd) < 1e-3f) ? fd * (1.0f - fd * (0.5f - fd * (1.0f / 3.0f)))
                                  : log_hw(fmaxf((float)detA, 1e-30f));
  double x = -(double)l1p;
  if (x > kXiMax) { xi = 0.0; ws = 0.0; return; }
  xi = x;
  ws = w * detA;
}
__device__ __forceinline__ void cell_jump(double kap, double gam, double w, double& xi, double& ws) {
  double detA = (1.0 - kap) * (1.0 - kap) - gam * gam;
  if (!(detA > 0.0)) { xi = 0.0; ws = 0.0; return; }
  double x = -log(detA);
  if (x > kXiMax) { xi = 0.0; ws = 0.0; return; }
  xi = x;
  ws = w * detA;
}

}  // namespace sgl
