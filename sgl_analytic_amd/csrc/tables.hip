// Lens-population tables on the device, batched over cosmologies.
//
// For every instance (cosmology + source redshift) the kernels below build
//   derived scalars     Omega_r, Omega_DE, H0, rho_c0, rho_m, chi(z_s), D_A(z_s), A_norm
//   slice table         z, chi, dchi, dchi (1+z)^2, D(z), Sigma_cr(z, z_s), E(z)^2, Dg(z)
//   mass table          M, sigma(M, 0), dln sigma / dln M
//   row table           per (slice, mass): k_s, r_s, x_t = c, r_vir, w_row, beta
//   slice covariance    C_ij of the smoothed 1-D linear density field
// following sgl.Cosmology (sgl.py:65-615) on the make_cosmology configuration
// (pipeline.py:36-72), the builder's per-(z, M) quantities
// (sgl_full.py:1848-1885, 541-543; subhalos.py:381-385, 472-481) and
// shell_covariance (sgl_full.py:2514-2538).
#include <cmath>

#include "sgl_amd_internal.h"
#include "ref_sub.h"

namespace sgl {

// ------------------------------------------------------------ background --
// One workgroup of kBgThreads per instance.  chi(z) = (c/H0) int dz / E on
// kPanels uniform z-panels with 8-point Gauss-Legendre, a fixed-order prefix
// sum (each thread runs over its own kPanels / kBgThreads consecutive panels,
// then a Hillis-Steele scan of the per-thread totals), then Newton on each
// slice midpoint chi_i = (i + 1/2) chi_s / n_z.
constexpr int kPanels = 2048;
constexpr int kBgThreads = 256;
constexpr int kBgPer = kPanels / kBgThreads;
static_assert(kPanels % kBgThreads == 0, "whole panel runs per thread");
__constant__ double kGl8X[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290,
                                -0.1834346424956498, 0.1834346424956498, 0.5255324099163290,
                                0.7966664774136267, 0.9602898564975363};
__constant__ double kGl8W[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873,
                                0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                                0.2223810344533745, 0.1012285362903763};

__device__ inline double inv_e_integral(const Cosmo& c, double Or, double OL, double a, double b) {
  double m = 0.5 * (a + b), r = 0.5 * (b - a), s = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) s += kGl8W[j] / sqrt(e2_of_z(c, Or, OL, m + r * kGl8X[j]));
  return s * r;
}

// chi at the kPanels + 1 panel edges j dz (j = 0..kPanels) into chi_e, by
// one workgroup of kBgThreads (run: kBgThreads doubles of LDS scratch).  Ends
// with every edge written but without a barrier after the last writes.
__device__ inline void chi_panels(const Cosmo& c, double Or, double OL, double dh, double dz,
                                  double* chi_e, double* run) {
  const int t = threadIdx.x;
  for (int p = t; p < kPanels; p += kBgThreads)
    chi_e[p + 1] = dh * inv_e_integral(c, Or, OL, p * dz, (p + 1) * dz);
  __syncthreads();
  double loc[kBgPer], acc = 0.0;
#pragma unroll
  for (int j = 0; j < kBgPer; ++j) {
    acc += chi_e[t * kBgPer + j + 1];
    loc[j] = acc;
  }
  run[t] = acc;
  __syncthreads();
  for (int off = 1; off < kBgThreads; off <<= 1) {
    const double add = (t >= off) ? run[t - off] : 0.0;
    __syncthreads();
    run[t] += add;
    __syncthreads();
  }
  const double base = (t > 0) ? run[t - 1] : 0.0;
#pragma unroll
  for (int j = 0; j < kBgPer; ++j) chi_e[t * kBgPer + j + 1] = base + loc[j];
  if (t == 0) chi_e[0] = 0.0;
}

__global__ __launch_bounds__(kBgThreads) void k_background(const Cosmo* __restrict__ cosmo,
                                                           int n_z_max, Derived* __restrict__ der,
                                                           double* __restrict__ slices) {
  __shared__ double chi_e[kPanels + 1];
  __shared__ double run[kBgThreads];
  const int b = blockIdx.x, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  // radiation "engine": Omega_r = Omega_m / (1 + z_eq); None: no radiation
  // (sgl.py:93-104)
  const double Or = (c.rad_none != 0.0) ? 0.0 : c.Om / (1.0 + c.zeq);
  const double OL = 1.0 - c.Om - c.Ok - Or;
  const double H0 = 100.0 * c.h;
  const double dh = kCkms / H0;
  const double zs = c.zs;
  const int n_z = (int)c.n_z;
  const double dz = zs / kPanels;
  chi_panels(c, Or, OL, dh, dz, chi_e, run);
  __syncthreads();
  const double chis = chi_e[kPanels];
  const double dchi = chis / n_z;
  const double DAs = f_k(c.Ok, H0, chis) / (1.0 + zs);
  if (t == 0) {
    Derived d = {};
    d.Or = Or; d.OL = OL; d.H0 = H0; d.rhoc0 = kRhoc0Unit * c.h * c.h;
    d.rhom = c.Om * d.rhoc0; d.chis = chis; d.DAs = DAs; d.dchi = dchi; d.n_z = n_z;
    der[b] = d;
  }
  for (int i = t; i < n_z_max; i += kBgThreads) {
    double* S = slices + ((size_t)b * n_z_max + i) * kSliceStride;
    if (i >= n_z) {
      for (int q = 0; q < kSliceStride; ++q) S[q] = 0.0;
      continue;
    }
    double target = (i + 0.5) * dchi;
    int lo = 0, hi = kPanels;  // chi_e[lo] <= target < chi_e[hi]
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (chi_e[mid] <= target) lo = mid; else hi = mid;
    }
    double z0 = lo * dz;
    double z = z0 + (target - chi_e[lo]) / (chi_e[lo + 1] - chi_e[lo]) * dz;
    for (int it = 0; it < 6; ++it) {
      double F = chi_e[lo] + dh * inv_e_integral(c, Or, OL, z0, z) - target;
      z -= F * sqrt(e2_of_z(c, Or, OL, z)) / dh;
    }
    double E2 = e2_of_z(c, Or, OL, z);
    double DAl = f_k(c.Ok, H0, target) / (1.0 + z);
    double DAls = f_k(c.Ok, H0, chis - target) / (1.0 + zs);
    S[kSlZ] = z;
    S[kSlChi] = target;
    S[kSlDchi] = dchi;
    S[kSlWz] = dchi * (1.0 + z) * (1.0 + z);
    S[kSlD] = 0.0;
    S[kSlScr] = kCkms * kCkms * DAs / (4.0 * kPi * kGnewt * DAl * DAls);
    S[kSlE2] = E2;
    S[kSlDg] = 0.0;
  }
}

// Reference grid (SGL_FLAG_REF_GRID): shells whose edges are the engine's
// log-z nodes z_j = exp(ln 0.01 + j (ln 12.3412 - ln 0.01) / 102) below z_s,
// plus 0 and z_s, one midpoint node each (sgl_full.py:471-518, zint="midpoint").
// The caller stores the shell count (nodes below z_s + 1) in the record's n_z
// slot.  Slice record: z = z_lo + (z_hi - z_lo)/2, chi(z), the comoving width
// dchi = c/H(z) dz (the Limber node width of sgl_full.py:1684), dchi (1+z)^2
// (the lens weight of sgl_full.py:1860 without its pi), Sigma_cr, E^2; D and
// Dg are filled by the growth kernel.
constexpr double kZminEng = 0.01, kZmaxEng = 12.341169644129371;
constexpr int kNzEng = 103;
// Node q (ascending) of the n-point Gauss-Legendre rule on [-1, 1] and its
// weight: Newton on P_n from cos(pi (q + 3/4) / (n + 1/2)) to convergence,
// w = 2 / ((1 - x^2) P_n'(x)^2) (numpy.polynomial.legendre.leggauss's rule).
__device__ inline void gl_node(int n, int q, double& x, double& w) {
  double z = -cos(kPi * (q + 0.75) / (n + 0.5));
  double dp = 1.0;
  for (int it = 0; it < 64; ++it) {
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    const double dz = p1 / dp;
    z -= dz;
    if (fabs(dz) <= 1e-16 * fmax(1.0, fabs(z))) break;
  }
  double p0 = 1.0, p1 = z;
  for (int k = 2; k <= n; ++k) {
    const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  dp = (n == 1) ? 1.0 : n * (z * p1 - p0) / (z * z - 1.0);
  x = z;
  w = 2.0 / ((1.0 - z * z) * dp * dp);
}

__device__ inline double engine_z(int j) {
  const double l0 = log(kZminEng), l1 = log(kZmaxEng);
  return exp((j == kNzEng - 1) ? l1 : l0 + j * ((l1 - l0) / (kNzEng - 1)));
}

__global__ __launch_bounds__(kBgThreads) void k_ref_background(const Cosmo* __restrict__ cosmo,
                                                               int n_z_max,
                                                               Derived* __restrict__ der,
                                                               double* __restrict__ slices) {
  __shared__ double chi_e[kPanels + 1];
  __shared__ double run[kBgThreads];
  const int b = blockIdx.x, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  const double Or = (c.rad_none != 0.0) ? 0.0 : c.Om / (1.0 + c.zeq);
  const double OL = 1.0 - c.Om - c.Ok - Or;
  const double H0 = 100.0 * c.h;
  const double dh = kCkms / H0;
  const double zs = c.zs;
  const int n_z = (int)c.n_z;
  const double dz = zs / kPanels;
  chi_panels(c, Or, OL, dh, dz, chi_e, run);
  __syncthreads();
  const double chis = chi_e[kPanels];
  const double DAs = f_k(c.Ok, H0, chis) / (1.0 + zs);
  if (t == 0) {
    Derived d = {};
    d.Or = Or; d.OL = OL; d.H0 = H0; d.rhoc0 = kRhoc0Unit * c.h * c.h;
    d.rhom = c.Om * d.rhoc0; d.chis = chis; d.DAs = DAs; d.dchi = 0.0; d.n_z = n_z;
    der[b] = d;
  }
  for (int i = t; i < n_z_max; i += kBgThreads) {
    double* S = slices + ((size_t)b * n_z_max + i) * kSliceStride;
    if (i >= n_z) {
      for (int q = 0; q < kSliceStride; ++q) S[q] = 0.0;
      continue;
    }
    // midpoint shells: edges 0, the engine nodes below z_s, z_s, one node
    // in the middle; gauss: the same shells at ng Gauss-Legendre nodes each
    // (slice i = node i % ng of shell i / ng); engine shells: [z_i, z_{i+1}]
    // of the engine grid, node at the upper edge
    const bool engine = c.zint == 1.0;
    const int ng = (c.zint == 2.0 && c.zint_n >= 1.0 && c.zint_n <= kRefMaxShells) ? (int)c.zint_n : 1;
    const int sh = i / ng, n_sh = n_z / ng;
    double t = 0.0, w = 2.0;
    if (ng > 1) gl_node(ng, i - sh * ng, t, w);
    const double zlo = engine ? engine_z(i) : (sh == 0) ? 0.0 : engine_z(sh - 1);
    const double zhi = engine ? engine_z(i + 1) : (sh == n_sh - 1) ? zs : engine_z(sh);
    const double z = engine ? zhi : zlo + 0.5 * (zhi - zlo) * (t + 1.0);
    const double dzs = engine ? zhi - zlo : 0.5 * (zhi - zlo) * w;
    const int lo = min((int)(z / dz), kPanels - 1);
    const double chi = chi_e[lo] + dh * inv_e_integral(c, Or, OL, lo * dz, z);
    const double E2 = e2_of_z(c, Or, OL, z);
    const double DAl = f_k(c.Ok, H0, chi) / (1.0 + z);
    const double DAls = f_k(c.Ok, H0, chis - chi) / (1.0 + zs);
    const double dchi = kCkms / (H0 * sqrt(E2)) * dzs;
    S[kSlZ] = z;
    S[kSlChi] = chi;
    S[kSlDchi] = dchi;
    S[kSlWz] = dchi * (1.0 + z) * (1.0 + z);
    S[kSlD] = 0.0;
    S[kSlScr] = kCkms * kCkms * DAs / (4.0 * kPi * kGnewt * DAl * DAls);
    S[kSlE2] = E2;
    S[kSlDg] = 0.0;
  }
}

// ---------------------------------------------------------------- growth --
// D'' + (2 + dlnH/dlna) D' - 3/2 Omega_m(a) D = 0 in x = ln a (sgl.py:310-347),
// classical RK4 on kGrowthSteps uniform steps from a = 1e-6, cubic Hermite
// dense output at every slice node, normalised by D(a = 1).
//
// The equation is linear, y' = A(x) y with y = (D, D') and
// A = [[0, 1], [beta(x), alpha(x)]], so one RK4 step is a 2x2 matrix T_n and
// the solve is the ordered product of 4096 of them.  One wave per instance:
// lane L forms the product of its kGrowthLane consecutive steps, a
// Hillis-Steele scan over the lanes (shuffles) gives every lane the product of
// the steps before its own, and each lane then steps its own chunk once more
// from its starting state, writing the slice nodes that fall in it.  The
// serial depth is 2 x 64 steps and 6 scan levels instead of 4096 steps.
constexpr int kGrowthLane = kGrowthSteps / 64;
static_assert(kGrowthSteps % 64 == 0, "whole chunks per lane");

struct GrowthRhs {
  double Om, Ok, Or, OL, w, wa;
  bool lambda_de;
  // alpha = -(2 + dlnH/dx), beta = 3/2 Omega_m(x)
  __device__ inline void coef(double x, double& alpha, double& beta) const {
    double e1 = exp(-x);
    double e2 = e1 * e1;
    double matter = Om * e2 * e1;
    double curv = Ok * e2;
    double rad = Or * e2 * e2;
    double de, wz;
    if (lambda_de) {
      de = OL;
      wz = -1.0;
    } else {
      double a = 1.0 / e1;
      de = OL * exp(-3.0 * (1.0 + w + wa) * x + 3.0 * wa * (a - 1.0));
      wz = w + wa * (1.0 - a);
    }
    double itot = 1.0 / (matter + rad + curv + de);
    double dh = -0.5 * (3.0 * matter + 4.0 * rad + 2.0 * curv + 3.0 * (1.0 + wz) * de) * itot;
    alpha = -(2.0 + dh);
    beta = 1.5 * matter * itot;
  }
};

// One RK4 step of y' = A y on (y0, y1), A at the step's start (a1, b1),
// midpoint (a2, b2) and end (a4, b4).
__device__ inline void rk4_lin(double h, double a1, double b1, double a2, double b2, double a4,
                               double b4, double& y0, double& y1) {
  const double k1a = y1, k1b = b1 * y0 + a1 * y1;
  const double u0 = y0 + 0.5 * h * k1a, u1 = y1 + 0.5 * h * k1b;
  const double k2a = u1, k2b = b2 * u0 + a2 * u1;
  const double v0 = y0 + 0.5 * h * k2a, v1 = y1 + 0.5 * h * k2b;
  const double k3a = v1, k3b = b2 * v0 + a2 * v1;
  const double w0 = y0 + h * k3a, w1 = y1 + h * k3b;
  const double k4a = w1, k4b = b4 * w0 + a4 * w1;
  y0 += h / 6.0 * (k1a + 2.0 * k2a + 2.0 * k3a + k4a);
  y1 += h / 6.0 * (k1b + 2.0 * k2b + 2.0 * k3b + k4b);
}

// Legacy growth (growth_mode="legacy", LambdaCDM only).  D(a) is the Heath
// integral E(a) I(a) / (E(1) I(1)), I(a) = int_0^a da' / (a' E(a'))^3, with E
// of matter + curvature + Lambda only, Omega_L = 1 - Omega_m - Omega_k
// (sgl.py:270-306: the integral never carries radiation).  With a' = a t^2,
// I(a) = 2 a^(5/2) int_0^1 t^4 (Om + Ok a t^2 + OL a^3 t^6)^(-3/2) dt, whose
// integrand is analytic: the 24-point Gauss-Legendre rule gives it to 1.3e-14
// relative over the parameter box (CPU check against 30-digit quadrature).
__device__ inline double heath_integral(double a, double Om, double Ok, double OLg) {
  double s = 0.0;
  for (int j = 0; j < 24; ++j) {
    const double t = kGl24X[j], t2 = t * t, x = a * t2;
    const double q = Om + Ok * x + OLg * x * x * x;
    s += kGl24W[j] * t2 * t2 / (q * sqrt(q));
  }
  return 2.0 * a * a * sqrt(a) * s;
}

__global__ __launch_bounds__(64) void k_growth(const Cosmo* __restrict__ cosmo, int n_z_max,
                                               Derived* __restrict__ der,
                                               double* __restrict__ slices,
                                               double* __restrict__ gdense) {
  __shared__ double xs[kGrowthMaxNodes];
  const int b = blockIdx.x, L = threadIdx.x;
  const Cosmo c = cosmo[b];
  const Derived d = der[b];
  const GrowthRhs f{c.Om, c.Ok, d.Or, d.OL, c.w0, c.wa, c.w0 == -1.0 && c.wa == 0.0};
  const int n_z = (int)c.n_z;
  double* Sb = slices + (size_t)b * n_z_max * kSliceStride;
  if (c.growth_legacy != 0.0) {
    const double OLg = 1.0 - c.Om - c.Ok;
    const double Eg1 = sqrt(c.Om + c.Ok + OLg);
    const double D0 = Eg1 * heath_integral(1.0, c.Om, c.Ok, OLg);
    const double g0 = cpt_growth(0.0, c.Om, c.Ok, d.Or, d.OL);
    for (int i = L; i < n_z; i += 64) {
      const double z = Sb[i * kSliceStride + kSlZ], a = 1.0 / (1.0 + z);
      const double Eg = sqrt(c.Om / (a * a * a) + c.Ok / (a * a) + OLg);
      Sb[i * kSliceStride + kSlD] = Eg * heath_integral(a, c.Om, c.Ok, OLg) / D0;
      Sb[i * kSliceStride + kSlDg] = cpt_growth(z, c.Om, c.Ok, d.Or, d.OL) / g0;
    }
    if (L == 0) der[b].D0 = D0;
    return;
  }
  for (int i = L; i < n_z; i += 64) xs[i] = -log1p(Sb[i * kSliceStride + kSlZ]);
  const double x0 = log(1e-6);
  const double h = -x0 / kGrowthSteps;
  const int n0 = L * kGrowthLane;
  // step n spans [x0 + n h, x0 + (n + 1) h]; the last step ends at exactly 0
  auto xat = [&](int n) { return (n == kGrowthSteps) ? 0.0 : x0 + n * h; };
  // phase 1: P = T_{n0 + 63} ... T_{n0}, column by column
  double p00 = 1.0, p10 = 0.0, p01 = 0.0, p11 = 1.0;
  double a1, b1;
  f.coef(xat(n0), a1, b1);
  for (int n = n0; n < n0 + kGrowthLane; ++n) {
    const double x = xat(n), xe = xat(n + 1), hh = xe - x;
    double a2, b2, a4, b4;
    f.coef(x + 0.5 * hh, a2, b2);
    f.coef(xe, a4, b4);
    rk4_lin(hh, a1, b1, a2, b2, a4, b4, p00, p10);
    rk4_lin(hh, a1, b1, a2, b2, a4, b4, p01, p11);
    a1 = a4;
    b1 = b4;
  }
  // phase 2: inclusive scan, lane L ends with M_L ... M_0
  for (int off = 1; off < 64; off <<= 1) {
    const double q00 = __shfl_up(p00, off), q10 = __shfl_up(p10, off);
    const double q01 = __shfl_up(p01, off), q11 = __shfl_up(p11, off);
    if (L >= off) {
      const double r00 = p00 * q00 + p01 * q10, r01 = p00 * q01 + p01 * q11;
      const double r10 = p10 * q00 + p11 * q10, r11 = p10 * q01 + p11 * q11;
      p00 = r00; p01 = r01; p10 = r10; p11 = r11;
    }
  }
  const double yi0 = exp(x0) + (2.0 / 3.0) * d.Or / c.Om, yi1 = exp(x0);
  // D(a = 1) from the whole product (lane 63)
  const double Dend = __shfl(p00 * yi0 + p01 * yi1, 63);
  // the state at the start of this lane's chunk: the exclusive product
  double e00 = __shfl_up(p00, 1), e10 = __shfl_up(p10, 1);
  double e01 = __shfl_up(p01, 1), e11 = __shfl_up(p11, 1);
  if (L == 0) { e00 = 1.0; e10 = 0.0; e01 = 0.0; e11 = 1.0; }
  double y0 = e00 * yi0 + e01 * yi1, y1 = e10 * yi0 + e11 * yi1;
  __syncthreads();                                 // xs[] complete
  // phase 3: this lane's slice nodes lie in (x_{n0}, x_{n0 + 64}] (slices are
  // ordered by increasing z, i.e. decreasing x); lane 0 also takes any node
  // below x0, as the first step of a sequential solve would
  const double xa = xat(n0), xb = xat(n0 + kGrowthLane);
  int next = n_z - 1;
  while (L > 0 && next >= 0 && xs[next] <= xa) --next;
  if (next >= 0 && xs[next] <= xb) {
    f.coef(xa, a1, b1);
    for (int n = n0; n < n0 + kGrowthLane && next >= 0 && xs[next] <= xb; ++n) {
      const double x = xat(n), xe = xat(n + 1), hh = xe - x;
      double a2, b2, a4, b4;
      f.coef(x + 0.5 * hh, a2, b2);
      f.coef(xe, a4, b4);
      double n0v = y0, n1v = y1;
      rk4_lin(hh, a1, b1, a2, b2, a4, b4, n0v, n1v);
      while (next >= 0 && xs[next] <= xe) {
        // cubic Hermite on [x, xe] with values y0, n0v and slopes y1, n1v
        const double s = (xs[next] - x) / hh;
        const double s2 = s * s, s3 = s2 * s;
        const double h00 = 2 * s3 - 3 * s2 + 1, h10 = s3 - 2 * s2 + s, h01 = -2 * s3 + 3 * s2,
                     h11 = s3 - s2;
        const double Dz = (h00 * y0 + h10 * hh * y1 + h01 * n0v + h11 * hh * n1v) / Dend;
        Sb[next * kSliceStride + kSlD] = Dz;
        Sb[next * kSliceStride + kSlDg] = Dz;             // the ODE growth is the bias growth
        --next;
      }
      y0 = n0v;
      y1 = n1v;
      a1 = a4;
      b1 = b4;
    }
  }
  if (L == 0) der[b].D0 = Dend;
  if (gdense) {
    // the state on every step node of this lane's chunk (and the end node),
    // normalised by D(a = 1): the dense table of D(z) for any z
    double* Gd = gdense + (size_t)b * (kGrowthSteps + 1) * 2;
    double g0 = e00 * yi0 + e01 * yi1, g1 = e10 * yi0 + e11 * yi1;
    double c1, d1;
    f.coef(xat(n0), c1, d1);
    const double iD = 1.0 / Dend;
    for (int n = n0; n < n0 + kGrowthLane; ++n) {
      Gd[2 * n] = g0 * iD;
      Gd[2 * n + 1] = g1 * iD;
      const double x = xat(n), xe = xat(n + 1), hh = xe - x;
      double a2, b2, a4, b4;
      f.coef(x + 0.5 * hh, a2, b2);
      f.coef(xe, a4, b4);
      rk4_lin(hh, c1, d1, a2, b2, a4, b4, g0, g1);
      c1 = a4;
      d1 = b4;
    }
    if (L == 63) {
      Gd[2 * kGrowthSteps] = g0 * iD;
      Gd[2 * kGrowthSteps + 1] = g1 * iD;
    }
  }
}

// ----------------------------------------------------------------- sigma --
// sigma(R) by the trapezoid over k = logspace(-5, 3, 4000) (sgl.py:458-460,
// 510-515), top-hat anchor at 8/h Mpc (sgl.py:517-538), smooth-k window
// Ws = 1/(1 + (0.43 x)^6) for sigma(M); dln sigma/dln M from the same sum
// with the analytic window derivative.
constexpr int kSigmaK = 4000;
constexpr int kSigChunk = 1000;   // k nodes staged in LDS at a time (16 KB)
// Window terms past y = (0.43 k R)^6 = kSigYCut are skipped: Ws^2 ~ y^-2
// falls by 5.7 % a node there while g grows about like k.  A CPU check over
// 1e8..1e16 Msun/h puts the skipped tail of either sum below 1e-17 of it.
constexpr double kSigYCut = 1e10;
static_assert(kSigmaK % kSigChunk == 0, "whole chunks");

// One workgroup per instance.  The k nodes go through LDS in chunks of
// kSigChunk (16 KB per workgroup, so that LDS does not cap occupancy): each
// chunk's integrand g = k^2 Pshape w_trap / (2 pi^2) and k^6 are formed, the
// top-hat anchor and every mass's window sums take their share of it, and
// the next chunk overwrites it.  The mass sums run on four independent
// accumulators (the FP64 add latency, not the issue rate, bounds one chain).
__global__ __launch_bounds__(256) void k_sigma(const Cosmo* __restrict__ cosmo,
                                               const double* __restrict__ m_nodes_h, int n_m,
                                               int m_in_msun, Derived* __restrict__ der,
                                               double* __restrict__ mass) {
  __shared__ double g[kSigChunk];
  __shared__ double k6[kSigChunk];
  __shared__ double red[256], pS[256], pSd[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  const EH98Const e = eh98_const(c);
  const double lstep = 8.0 / (kSigmaK - 1);
  const double kLn10 = 2.302585092994046;
  // trapezoid weights on k = 10^(-5 + lstep i): (k_{i+1} - k_{i-1}) / 2 =
  // k_i (r - 1/r) / 2 with r = 10^lstep, half intervals at the two ends
  const double r = exp(kLn10 * lstep), ir = 1.0 / r;
  const double R8 = 8.0 / c.h;
  const double rhom = c.Om * kRhoc0Unit * c.h * c.h;
  const double a6 = pow(0.43, 6.0);
  // G threads per mass when n_m < 256 (G = 256 / nm2, nm2 the power of two
  // >= n_m): thread (g, j) sums the g-th contiguous part of every chunk for
  // mass j; the parts are added in order afterwards.
  int nm2 = 1;
  while (nm2 < n_m) nm2 <<= 1;
  const int G = 256 / nm2;
  const int j = t % nm2, gq = t / nm2;
  const int ia = gq * kSigChunk / G, ib = (gq + 1) * kSigChunk / G;
  double aR = 0.0;
  int qcut = kSigmaK;   // first node with y = aR k^6 > kSigYCut
  if (j < n_m) {
    const double M = m_in_msun ? m_nodes_h[j] : m_nodes_h[j] / c.h;
    const double R = cbrt(3.0 * M / (4.0 * kPi * rhom));
    const double R2 = R * R;
    aR = a6 * R2 * R2 * R2;
    // log10 k > (ln kSigYCut - ln aR) / (6 ln 10), i.e. node index above
    const double qc = ((log(kSigYCut) - log(aR)) / (6.0 * kLn10) + 5.0) / lstep;
    qcut = (qc < kSigmaK) ? (int)fmax(ceil(qc), 0.0) : kSigmaK;
  }
  // S = sum g Ws^2 and, from dWs/dx = -6 y Ws^2 / x, dS/dln R = -12 sum g y Ws^3,
  // so dln sigma / dln M = (dS/dln R) / (2 S) / 3 = -2 sum g y Ws^3 / S
  double S0 = 0.0, S1 = 0.0, S2 = 0.0, S3 = 0.0;
  double D0 = 0.0, D1 = 0.0, D2 = 0.0, D3 = 0.0;
  double sa = 0.0;                                   // top-hat anchor partial
  for (int i0 = 0; i0 < kSigmaK; i0 += kSigChunk) {
    __syncthreads();                                 // the previous chunk is consumed
    for (int q = t; q < kSigChunk; q += blockDim.x) {
      const int i = i0 + q;
      const double lk = (-5.0 + lstep * i) * kLn10;  // ln k
      const double k = exp(lk);
      const double T = eh98_T(e, k, lk);
      const double wl = (i == 0) ? 0.5 * k * (r - 1.0)
                        : (i == kSigmaK - 1) ? 0.5 * k * (1.0 - ir) : 0.5 * k * (r - ir);
      const double gi = k * k * exp(c.ns * lk) * T * T * wl / (2.0 * kPi * kPi);
      g[q] = gi;
      const double k2 = k * k;
      k6[q] = k2 * k2 * k2;
      const double x = k * R8;
      double sx, cx;
      sincos(x, &sx, &cx);
      const double W = 3.0 * (sx - x * cx) / (x * x * x);
      sa += gi * W * W;
    }
    __syncthreads();
    if (j < n_m) {
      auto term = [&](int q, double& A, double& Dd) {
        const double y = aR * k6[q];
        const double den = 1.0 + y;
        // 1 / (1 + y): hardware reciprocal and two Newton steps (full FP64)
        double Ws = __builtin_amdgcn_rcp(den);
        Ws = Ws * (2.0 - den * Ws);
        Ws = Ws * (2.0 - den * Ws);
        const double gw = g[q] * Ws * Ws;
        A += gw;
        Dd += gw * y * Ws;
      };
      int q = ia;
      const int qe = min(ib, max(qcut - i0, ia));
      for (; q + 4 <= qe; q += 4) {
        term(q, S0, D0);
        term(q + 1, S1, D1);
        term(q + 2, S2, D2);
        term(q + 3, S3, D3);
      }
      for (; q < qe; ++q) term(q, S0, D0);
    }
  }
  red[t] = sa;
  if (j < n_m) {
    pS[t] = (S0 + S1) + (S2 + S3);
    pSd[t] = (D0 + D1) + (D2 + D3);
  }
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double Anorm = c.s8 / sqrt(red[0]);
  if (t == 0) der[b].Anorm = Anorm;
  if (gq == 0 && j < n_m) {
    const double M = m_in_msun ? m_nodes_h[j] : m_nodes_h[j] / c.h;
    double S = 0.0, Sd = 0.0;
    for (int q = 0; q < G; ++q) {
      S += pS[q * nm2 + j];
      Sd += pSd[q * nm2 + j];
    }
    double* Mr = mass + ((size_t)b * n_m + j) * kMassStride;
    Mr[0] = M;
    Mr[1] = Anorm * sqrt(S);
    Mr[2] = -2.0 * Sd / S;
    Mr[3] = log(Anorm) + 0.5 * log(S);           // ln sigma(M, 0), read by k_rows
  }
}

// sigma(M, 0) on M = logspace(0, 17, kMstarN) Msun for the M*(z) root of the
// halo ellipticity (SGL_FLAG_ELL; sgl_full.py:551-567): the same smooth-k
// window sums as k_sigma on the same 4000 k nodes (the kernel-shape tables;
// the reference grid uses the reference's own spline, k_ref_mstar_table).
// Runs after k_sigma (A_norm).
__global__ __launch_bounds__(256) void k_mstar_table(const Cosmo* __restrict__ cosmo,
                                                     const Derived* __restrict__ der,
                                                     double* __restrict__ sigtab) {
  __shared__ double g[kSigChunk];
  __shared__ double k6[kSigChunk];
  constexpr int kPer = (kMstarN + 255) / 256;
  const int b = blockIdx.x, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  const EH98Const e = eh98_const(c);
  const double lstep = 8.0 / (kSigmaK - 1);
  const double kLn10 = 2.302585092994046;
  const double r = exp(kLn10 * lstep), ir = 1.0 / r;
  const double rhom = c.Om * kRhoc0Unit * c.h * c.h;
  const double a6 = pow(0.43, 6.0);
  double aR[kPer], S[kPer];
#pragma unroll
  for (int m = 0; m < kPer; ++m) {
    const int q = t + 256 * m;
    const double M = exp(kLn10 * 17.0 * q / (kMstarN - 1));
    const double R = cbrt(3.0 * M / (4.0 * kPi * rhom)), R2 = R * R;
    aR[m] = a6 * R2 * R2 * R2;
    S[m] = 0.0;
  }
  for (int i0 = 0; i0 < kSigmaK; i0 += kSigChunk) {
    __syncthreads();
    for (int q = t; q < kSigChunk; q += blockDim.x) {
      const int i = i0 + q;
      const double lk = (-5.0 + lstep * i) * kLn10;
      const double k = exp(lk);
      const double T = eh98_T(e, k, lk);
      const double wl = (i == 0) ? 0.5 * k * (r - 1.0)
                        : (i == kSigmaK - 1) ? 0.5 * k * (1.0 - ir) : 0.5 * k * (r - ir);
      g[q] = k * k * exp(c.ns * lk) * T * T * wl / (2.0 * kPi * kPi);
      const double k2 = k * k;
      k6[q] = k2 * k2 * k2;
    }
    __syncthreads();
    for (int q = 0; q < kSigChunk; ++q) {
      const double gq = g[q], kq = k6[q];
#pragma unroll
      for (int m = 0; m < kPer; ++m) {
        const double den = 1.0 + aR[m] * kq;
        double Ws = __builtin_amdgcn_rcp(den);
        Ws = Ws * (2.0 - den * Ws);
        Ws = Ws * (2.0 - den * Ws);
        S[m] += gq * Ws * Ws;
      }
    }
  }
  const double A = der[b].Anorm;
#pragma unroll
  for (int m = 0; m < kPer; ++m) {
    const int q = t + 256 * m;
    if (q < kMstarN) sigtab[(size_t)b * kMstarN + q] = A * sqrt(S[m]);
  }
}

// The reference's sigma(M) spline (sgl.py:536-538, SGL_FLAG_SUBHALOS): knots
// M = logspace(4, 17, 500) Msun (10^(4 + i 13/499), ends exact), filled here,
// then their window sums by k_sigma (two launches of 250 masses), then
// scipy's not-a-knot CubicSpline slopes in x = ln M (scipy/interpolate/
// _cubic.py: the tridiagonal slope system with the not-a-knot end rows),
// solved by one thread per instance with the Thomas algorithm.
__global__ void k_sigspl_nodes(double* __restrict__ m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kSigSplN)
    m[i] = (i == 0) ? 1e4 : (i == kSigSplN - 1) ? 1e17
                                                : pow(10.0, i * (13.0 / (kSigSplN - 1)) + 4.0);
}

__global__ void k_sigspl_solve(double* __restrict__ spl, int B) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double* S = spl + (size_t)b * kSigSplDoubles;
  double* x = S + kSigSplN * kMassStride;
  double* y = x + kSigSplN;
  double* s = y + kSigSplN;
  constexpr int n = kSigSplN;
  for (int i = 0; i < n; ++i) {
    x[i] = log(S[i * kMassStride]);
    y[i] = S[i * kMassStride + 3];
  }
  // rows: lower l[i] s[i-1] + diag d[i] s[i] + upper u[i] s[i+1] = r[i]; the
  // forward sweep keeps the modified upper diagonal in s[] and the right side
  // in the caller-visible slope array's place until the back substitution
  auto h = [&](int i) { return x[i + 1] - x[i]; };
  auto dl = [&](int i) { return (y[i + 1] - y[i]) / h(i); };
  double cp[n];
  double rp;
  {
    const double h0 = h(0), h1 = h(1), d = x[2] - x[0];
    const double diag = h1, up = d;
    const double r = ((h0 + 2.0 * d) * h1 * dl(0) + h0 * h0 * dl(1)) / d;
    cp[0] = up / diag;
    s[0] = r / diag;
  }
  for (int i = 1; i < n - 1; ++i) {
    const double lo = h(i), di = 2.0 * (h(i - 1) + h(i)), up = h(i - 1);
    const double r = 3.0 * (h(i) * dl(i - 1) + h(i - 1) * dl(i));
    const double den = di - lo * cp[i - 1];
    cp[i] = up / den;
    s[i] = (r - lo * s[i - 1]) / den;
  }
  {
    const double hm = h(n - 2), hm1 = h(n - 3), d = x[n - 1] - x[n - 3];
    const double lo = d, di = hm1;
    const double r = (hm * hm * dl(n - 3) + (2.0 * d + hm) * hm1 * dl(n - 2)) / d;
    const double den = di - lo * cp[n - 2];
    rp = (r - lo * s[n - 2]) / den;
  }
  s[n - 1] = rp;
  for (int i = n - 2; i >= 0; --i) s[i] -= cp[i] * s[i + 1];
}

// sigma(M, 0) on M = logspace(0, 17, kMstarN) Msun from the reference's
// spline (the reference-grid tables): Cosmology.Mstar tabulates sigmaM, the
// not-a-knot spline on 1e4..1e17 Msun extrapolated below 1e4 by its first
// piece, and keeps the largest suffix on which it strictly decreases
// (sgl_full.py:551-567).  The entries before that suffix are set to its first
// value, so ln_mstar's plateau rule returns the suffix's first mass for any
// target at or above it, as np.interp clamps to its first node.
__global__ __launch_bounds__(256) void k_ref_mstar_table(const double* __restrict__ sigspl,
                                                         double* __restrict__ sigtab) {
  __shared__ int last_up;
  const int t = threadIdx.x;
  const double* spl = sigspl + (size_t)blockIdx.x * kSigSplDoubles;
  double* tab = sigtab + (size_t)blockIdx.x * kMstarN;
  const double kLn10 = 2.302585092994046;
  if (t == 0) last_up = -1;
  for (int q = t; q < kMstarN; q += blockDim.x) {
    const double lnM = (q == 0) ? 0.0 : (q == kMstarN - 1) ? 17.0 * kLn10
                                                         : kLn10 * (17.0 * q / (kMstarN - 1));
    tab[q] = sigma_spl(spl, lnM);
  }
  __syncthreads();
  for (int q = t; q < kMstarN - 1; q += blockDim.x)
    if (tab[q + 1] >= tab[q]) atomicMax(&last_up, q);
  __syncthreads();
  const int i0 = last_up + 1;
  const double top = tab[i0];
  __syncthreads();
  for (int q = t; q < i0; q += blockDim.x) tab[q] = top;
}

// ------------------------------------------------------------------ rows --
// One workgroup per (instance, slice), threads over masses: the slice's
// redshift-only factors (Ludlow+16 parameters, rho_c, Delta_vir) are formed
// once in LDS.  a_st and dc0 are the Sheth-Tormen normalisation and the
// bias delta_c constant, computed on the host.  Rows of slices past the
// instance's n_z are not written (no kernel reads them).
__global__ __launch_bounds__(256) void k_rows(const Cosmo* __restrict__ cosmo,
                                              const Derived* __restrict__ der,
                                              const double* __restrict__ slices,
                                              const double* __restrict__ mass, int n_z_max,
                                              int n_m, double dlnM, double a_st, double dc0,
                                              int edge_rvir, const double* __restrict__ sigtab,
                                              double* __restrict__ rows,
                                              double* __restrict__ frows) {
  __shared__ double sk[15];
  // (sk[8] = rho_c0 for the filament rows)
  const int i = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  if (i >= (int)c.n_z) return;
  double* Rs = rows + ((size_t)b * n_z_max + i) * n_m * kRowStride;
  double* Fs = frows ? frows + ((size_t)b * n_z_max + i) * n_m * kFilRowStride : nullptr;
  const double* S = slices + ((size_t)b * n_z_max + i) * kSliceStride;
  if (t < 9) {
    // Ludlow+16 with the z <= 7 clamp (sgl.py:590-606) and the Bryan-Norman
    // Delta_vir at Omega_m(z) (subhalos.py:381-385).  The four powers a^e,
    // the two cube roots and the three logs are one log and one exp on nine
    // lanes: lane q takes v = f_q exp(e_q ln u_q) (q < 6) or ln u_q.
    const Derived d = der[b];
    const double z = S[kSlZ], E2 = S[kSlE2];
    const double zc = fmin(z, 7.0), a = 1.0 + zc;
    const double poly = 4.135 - 0.564 * a - 0.210 * a * a + 0.0557 * a * a * a -
                        0.00348 * a * a * a * a;
    const double rhoc = d.rhoc0 * E2;
    const double dd = c.Om * (1.0 + z) * (1.0 + z) * (1.0 + z) / E2 - 1.0;
    const double Dvir = 18.0 * kPi * kPi + 82.0 * dd - 39.0 * dd * dd;
    const double f = (t == 0) ? 3.395 : (t == 1) ? 0.307 : (t == 2) ? 0.628 : (t == 3) ? 0.317 : 1.0;
    const double e = (t == 0) ? -0.215 : (t == 1) ? 0.540 : (t == 2) ? -0.047
                     : (t == 3) ? -0.893 : 1.0 / 3.0;
    const int slot = (t < 4) ? t + 1 : (t == 4) ? 9 : (t == 5) ? 12 : (t == 6) ? 10
                     : (t == 7) ? 11 : 13;
    const double u = (t < 4) ? a
                     : (t == 4) ? 200.0 / Dvir
                     : (t == 5) ? 3.0 / (4.0 * kPi * 200.0 * rhoc)
                     : (t == 6) ? poly : (t == 7) ? S[kSlD] : S[kSlDg];
    const double lu = log(u);
    sk[slot] = (t < 6) ? f * exp(e * lu) : lu;
    if (t == 0) {
      sk[0] = poly;
      sk[5] = rhoc;                              // rho_c(z)
      sk[6] = Dvir;
      sk[7] = d.rhom;
      sk[8] = d.rhoc0;
    }
    // ln M*(z) for the halo ellipticity (sigtab is null without it)
    if (t == 1) sk[14] = sigtab ? ln_mstar(sigtab + (size_t)b * kMstarN, dc0 / S[kSlDg]) : 0.0;
  }
  __syncthreads();
  // D: growth of sigma(M, z) in the mass function; Dg: growth in the bias and
  // beta (sgl_full.py:541-548, 1861), the same as D on the ODE growth
  const double D = S[kSlD], Dg = S[kSlDg];
  const double c0 = sk[1], beta_ = sk[2], g1 = sk[3], g2 = sk[4];
  const double Dvir = sk[6], rhom = sk[7], rhoc0 = sk[8], cr = sk[9];
  const double lpoly = sk[10], lD = sk[11], r200f = sk[12], rhoc = sk[5], lDg = sk[13];
  const double lmstar = sk[14];
  const double ldc = log(kDeltaC), lqst = log(kQst), l08dc0 = log(0.8) + 2.0 * log(dc0);
  for (int j = t; j < n_m; j += blockDim.x) {
    double* R = Rs + (size_t)j * kRowStride;
    const double* Mr = mass + ((size_t)b * n_m + j) * kMassStride;
    const double M = Mr[0], sig0 = Mr[1], dls = Mr[2], lsig0 = Mr[3];
    // powers as exp(b ln a) from the logs of sigma, D and the Ludlow
    // polynomial (relative error ~1e-15 for these exponents and arguments)
    const double lsig = lsig0 + lD;
    // Sheth-Tormen dn/dlnM (sgl.py:558-567)
    const double sig = sig0 * D;
    const double nu = (kDeltaC / sig) * (kDeltaC / sig);
    const double qn = kQst * nu;
    const double lqn = lqst + 2.0 * (ldc - lsig);
    const double mult = a_st * (1.0 + exp(-kPst * lqn)) * sqrt(qn / (2.0 * kPi)) * exp(-0.5 * qn);
    const double dndlnM = (rhom / M) * mult * (-2.0 * dls);
    // concentration: c0 q^-g1 (1 + q^(1/beta))^(-beta (g2 - g1)), q = delta_c / (sigma0 poly)
    const double lq = ldc - lsig0 - lpoly;
    const double C = c0 * exp(-g1 * lq - beta_ * (g2 - g1) * log1p(exp(lq / beta_)));
    // nfw_params (sgl.py:608-615)
    const double r200 = cbrt(M) * r200f;
    const double rs = r200 / C;
    const double fC = log1p(C) - C / (1.0 + C);
    const double rhos = (200.0 / 3.0) * rhoc * C * C * C / fC;
    const double ks = rhos * rs / S[kSlScr];
    // r_vir (subhalos.py:381-385); truncation at r200 (x_t = c) or at r_vir
    // (sgl_full.py:1874-1880)
    const double eta = eta_rvir(C, Dvir, fC, cr);
    const double rvir = eta * C * rs;
    // linear halo bias with the bias growth Dg (sgl_full.py:541-543)
    const double qn2 = 0.8 * (dc0 / Dg / sig0) * (dc0 / Dg / sig0);
    const double bh = 1.0 + (qn2 - 1.0) / dc0 +
                      0.6 / (dc0 * (1.0 + exp(0.3 * (l08dc0 - 2.0 * (lsig0 + lDg)))));
    R[kRowKs] = ks;
    R[kRowRs] = rs;
    R[kRowXt] = edge_rvir ? eta * C : C;
    R[kRowRvir] = rvir;
    R[kRowW] = S[kSlWz] * dndlnM * dlnM;
    R[kRowBeta] = Dg * bh;
    // halo ellipticity eps = max(0, (1 - s) / (1 + s)), s = 0.54 (M / M*)^-0.05
    // (Allgood+06, sgl_full.py:570-579); 0 without it
    double eps = 0.0;
    if (sigtab) {
      const double s = 0.54 * exp(-0.05 * (log(M) - lmstar));
      eps = fmax(0.0, (1.0 - s) / (1.0 + s));
    }
    R[6] = eps;
    R[7] = 0.0;
    if (Fs) {
      // filament row (sgl_full.py:1984-2002): flat-barrier mass function
      // (p, q, A) = (0, 0.7, 1/2) (filaments.py:43-50), r_s = (M/1e14)^(1/3)
      // Mpc, L = 20 r_s, k0base = r_s 14.4 rho_c0 L / Sigma_cr, bias of the
      // flat barrier (sgl_full.py:546-548)
      const double qf = 0.7 * nu;
      const double multf = 0.5 * 2.0 * sqrt(qf / (2.0 * kPi)) * exp(-0.5 * qf);
      const double dnf = (rhom / M) * multf * (-2.0 * dls);
      const double rsf = cbrt(M / 1e14);
      const double qb = 0.7 * (dc0 / Dg / sig0) * (dc0 / Dg / sig0);
      double* F = Fs + (size_t)j * kFilRowStride;
      F[0] = rsf * 14.4 * rhoc0 * (20.0 * rsf) / S[kSlScr];
      F[1] = rsf;
      F[2] = S[kSlWz] * dnf * dlnM;
      F[3] = Dg * (1.0 + (qb - 1.0) / dc0);
    }
  }
}

// ------------------------------------------------------------ covariance --
// C_ij = Cov of slice averages of the 1-D field smoothed with a top hat of
// Rs = 20 Mpc: modes k_q = 2 pi q / L, L = 1.05 chi_s, q = 1..max(4, L/Rs),
// P_1D(k) = (1/2pi) int_k^Kmax K P0(K) W_TH^2(K Rs) dK from k_1 / 2 to
// 200 / Rs (sgl_full.py:2514-2538).  The reference integrates with the
// trapezoid on 20000 log-spaced K and interpolates the running integral
// linearly at k_q.  Here the integral runs in s = ln K on kCovN uniform
// nodes with the fourth-order interval rule (the cubic through four
// neighbouring nodes), and the running integral is interpolated at ln k_q by
// the cubic Hermite with its slope g = K^2 P0 W^2: a CPU check against a
// 400000-node trapezoid puts this at 1.6e-10 of max |C| with 1000 nodes,
// where the reference's own rule is 5.6e-8 from it.
constexpr int kCovN = 1000;
constexpr int kCovQChunk = 32;
constexpr int kCovPad = (kMaxSlices + 3) / 4 * 4;   // slices padded to whole 4-blocks

__global__ __launch_bounds__(256) void k_cov(const Cosmo* __restrict__ cosmo,
                                             const Derived* __restrict__ der, int n_z_max,
                                             int* __restrict__ status, double* __restrict__ cov) {
  __shared__ double part[256];
  __shared__ double gs[kCovN], cum[kCovN];
  __shared__ double Yc[kCovQChunk][kCovPad], Ys[kCovQChunk][kCovPad];
  __shared__ double rwq[kCovQChunk], hsq[kCovQChunk];
  const int b = blockIdx.x, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  const Derived d = der[b];
  const int n_z = (int)c.n_z;
  const double Rs = 20.0;
  const double L = 1.05 * d.chis;
  int nq = (int)floor(L / Rs);
  if (nq < 4) nq = 4;
  if (nq > kCovMaxModes) {
    if (t == 0) atomicOr(status, kStatusCovModes);
    nq = kCovMaxModes;
  }
  const double k1 = 2.0 * kPi / L;
  const double K0 = 0.5 * k1, K1 = 200.0 / Rs;
  const double s0 = log(K0), s1 = log(K1);
  const double hs = (s1 - s0) / (kCovN - 1);
  const EH98Const e = eh98_const(c);
  // integrand in s = ln K: g = K f = K^2 P0(K) W^2(K Rs)
  for (int i = t; i < kCovN; i += blockDim.x) {
    const double lK = (i == kCovN - 1) ? s1 : s0 + hs * i;
    const double K = (i == kCovN - 1) ? K1 : exp(lK);
    const double x = K * Rs;
    double W;
    if (x < 1e-2) {
      W = 1.0 - x * x / 10.0 * (1.0 - x * x / 28.0);
    } else {
      double sx, cx;
      sincos(x, &sx, &cx);
      W = 3.0 * (sx - x * cx) / (x * x * x);
    }
    const double T = eh98_T(e, K, lK);
    gs[i] = K * K * d.Anorm * d.Anorm * exp(c.ns * lK) * T * T * W * W;
  }
  __syncthreads();
  // interval integrals: h/24 (-g_{i-1} + 13 g_i + 13 g_{i+1} - g_{i+2}), one-sided
  // h/24 (9 g_0 + 19 g_1 - 5 g_2 + g_3) at the ends; each thread sums a run of
  // intervals, the runs are scanned in order, then every thread writes the
  // running integral over its run
  auto interval = [&](int i) {
    if (i == 0) return hs / 24.0 * (9.0 * gs[0] + 19.0 * gs[1] - 5.0 * gs[2] + gs[3]);
    if (i == kCovN - 2)
      return hs / 24.0 * (9.0 * gs[kCovN - 1] + 19.0 * gs[kCovN - 2] - 5.0 * gs[kCovN - 3] +
                          gs[kCovN - 4]);
    return hs / 24.0 * (-gs[i - 1] + 13.0 * gs[i] + 13.0 * gs[i + 1] - gs[i + 2]);
  };
  const int per = (kCovN - 1 + 255) / 256;
  const int i0 = t * per, i1 = min(kCovN - 1, i0 + per);
  double loc = 0.0;
  for (int i = i0; i < i1; ++i) loc += interval(i);
  part[t] = loc;
  __syncthreads();
  if (t == 0) {
    double run = 0.0;
    for (int q = 0; q < 256; ++q) { double v = part[q]; part[q] = run; run += v; }
  }
  __syncthreads();
  {
    double run = part[t];
    for (int i = i0; i < i1; ++i) { run += interval(i); cum[i + 1] = run; }
    if (t == 0) cum[0] = 0.0;
  }
  __syncthreads();
  const double total = cum[kCovN - 1];
  // C_ij = sum_q r_q^2 (c_qi c_qj + s_qi s_qj): per chunk of modes the
  // weighted differences Yc = r (sin k chi_{i+1} - sin k chi_i), Ys = r (cos
  // k chi_i - cos k chi_{i+1}) go to LDS once; each thread accumulates 4 x 4
  // blocks (i0.., j0..) of the upper triangle in registers, 2 FMAs per pair
  // and mode; C_ji is written from C_ij at the end.
  // With ntile <= 256 blocks, G = 256 / ntile thread groups split the modes
  // of every chunk (group gm takes modes gm, gm + G, ..) and are summed in
  // group order at the end; beyond 256 blocks each thread takes two.
  const int nt = (n_z + 3) / 4;                  // 4-slice blocks per side
  const int ntile = nt * (nt + 1) / 2;           // blocks of the upper triangle
  constexpr int kTilesPerThread = ((kMaxSlices + 3) / 4 * ((kMaxSlices + 3) / 4 + 1) / 2 + 255) / 256;
  const int G = (ntile <= 256) ? 256 / ntile : 1;
  const int gm = (ntile <= 256) ? t / ntile : 0;
  double acc[kTilesPerThread][16];
  int ti[kTilesPerThread], tj[kTilesPerThread];
  for (int a = 0; a < kTilesPerThread; ++a) {
    for (int e = 0; e < 16; ++e) acc[a][e] = 0.0;
    int p = (ntile <= 256) ? ((a == 0 && gm < G) ? t % ntile : ntile) : t + a * 256;
    int bi = 0;
    if (p < ntile) {                             // p -> (bi, bj), bi <= bj, row-major
      while (p >= nt - bi) { p -= nt - bi; ++bi; }
      ti[a] = 4 * bi;
      tj[a] = 4 * (bi + p);
    } else {
      ti[a] = tj[a] = -1;
    }
  }
  const double dchi = d.dchi;
  for (int q0 = 0; q0 < nq; q0 += kCovQChunk) {
    const int qn = min(kCovQChunk, nq - q0);
    if (t < qn) {
      const double kq = k1 * (q0 + t + 1);
      const double sq = log(kq);
      int j = (int)floor((sq - s0) / hs);
      j = max(0, min(j, kCovN - 2));
      const double u = (sq - (s0 + hs * j)) / hs, u2 = u * u, u3 = u2 * u;
      const double cv = (2.0 * u3 - 3.0 * u2 + 1.0) * cum[j] + (u3 - 2.0 * u2 + u) * hs * gs[j] +
                        (-2.0 * u3 + 3.0 * u2) * cum[j + 1] + (u3 - u2) * hs * gs[j + 1];
      const double P1D = (total - cv) / (2.0 * kPi);
      rwq[t] = sqrt(fmax(2.0 * P1D / L, 0.0)) / (kq * dchi);
      hsq[t] = 2.0 * sin(0.5 * kq * dchi);
    }
    __syncthreads();
    // Yc = r (sin k chi_{i+1} - sin k chi_i) = r 2 sin(k dchi / 2) cos(k chi_mid),
    // Ys = r (cos k chi_i - cos k chi_{i+1}) = r 2 sin(k dchi / 2) sin(k chi_mid)
    // (uniform slices chi_i = i dchi; padding slices i >= n_z are zero)
    for (int p = t; p < qn * kCovPad; p += blockDim.x) {
      const int ql = p / kCovPad, i = p % kCovPad;
      double yc = 0.0, ys = 0.0;
      if (i < n_z) {
        const double kq = k1 * (q0 + ql + 1);
        double sn, cs;
        sincos(kq * ((i + 0.5) * dchi), &sn, &cs);
        const double r = rwq[ql] * hsq[ql];
        yc = r * cs;
        ys = r * sn;
      }
      Yc[ql][i] = yc;
      Ys[ql][i] = ys;
    }
    __syncthreads();
    for (int a = 0; a < kTilesPerThread; ++a) {
      if (ti[a] < 0) continue;
      const int i0 = ti[a], j0 = tj[a];
      for (int ql = gm; ql < qn; ql += G) {
        double ci[4], si[4], cj[4], sj[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ci[e] = Yc[ql][i0 + e];
          si[e] = Ys[ql][i0 + e];
          cj[e] = Yc[ql][j0 + e];
          sj[e] = Ys[ql][j0 + e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
          for (int f = 0; f < 4; ++f) acc[a][4 * e + f] = fma(ci[e], cj[f], fma(si[e], sj[f], acc[a][4 * e + f]));
      }
    }
    __syncthreads();
  }
  double* C = cov + (size_t)b * n_z_max * n_z_max;
  for (int p = t; p < n_z_max * n_z_max; p += blockDim.x) C[p] = 0.0;
  if (G > 1) {                                   // sum the mode groups in order
    for (int e = 0; e < 16; ++e) {
      __syncthreads();
      part[t] = (ti[0] >= 0) ? acc[0][e] : 0.0;
      __syncthreads();
      if (gm == 0 && ti[0] >= 0) {
        double s = 0.0;
        for (int g = 0; g < G; ++g) s += part[t + g * ntile];
        acc[0][e] = s;
      }
    }
    if (gm != 0) ti[0] = -1;                      // group 0 writes the blocks
  }
  __syncthreads();
  for (int a = 0; a < kTilesPerThread; ++a) {
    if (ti[a] < 0) continue;
    for (int e = 0; e < 4; ++e)
      for (int f = 0; f < 4; ++f) {
        const int i = ti[a] + e, j = tj[a] + f;
        if (i < n_z && j < n_z && i <= j) {
          C[i * n_z_max + j] = acc[a][4 * e + f];
          C[j * n_z_max + i] = acc[a][4 * e + f];
        }
      }
  }
}

// Side stream of the table build, one per device, created on the first call
// that is not under graph capture (never during capture; without it the
// build runs serially on the caller's stream).  The growth ODE (one thread
// per instance, latency-bound) overlaps sigma(M), and the slice covariance
// overlaps the rows; the caller's stream waits for both before the build
// returns, so the result is ordered on the caller's stream as before.
// Host-side state: not for concurrent calls from several host threads.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t e_bg = nullptr, e_sig = nullptr, j_growth = nullptr, j_cov = nullptr;
  bool ok = false, tried = false;
};
constexpr int kMaxDevices = 64;
static SideStream g_side[kMaxDevices];

static SideStream* side_stream(hipStream_t st) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  SideStream& S = g_side[dev];
  if (S.ok) return &S;
  if (S.tried) return nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return nullptr;
  S.tried = true;
  if (hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking) != hipSuccess) return nullptr;
  hipEvent_t* ev[4] = {&S.e_bg, &S.e_sig, &S.j_growth, &S.j_cov};
  for (hipEvent_t* e : ev)
    if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) return nullptr;
  S.ok = true;
  return &S;
}

int tables_build(const double* cosmo, int B, const sgl_shape& sh, const double* m_nodes_h,
                 double* derived, double* slices, double* mass, double* rows, double* frows,
                 double* cov, double* sigtab, int* status, hipStream_t st) {
  const Cosmo* C = reinterpret_cast<const Cosmo*>(cosmo);
  Derived* D = reinterpret_cast<Derived*>(derived);
  const double dlnM = (log(sh.m_hi_h) - log(sh.m_lo_h)) / sh.n_m;
  SideStream* S = side_stream(st);
  const bool cov_on = (sh.flags & SGL_FLAG_CLUSTERING) != 0;
  bool fail = false;
  auto ck = [&](hipError_t e) { fail |= (e != hipSuccess); };
  hipLaunchKernelGGL(k_background, dim3(B), dim3(kBgThreads), 0, st, C, sh.n_z_max, D, slices);
  // growth needs the slices and Omega_r, Omega_DE (background); it writes
  // D(z) into the slices and D0 into the derived record
  hipStream_t sg = st;
  if (S) {
    ck(hipEventRecord(S->e_bg, st));
    ck(hipStreamWaitEvent(S->s, S->e_bg, 0));
    sg = S->s;
  }
  hipLaunchKernelGGL(k_growth, dim3(B), dim3(64), 0, sg, C, sh.n_z_max, D, slices,
                     (double*)nullptr);
  if (S) ck(hipEventRecord(S->j_growth, S->s));
  // sigma(M) needs the background only; it writes A_norm and the mass table
  hipLaunchKernelGGL(k_sigma, dim3(B), dim3(256), 0, st, C, m_nodes_h, sh.n_m, 0, D, mass);
  const bool ell = (sh.flags & SGL_FLAG_ELL) != 0;
  if (ell) hipLaunchKernelGGL(k_mstar_table, dim3(B), dim3(256), 0, st, C, D, sigtab);
  // the covariance needs chi and A_norm (background, sigma), not the growth
  if (cov_on) {
    hipStream_t sc = st;
    if (S) {
      ck(hipEventRecord(S->e_sig, st));
      ck(hipStreamWaitEvent(S->s, S->e_sig, 0));
      sc = S->s;
    }
    hipLaunchKernelGGL(k_cov, dim3(B), dim3(256), 0, sc, C, D, sh.n_z_max, status, cov);
    if (S) ck(hipEventRecord(S->j_cov, S->s));
  }
  // Sheth-Tormen normalisation 1 / (1 + 2^-p Gamma(1/2 - p) / sqrt(pi)), p = 0.3, and
  // the bias delta_c constant 0.6 (3 pi / 2)^(2/3)
  const double a_st = 1.0 / (1.0 + std::pow(2.0, -kPst) * std::tgamma(0.5 - kPst) / std::sqrt(kPi));
  const double dc0 = 0.6 * std::pow(1.5 * kPi, 2.0 / 3.0);
  if (S) ck(hipStreamWaitEvent(st, S->j_growth, 0));
  const int rows_tpb = std::min(256, (sh.n_m + 63) / 64 * 64);
  hipLaunchKernelGGL(k_rows, dim3(sh.n_z_max, B), dim3(rows_tpb), 0, st, C, D, slices, mass,
                     sh.n_z_max, sh.n_m, dlnM, a_st, dc0, (sh.flags & SGL_FLAG_EDGE_RVIR) ? 1 : 0,
                     ell ? (const double*)sigtab : (const double*)nullptr, rows,
                     (sh.flags & SGL_FLAG_FILAMENTS) ? frows : (double*)nullptr);
  if (S && cov_on) ck(hipStreamWaitEvent(st, S->j_cov, 0));
  if (fail) return SGL_ERR_HIP;
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

// ------------------------------------------------------------ limber lss --
// limber_lss_var (sgl_full.py:2541-2572): the Limber variance of a Gaussian
// large-scale-structure convergence with the chain's own linear P0,
//   v = int dchi W^2 Dg^2 int_{lmin/f_K(chi)}^{lmax/f_K(chi)} dk k P0(k) / (2 pi),
//   W = 3/2 Omega_m H0^2 (1 + z) f_K(chi) f_K(chi_s - chi) / f_K(chi_s),
// lmin = 1, lmax = 1e7, on the reference's own nodes: z(chi) linear in the
// table of chi on 400 uniform z in [0, z_s]; 600 uniform chi in
// [1e-4, 1 - 1e-4] chi_s; the k integral a cumulative trapezoid on 40000
// geometric k in [1e-7, 1e5], read by linear interpolation clamped at both
// ends; the chi integral a trapezoid.  Dg is the growth of the bias field:
// Carroll-Press-Turner normalised at z = 0 under the legacy growth, the ODE
// growth otherwise (sgl_full.py:521-538).  One workgroup of kBgThreads for
// instance 0 of the tables; kn and cum (kLssK doubles each, device memory)
// take the k nodes and the running integral; the variance goes to *var.
constexpr int kLssZ = 400, kLssChi = 600, kLssK = 40000;
constexpr int kLssPer = (kLssK + kBgThreads - 1) / kBgThreads;

// linear interpolation between the bracketing nodes of the increasing xp,
// clamped to fp[0] / fp[n - 1] outside [xp[0], xp[n - 1]] (np.interp)
__device__ inline double lin_interp(double x, const double* xp, const double* fp, int n) {
  if (!(x > xp[0])) return fp[0];
  if (x >= xp[n - 1]) return fp[n - 1];
  int lo = 0, hi = n - 1;
  while (hi - lo > 1) {
    const int m = (lo + hi) >> 1;
    if (xp[m] <= x) lo = m; else hi = m;
  }
  return (fp[hi] - fp[lo]) / (xp[hi] - xp[lo]) * (x - xp[lo]) + fp[lo];
}

__global__ __launch_bounds__(kBgThreads) void k_ref_lss(const Cosmo* __restrict__ cosmo,
                                                        const Derived* __restrict__ der,
                                                        const double* __restrict__ gd,
                                                        double* __restrict__ kn,
                                                        double* __restrict__ cum,
                                                        double* __restrict__ var) {
  __shared__ double chi_e[kPanels + 1];
  __shared__ double run[kBgThreads];
  __shared__ double zt[kLssZ], ct[kLssZ], xc[kLssChi], yc[kLssChi];
  const int t = threadIdx.x;
  const Cosmo c = cosmo[0];
  const Derived d = der[0];
  const double dh = kCkms / d.H0;
  const double zs = c.zs;
  const double dz = zs / kPanels;
  chi_panels(c, d.Or, d.OL, dh, dz, chi_e, run);
  // the k table: nodes 10^(-7 + 12 i / 39999) with exact end points, and
  // this thread's run of trapezoid intervals [i, i + 1], i in [n0, n1)
  const EH98Const e = eh98_const(c);
  const double l0 = log10(1e-7), l1 = log10(1e5);
  const double lstep = (l1 - l0) / (kLssK - 1);
  auto knode = [&](int i) {
    return (i == 0) ? 1e-7 : (i == kLssK - 1) ? 1e5 : pow(10.0, i * lstep + l0);
  };
  auto fk = [&](double K) {
    const double lK = log(K);
    const double T = eh98_T(e, K, lK);
    return K * (d.Anorm * d.Anorm * exp(c.ns * lK) * T * T) / (2.0 * kPi);
  };
  const int n0 = t * kLssPer, n1 = min(kLssK - 1, n0 + kLssPer);
  double acc = 0.0;
  if (n0 < kLssK) {
    double kp = knode(n0), fp = fk(kp);
    kn[n0] = kp;
    for (int i = n0; i < n1; ++i) {
      const double kq = knode(i + 1), fq = fk(kq);
      acc += 0.5 * (fq + fp) * (kq - kp);
      cum[i + 1] = acc;
      if (i + 1 < n0 + kLssPer) kn[i + 1] = kq;
      kp = kq;
      fp = fq;
    }
  }
  __syncthreads();                               // chi_e complete, run[] free
  run[t] = acc;
  __syncthreads();
  if (t == 0) {
    double r = 0.0;
    for (int q = 0; q < kBgThreads; ++q) {
      const double v = run[q];
      run[q] = r;
      r += v;
    }
    cum[0] = 0.0;
  }
  __syncthreads();
  for (int i = n0; i < n1; ++i) cum[i + 1] += run[t];
  // chi on the z table
  for (int i = t; i < kLssZ; i += kBgThreads) {
    const double z = (i == kLssZ - 1) ? zs : i * (zs / (kLssZ - 1));
    const int lo = min((int)(z / dz), kPanels - 1);
    zt[i] = z;
    ct[i] = chi_e[lo] + dh * inv_e_integral(c, d.Or, d.OL, lo * dz, z);
  }
  __syncthreads();                               // z table, k table complete
  const double chis = d.chis;
  const double H0 = d.H0 / kCkms, H02 = H0 * H0;
  const double a0 = chis * 1e-4, a1 = chis * (1.0 - 1e-4);
  const double cstep = (a1 - a0) / (kLssChi - 1);
  const bool flat = c.Ok == 0.0;
  const double fks = f_k(c.Ok, d.H0, chis);
  const bool legacy = c.growth_legacy != 0.0;
  const double g0 = cpt_growth(0.0, c.Om, c.Ok, d.Or, d.OL);
  for (int j = t; j < kLssChi; j += kBgThreads) {
    const double chi = (j == kLssChi - 1) ? a1 : j * cstep + a0;
    const double z = lin_interp(chi, ct, zt, kLssZ);
    const double tr = f_k(c.Ok, d.H0, chi);
    const double W = flat ? 1.5 * c.Om * H02 * (1.0 + z) * chi * (chis - chi) / chis
                          : 1.5 * c.Om * H02 * (1.0 + z) * tr * f_k(c.Ok, d.H0, chis - chi) / fks;
    const double D = legacy ? cpt_growth(z, c.Om, c.Ok, d.Or, d.OL) / g0 : growth_dense_at(gd, z);
    const double I = lin_interp(1e7 / tr, kn, cum, kLssK) - lin_interp(1.0 / tr, kn, cum, kLssK);
    xc[j] = chi;
    yc[j] = W * W * (D * D) * I;
  }
  __syncthreads();
  if (t == 0) {
    double sum = 0.0;
    for (int j = 0; j + 1 < kLssChi; ++j) sum += (xc[j + 1] - xc[j]) * (yc[j + 1] + yc[j]) / 2.0;
    *var = sum;
  }
}

size_t ref_lss_doubles() { return 2 * (size_t)kLssK; }

int launch_ref_lss(const double* cosmo, const double* derived, const double* gdense, double* ws,
                   double* var, hipStream_t st) {
  hipLaunchKernelGGL(k_ref_lss, dim3(1), dim3(kBgThreads), 0, st,
                     reinterpret_cast<const Cosmo*>(cosmo),
                     reinterpret_cast<const Derived*>(derived), gdense, ws, ws + kLssK, var);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

// Tables on the reference grid (SGL_FLAG_REF_GRID): the reference-grid
// background, then the growth, sigma(M) (masses in Msun), M* and the rows
// kernels of the kernel-shape tables, with dlnM the engine grid's step.
int ref_tables_build(const double* cosmo, int B, const sgl_shape& sh, const double* m_nodes,
                     double* derived, double* slices, double* mass, double* rows, double* frows,
                     double* sigtab, double* sigspl, double* gdense, hipStream_t st) {
  const Cosmo* C = reinterpret_cast<const Cosmo*>(cosmo);
  Derived* D = reinterpret_cast<Derived*>(derived);
  const double dlnM = (log(1e17) - log(1e7)) / 99.0;
  const bool sub = (sh.flags & SGL_FLAG_SUBHALOS) != 0;
  const bool ell = (sh.flags & SGL_FLAG_ELL) != 0;
  hipLaunchKernelGGL(k_ref_background, dim3(B), dim3(kBgThreads), 0, st, C, sh.n_z_max, D, slices);
  // the dense growth table: the subhalo model's and the Limber term's D(z)
  hipLaunchKernelGGL(k_growth, dim3(B), dim3(64), 0, st, C, sh.n_z_max, D, slices, gdense);
  hipLaunchKernelGGL(k_sigma, dim3(B), dim3(256), 0, st, C, m_nodes, sh.n_m, 1, D, mass);
  if (sub || ell) {
    // spline knots into the x[] slot of instance 0 (overwritten by the solve)
    double* knots = sigspl + kSigSplN * kMassStride;
    hipLaunchKernelGGL(k_sigspl_nodes, dim3((kSigSplN + 255) / 256), dim3(256), 0, st, knots);
    for (int b = 0; b < B; ++b) {
      double* Sb = sigspl + (size_t)b * kSigSplDoubles;
      for (int half = 0; half < 2; ++half)
        hipLaunchKernelGGL(k_sigma, dim3(1), dim3(256), 0, st, C + b, knots + half * (kSigSplN / 2),
                           kSigSplN / 2, 1, D + b, Sb + half * (kSigSplN / 2) * kMassStride);
    }
    hipLaunchKernelGGL(k_sigspl_solve, dim3((B + 63) / 64), dim3(64), 0, st, sigspl, B);
  }
  if (ell) hipLaunchKernelGGL(k_ref_mstar_table, dim3(B), dim3(256), 0, st, sigspl, sigtab);
  const double a_st = 1.0 / (1.0 + std::pow(2.0, -kPst) * std::tgamma(0.5 - kPst) / std::sqrt(kPi));
  const double dc0 = 0.6 * std::pow(1.5 * kPi, 2.0 / 3.0);
  hipLaunchKernelGGL(k_rows, dim3(sh.n_z_max, B), dim3(128), 0, st, C, D, slices, mass,
                     sh.n_z_max, sh.n_m, dlnM, a_st, dc0, (sh.flags & SGL_FLAG_EDGE_RVIR) ? 1 : 0,
                     ell ? (const double*)sigtab : (const double*)nullptr, rows,
                     (sh.flags & SGL_FLAG_FILAMENTS) ? frows : (double*)nullptr);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

}  // namespace sgl
