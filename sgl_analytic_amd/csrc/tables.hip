// Lens-population tables on the device, batched over cosmologies.
//
// For every instance (cosmology + source redshift) the kernels below build
//   derived scalars     Omega_r, Omega_DE, H0, rho_c0, rho_m, chi(z_s), D_A(z_s), A_norm
//   slice table         z, chi, dchi, dchi (1+z)^2, D(z), Sigma_cr(z, z_s), E(z)
//   mass table          M, sigma(M, 0), dln sigma / dln M
//   row table           per (slice, mass): k_s, r_s, x_t = c, r_vir, w_row, beta
//   slice covariance    C_ij of the smoothed 1-D linear density field
// following sgl.Cosmology (sgl.py:65-615) on the make_cosmology configuration
// (pipeline.py:36-72), the builder's per-(z, M) quantities
// (sgl_full.py:1848-1885, 541-543; subhalos.py:381-385, 472-481) and
// shell_covariance (sgl_full.py:2514-2538).
#include <cmath>

#include "sgl_amd_internal.h"

namespace sgl {

// ------------------------------------------------------------ background --
// One workgroup per instance.  chi(z) = (c/H0) int dz / E on NPANEL uniform
// z-panels with 8-point Gauss-Legendre, fixed-order prefix sum, then Newton
// on each slice midpoint chi_i = (i + 1/2) chi_s / n_z.
constexpr int kPanels = 2048;
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

__global__ __launch_bounds__(256) void k_background(const Cosmo* __restrict__ cosmo, int n_z_max,
                                                    Derived* __restrict__ der,
                                                    double* __restrict__ slices) {
  __shared__ double chi_e[kPanels + 1];
  const int b = blockIdx.x, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  const double Or = c.Om / (1.0 + c.zeq);
  const double OL = 1.0 - c.Om - c.Ok - Or;
  const double H0 = 100.0 * c.h;
  const double dh = kCkms / H0;
  const double zs = c.zs;
  const int n_z = (int)c.n_z;
  const double dz = zs / kPanels;
  for (int p = t; p < kPanels; p += blockDim.x)
    chi_e[p + 1] = dh * inv_e_integral(c, Or, OL, p * dz, (p + 1) * dz);
  __syncthreads();
  if (t == 0) {
    chi_e[0] = 0.0;
    for (int p = 1; p <= kPanels; ++p) chi_e[p] += chi_e[p - 1];
  }
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
  for (int i = t; i < n_z_max; i += blockDim.x) {
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
    S[7] = 0.0;
  }
}

// ---------------------------------------------------------------- growth --
// D'' + (2 + dlnH/dlna) D' - 3/2 Omega_m(a) D = 0 in x = ln a (sgl.py:310-347),
// classical RK4 on kGrowthSteps uniform steps from a = 1e-6, cubic Hermite
// dense output at every slice node, normalised by D(a = 1).
constexpr int kGrowthSteps = 4096;

struct GrowthRhs {
  double Om, Ok, Or, OL, w, wa;
  bool lambda_de;
  __device__ inline void operator()(double x, double y0, double y1, double& d0, double& d1) const {
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
    double total = matter + rad + curv + de;
    double dh = -0.5 * (3.0 * matter + 4.0 * rad + 2.0 * curv + 3.0 * (1.0 + wz) * de) / total;
    d0 = y1;
    d1 = -(2.0 + dh) * y1 + 1.5 * (matter / total) * y0;
  }
};

__global__ __launch_bounds__(64) void k_growth(const Cosmo* __restrict__ cosmo, int B, int n_z_max,
                                               Derived* __restrict__ der,
                                               double* __restrict__ slices) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const Cosmo c = cosmo[b];
  const Derived d = der[b];
  GrowthRhs f{c.Om, c.Ok, d.Or, d.OL, c.w0, c.wa, c.w0 == -1.0 && c.wa == 0.0};
  const int n_z = (int)c.n_z;
  const double x0 = log(1e-6);
  const double h = -x0 / kGrowthSteps;
  double y0 = exp(x0) + (2.0 / 3.0) * d.Or / c.Om, y1 = exp(x0);
  double* Sb = slices + (size_t)b * n_z_max * kSliceStride;
  int next = n_z - 1;  // slices are ordered by increasing z, i.e. decreasing x
  double xnext = (next >= 0) ? -log1p(Sb[next * kSliceStride + kSlZ]) : 1.0;
  for (int n = 0; n < kGrowthSteps; ++n) {
    double x = x0 + n * h;
    double k1a, k1b, k2a, k2b, k3a, k3b, k4a, k4b;
    f(x, y0, y1, k1a, k1b);
    f(x + 0.5 * h, y0 + 0.5 * h * k1a, y1 + 0.5 * h * k1b, k2a, k2b);
    f(x + 0.5 * h, y0 + 0.5 * h * k2a, y1 + 0.5 * h * k2b, k3a, k3b);
    f(x + h, y0 + h * k3a, y1 + h * k3b, k4a, k4b);
    double n0 = y0 + h / 6.0 * (k1a + 2.0 * k2a + 2.0 * k3a + k4a);
    double n1 = y1 + h / 6.0 * (k1b + 2.0 * k2b + 2.0 * k3b + k4b);
    double xe = (n == kGrowthSteps - 1) ? 0.0 : x + h;
    while (next >= 0 && xnext <= xe) {
      // cubic Hermite on [x, xe] with values y0, n0 and slopes y1, n1
      double hh = xe - x, s = (xnext - x) / hh;
      double s2 = s * s, s3 = s2 * s;
      double h00 = 2 * s3 - 3 * s2 + 1, h10 = s3 - 2 * s2 + s, h01 = -2 * s3 + 3 * s2,
             h11 = s3 - s2;
      Sb[next * kSliceStride + kSlD] = h00 * y0 + h10 * hh * y1 + h01 * n0 + h11 * hh * n1;
      --next;
      xnext = (next >= 0) ? -log1p(Sb[next * kSliceStride + kSlZ]) : 1.0;
    }
    y0 = n0;
    y1 = n1;
  }
  for (int i = 0; i < n_z; ++i) Sb[i * kSliceStride + kSlD] /= y0;
  der[b].D0 = y0;
}

// ----------------------------------------------------------------- sigma --
// sigma(R) by the trapezoid over k = logspace(-5, 3, 4000) (sgl.py:458-460,
// 510-515), top-hat anchor at 8/h Mpc (sgl.py:517-538), smooth-k window
// Ws = 1/(1 + (0.43 x)^6) for sigma(M); dln sigma/dln M from the same sum
// with the analytic window derivative.
constexpr int kSigmaK = 4000;

__global__ __launch_bounds__(256) void k_sigma(const Cosmo* __restrict__ cosmo,
                                               const double* __restrict__ m_nodes_h, int n_m,
                                               Derived* __restrict__ der,
                                               double* __restrict__ mass) {
  __shared__ double g[kSigmaK];   // k^2 Pshape w_trap / (2 pi^2)
  __shared__ double kk[kSigmaK];
  __shared__ double red[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  const EH98Const e = eh98_const(c);
  const double lstep = 8.0 / (kSigmaK - 1);
  for (int i = t; i < kSigmaK; i += blockDim.x) kk[i] = pow(10.0, -5.0 + lstep * i);
  __syncthreads();
  for (int i = t; i < kSigmaK; i += blockDim.x) {
    double k = kk[i];
    const double lk = (-5.0 + lstep * i) * 2.302585092994046;   // ln k
    double T = eh98_T(e, k, lk);
    double wl = (i == 0) ? 0.5 * (kk[1] - kk[0])
                : (i == kSigmaK - 1) ? 0.5 * (kk[i] - kk[i - 1]) : 0.5 * (kk[i + 1] - kk[i - 1]);
    g[i] = k * k * exp(c.ns * lk) * T * T * wl / (2.0 * kPi * kPi);
  }
  __syncthreads();
  // top-hat anchor
  const double R8 = 8.0 / c.h;
  double s = 0.0;
  for (int i = t; i < kSigmaK; i += blockDim.x) {
    const double x = kk[i] * R8;
    double sx, cx;
    sincos(x, &sx, &cx);
    const double W = 3.0 * (sx - x * cx) / (x * x * x);
    s += g[i] * W * W;
  }
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double Anorm = c.s8 / sqrt(red[0]);
  const double rhom = c.Om * kRhoc0Unit * c.h * c.h;
  if (t == 0) der[b].Anorm = Anorm;
  const double a6 = pow(0.43, 6.0);
  // the mass sums need k only as k^6 (y = a6 (k R)^6 = a6 R^6 k^6): kk becomes k^6
  __syncthreads();
  for (int i = t; i < kSigmaK; i += blockDim.x) {
    const double k2 = kk[i] * kk[i];
    kk[i] = k2 * k2 * k2;
  }
  __syncthreads();
  // G threads per mass when n_m < 256 (G = 256 / nm2, nm2 the power of two
  // >= n_m): thread (g, j) sums the g-th contiguous chunk of the k nodes for
  // mass j; the chunks are added in order afterwards.
  int nm2 = 1;
  while (nm2 < n_m) nm2 <<= 1;
  const int G = 256 / nm2;
  const int j = t % nm2, gq = t / nm2;
  const int ia = gq * kSigmaK / G, ib = (gq + 1) * kSigmaK / G;
  __shared__ double pS[256], pSd[256];
  // S = sum g Ws^2 and, from dWs/dx = -6 y Ws^2 / x, dS/dln R = -12 sum g y Ws^3,
  // so dln sigma / dln M = (dS/dln R) / (2 S) / 3 = -2 sum g y Ws^3 / S
  if (j < n_m) {
    const double M = m_nodes_h[j] / c.h;
    const double R = cbrt(3.0 * M / (4.0 * kPi * rhom));
    const double R2 = R * R;
    const double aR = a6 * R2 * R2 * R2;
    double S = 0.0, Sd = 0.0;
    for (int i = ia; i < ib; ++i) {
      const double y = aR * kk[i];
      // 1 / (1 + y): hardware reciprocal and two Newton steps (full FP64)
      const double den = 1.0 + y;
      double Ws = __builtin_amdgcn_rcp(den);
      Ws = Ws * (2.0 - den * Ws);
      Ws = Ws * (2.0 - den * Ws);
      const double gw = g[i] * Ws * Ws;
      S += gw;
      Sd += gw * y * Ws;
    }
    pS[t] = S;
    pSd[t] = Sd;
  }
  __syncthreads();
  if (gq == 0 && j < n_m) {
    const double M = m_nodes_h[j] / c.h;
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

// ------------------------------------------------------------------ rows --
__device__ inline double mu_nfw(double y) { return log1p(y) - y / (1.0 + y); }

// r_vir / r200 = eta solving m^3 / mu(m) = 200 c^3 / (mu(c) Delta_vir) on
// [c, 3c] (subhalos.py:409-441): safeguarded Newton on ln(m^3 / mu(m)).
// mu_c = mu(c) and cr = cbrt(200 / Delta_vir) come from the caller; Newton
// starts at m = c cr (the root if mu(m) = mu(c), a few per cent off).
__device__ inline double eta_rvir(double c, double Dvir, double mu_c, double cr) {
  double lt = log(200.0 * c * c * c / mu_c / Dvir);
  double lo = c, hi = 3.0 * c, m = fmin(fmax(c * cr, lo), hi);
  for (int it = 0; it < 80; ++it) {
    double mu = mu_nfw(m);
    double f = log(m * m * m / mu) - lt;
    double dmu = m / ((1.0 + m) * (1.0 + m));
    double fp = 3.0 / m - dmu / mu;
    double step = f / fp;
    // Newton is quadratic here: a step below 1e-13 m leaves an error far below
    // 1 ulp.  The test comes before the bracket update, so that a point where
    // f rounds to 0 (the step lands on the new bracket end) ends the solve
    // instead of falling back to bisection for the last bits.
    if (fabs(step) <= 1e-13 * m) { m -= step; break; }
    if (f < 0.0) lo = m; else hi = m;
    double mn = m - step;
    if (!(lo < mn && mn < hi)) mn = 0.5 * (lo + hi);
    if (hi - lo <= 4e-16 * m) { m = mn; break; }
    m = mn;
  }
  return m / c;
}

// One workgroup per (instance, slice), threads over masses: the slice's
// redshift-only factors (Ludlow+16 parameters, rho_c, Delta_vir) are formed
// once in LDS.  a_st and dc0 are the Sheth-Tormen normalisation and the
// bias delta_c constant, computed on the host.
__global__ __launch_bounds__(256) void k_rows(const Cosmo* __restrict__ cosmo,
                                              const Derived* __restrict__ der,
                                              const double* __restrict__ slices,
                                              const double* __restrict__ mass, int n_z_max,
                                              int n_m, double dlnM, double a_st, double dc0,
                                              int edge_rvir, double* __restrict__ rows,
                                              double* __restrict__ frows) {
  __shared__ double sk[13];
  // (sk[8] = rho_c0 for the filament rows)
  const int i = blockIdx.x, b = blockIdx.y, t = threadIdx.x;
  const Cosmo c = cosmo[b];
  double* Rs = rows + ((size_t)b * n_z_max + i) * n_m * kRowStride;
  double* Fs = frows ? frows + ((size_t)b * n_z_max + i) * n_m * kFilRowStride : nullptr;
  if (i >= (int)c.n_z) {
    for (int q = t; q < n_m * kRowStride; q += blockDim.x) Rs[q] = 0.0;
    if (Fs)
      for (int q = t; q < n_m * kFilRowStride; q += blockDim.x) Fs[q] = 0.0;
    return;
  }
  const double* S = slices + ((size_t)b * n_z_max + i) * kSliceStride;
  if (t == 0) {
    const Derived d = der[b];
    const double z = S[kSlZ], E2 = S[kSlE2];
    // Ludlow+16 with the z <= 7 clamp (sgl.py:590-606): redshift-only factors
    const double zc = fmin(z, 7.0), a = 1.0 + zc;
    sk[0] = 4.135 - 0.564 * a - 0.210 * a * a + 0.0557 * a * a * a - 0.00348 * a * a * a * a;
    sk[1] = 3.395 * pow(a, -0.215);
    sk[2] = 0.307 * pow(a, 0.540);
    sk[3] = 0.628 * pow(a, -0.047);
    sk[4] = 0.317 * pow(a, -0.893);
    sk[5] = d.rhoc0 * E2;                       // rho_c(z)
    // Bryan-Norman Delta_vir at Omega_m(z) (subhalos.py:381-385)
    const double dd = c.Om * (1.0 + z) * (1.0 + z) * (1.0 + z) / E2 - 1.0;
    sk[6] = 18.0 * kPi * kPi + 82.0 * dd - 39.0 * dd * dd;
    sk[7] = d.rhom;
    sk[8] = d.rhoc0;
    sk[9] = cbrt(200.0 / sk[6]);
    sk[10] = log(sk[0]);                         // ln of the Ludlow polynomial
    sk[11] = log(S[kSlD]);                       // ln D(z)
    sk[12] = cbrt(3.0 / (4.0 * kPi * 200.0 * sk[5]));
  }
  __syncthreads();
  const double D = S[kSlD];
  const double c0 = sk[1], beta_ = sk[2], g1 = sk[3], g2 = sk[4];
  const double Dvir = sk[6], rhom = sk[7], rhoc0 = sk[8], cr = sk[9];
  const double lpoly = sk[10], lD = sk[11], r200f = sk[12], rhoc = sk[5];
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
    // linear halo bias with the growth D (sgl_full.py:541-543)
    const double qn2 = 0.8 * (dc0 / D / sig0) * (dc0 / D / sig0);
    const double bh = 1.0 + (qn2 - 1.0) / dc0 + 0.6 / (dc0 * (1.0 + exp(0.3 * (l08dc0 - 2.0 * lsig))));
    R[kRowKs] = ks;
    R[kRowRs] = rs;
    R[kRowXt] = edge_rvir ? eta * C : C;
    R[kRowRvir] = rvir;
    R[kRowW] = S[kSlWz] * dndlnM * dlnM;
    R[kRowBeta] = D * bh;
    R[6] = 0.0;
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
      const double qb = 0.7 * (dc0 / D / sig0) * (dc0 / D / sig0);
      double* F = Fs + (size_t)j * kFilRowStride;
      F[0] = rsf * 14.4 * rhoc0 * (20.0 * rsf) / S[kSlScr];
      F[1] = rsf;
      F[2] = S[kSlWz] * dnf * dlnM;
      F[3] = D * (1.0 + (qb - 1.0) / dc0);
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
                 double* cov, int* status, hipStream_t st) {
  const Cosmo* C = reinterpret_cast<const Cosmo*>(cosmo);
  Derived* D = reinterpret_cast<Derived*>(derived);
  const double dlnM = (log(sh.m_hi_h) - log(sh.m_lo_h)) / sh.n_m;
  SideStream* S = side_stream(st);
  const bool cov_on = (sh.flags & SGL_FLAG_CLUSTERING) != 0;
  bool fail = false;
  auto ck = [&](hipError_t e) { fail |= (e != hipSuccess); };
  hipLaunchKernelGGL(k_background, dim3(B), dim3(256), 0, st, C, sh.n_z_max, D, slices);
  // growth needs the slices and Omega_r, Omega_DE (background); it writes
  // D(z) into the slices and D0 into the derived record
  hipStream_t sg = st;
  if (S) {
    ck(hipEventRecord(S->e_bg, st));
    ck(hipStreamWaitEvent(S->s, S->e_bg, 0));
    sg = S->s;
  }
  hipLaunchKernelGGL(k_growth, dim3((B + 63) / 64), dim3(64), 0, sg, C, B, sh.n_z_max, D, slices);
  if (S) ck(hipEventRecord(S->j_growth, S->s));
  // sigma(M) needs the background only; it writes A_norm and the mass table
  hipLaunchKernelGGL(k_sigma, dim3(B), dim3(256), 0, st, C, m_nodes_h, sh.n_m, D, mass);
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
                     sh.n_z_max, sh.n_m, dlnM, a_st, dc0, (sh.flags & SGL_FLAG_EDGE_RVIR) ? 1 : 0, rows,
                     (sh.flags & SGL_FLAG_FILAMENTS) ? frows : (double*)nullptr);
  if (S && cov_on) ck(hipStreamWaitEvent(st, S->j_cov, 0));
  if (fail) return SGL_ERR_HIP;
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

}  // namespace sgl
