// Subhalo decoration of the reference-shape path (SGL_FLAG_SUBHALOS): the
// per-host workspace layout and the device functions shared by the host /
// coarse-table kernels (sub.hip) and the cell emission (ref2d.hip).
#pragma once
#include "sgl_amd_internal.h"

namespace sgl {

// The reference's defaults (subhalos.py:353-355, 541; subhalos_exact.py:
// 1534-1536, 70-74)
constexpr double kSubAlpha = -0.82, kSubBeta = 50.0, kSubOmega = 4.0;
constexpr double kSubPsiRes = 1e-4;
constexpr double kSubX0 = 0.86, kSubBiasExp = 2.5;
constexpr double kSubMFloor = 1e7, kSubMhostMin = 1e9, kSubMmin = 1e7;
constexpr double kSubEpsRare = 3e-3;
constexpr double kSubG3TolT = 1e-10, kSubG3TolW = 1e-12, kSubMtWfloor = 1e-12, kSubKsRidge = 1e-10;
constexpr int kSubNpsi = 28;     // clump mass bins
constexpr int kSubNd = 90;       // ray-distance cells of _dgrid
constexpr int kSubNth = 48;      // ring angles
constexpr int kSubNR = 256;      // projected clump-density radii (_n3_proj_nb)
constexpr int kSubNu = 160;      // its line-of-sight nodes
constexpr int kSubNc = 128;      // coarse host radii
constexpr int kSubNbr = 3;       // rare bins
constexpr int kSubNq = 9 + 2 * kSubNbr;
constexpr int kSubNmt = 18;      // host-mass table (17 log nodes, plus M)
constexpr int kSubNtab = 17;
constexpr int kSubCells = kSubNpsi * kSubNd;

// Host record (doubles), one per (shell, mass) row of the reference grid.
constexpr int kHrValid = 0, kHrF = 1, kHrEta = 2, kHrMvir = 3, kHrXsub = 4, kHrDmax = 5,
              kHrRs = 6, kHrM = 7, kHrXref = 8, kHrNq = 9, kHrMs = 10, kHrKstar = 11,
              kHrNmt = 12, kHrXt = 13, kHrR200 = 14, kHrC = 15, kHrZ = 16, kHrKs = 17,
              kHrXin = 18, kHrG = 19;
constexpr int kHrDN = 32;                     // dN [28]
constexpr int kHrMc = kHrDN + kSubNpsi;       // clump mass m / v_r [28]
constexpr int kHrD = kHrMc + kSubNpsi;        // reach D = r200 of the clump [28]
constexpr int kHrRsc = kHrD + kSubNpsi;       // clump r_s [28]
constexpr int kHrKsc = kHrRsc + kSubNpsi;     // clump k_s [28]
constexpr int kHrCc = kHrKsc + kSubNpsi;      // clump c = x_t [28]
constexpr int kHrN2 = kHrCc + kSubNpsi;       // n2 [256], per Mpc^2
constexpr int kHrDc = kHrN2 + kSubNR;         // dc [90]
constexpr int kHrDA = kHrDc + kSubNd;         // dA [90]
constexpr int kHrCk = kHrDA + kSubNd;         // clump kappa [28][90] (cells dc < D)
constexpr int kHrCg = kHrCk + kSubCells;      // clump gamma [28][90]
constexpr int kHostStride = (kHrCg + kSubCells + 31) / 32 * 32;

// Coarse tables per row: ln x on the coarse radii, the point tables
// [kSubNc][kSubNq] (K offset, removed mass S, weight, A1, A2), and the
// rebuilt host's (ln M, ln r_s, ln k_s, eps) on the host-mass table.
constexpr int kCmLx = 0;
constexpr int kCmK = kCmLx + kSubNc;
constexpr int kCmS = kCmK + kSubNc * kSubNq;
constexpr int kCmW = kCmS + kSubNc * kSubNq;
constexpr int kCmA1 = kCmW + kSubNc * kSubNq;
constexpr int kCmA2 = kCmA1 + kSubNc * kSubNq;
constexpr int kCmLMt = kCmA2 + kSubNc * kSubNq;
constexpr int kCmLrs = kCmLMt + kSubNmt;
constexpr int kCmLks = kCmLrs + kSubNmt;
constexpr int kCmEpt = kCmLks + kSubNmt;
constexpr int kCoarseStride = (kCmEpt + kSubNmt + 31) / 32 * 32;
// ring and mean cosine [kSubNc][kSubNd] each
constexpr int kRingStride = 2 * kSubNc * kSubNd;

constexpr int kSubRows = kRefMaxShells * kRefMasses;
// after the host records: per (row, orientation) task the decorated inner and
// tail cell counts, the inner cell count and nq (4 ints, written by the count
// pass of the cell kernel, read by the point expansion)
constexpr int kSubTaskInts = 4;
constexpr size_t kSubTaskDoubles = ((size_t)kSubRows * kEllPhi * kSubTaskInts / 2 + 31) / 32 * 32;
struct SubLayout {
  size_t host, tasks, coarse, ring, total;
};
inline SubLayout sub_layout() {
  SubLayout L;
  L.host = 0;
  L.tasks = L.host + (size_t)kSubRows * kHostStride;
  L.coarse = L.tasks + kSubTaskDoubles;
  L.ring = L.coarse + (size_t)kSubRows * kCoarseStride;
  L.total = L.ring + (size_t)kSubRows * kRingStride;
  return L;
}

// np.geomspace(a, b, n)[i]: 10^(log10 a + i (log10 b - log10 a)/(n-1)), endpoints exact
__device__ inline double geom_np(double a, double b, int n, int i) {
  if (i == 0) return a;
  if (i == n - 1) return b;
  const double la = log10(a), lb = log10(b);
  return pow(10.0, la + i * ((lb - la) / (n - 1)));
}

// sigma(M, 0) of the reference's spline at ln M (scipy PPoly evaluation of the
// not-a-knot CubicSpline: the interval by searchsorted(x, lnM, "right") - 1,
// clamped to the end intervals, which extrapolate)
__device__ inline double sigma_spl(const double* __restrict__ spl, double lnM) {
  const double* x = spl + kSigSplN * kMassStride;
  const double* y = x + kSigSplN;
  const double* s = y + kSigSplN;
  int lo = 0, hi = kSigSplN - 1;   // last knot <= lnM
  if (lnM < x[0]) {
    hi = 1;
  } else {
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (x[mid] <= lnM) lo = mid; else hi = mid;
    }
  }
  int i = lo;
  if (i > kSigSplN - 2) i = kSigSplN - 2;
  const double h = x[i + 1] - x[i], d = (y[i + 1] - y[i]) / h;
  const double c3 = (s[i] + s[i + 1] - 2.0 * d) / (h * h);
  const double c2 = (3.0 * d - 2.0 * s[i] - s[i + 1]) / h;
  const double t = lnM - x[i];
  return exp(((c3 * t + c2) * t + s[i]) * t + y[i]);
}

// D(z) / D(0) of the growth ODE from the dense table (cubic Hermite on the
// RK4 step through x = -ln(1 + z))
__device__ inline double growth_dense_at(const double* __restrict__ gd, double z) {
  const double x0 = log(1e-6), h = -x0 / kGrowthSteps;
  const double x = -log1p(z);
  int n = (int)floor((x - x0) / h);
  if (n < 0) n = 0;
  if (n > kGrowthSteps - 1) n = kGrowthSteps - 1;
  const double xa = x0 + n * h, xb = (n + 1 == kGrowthSteps) ? 0.0 : x0 + (n + 1) * h;
  const double hh = xb - xa, s = (x - xa) / hh;
  const double s2 = s * s, s3 = s2 * s;
  return (2 * s3 - 3 * s2 + 1) * gd[2 * n] + (s3 - 2 * s2 + s) * hh * gd[2 * n + 1] +
         (-2 * s3 + 3 * s2) * gd[2 * n + 2] + (s3 - s2) * hh * gd[2 * n + 3];
}

// One point of a decorated host cell (subhalos_exact.py:802-925): the node-q
// tables interpolated linearly in ln x on the coarse radii (clamped), the
// host rebuilt at max(M - S, 1e7 Msun) with (ln r_s, ln k_s, eps) linear in
// ln M on the host-mass table (clamped), kept at the host's physical
// truncation radius, its (pseudo-elliptical) truncated kappa, gamma at the
// cell, then K = kappa_h + K_q, G = sqrt(max(gamma_h^2 + 2 gamma_h A1 + A2, 0)),
// weight factor w_q.  circ: the circular profile (phi = 0, eps = 0).
__device__ inline void sub_point(const double* __restrict__ H, const double* __restrict__ Cm,
                                 double x, double cph, double sph, bool circ, int q, double& kap,
                                 double& gam, double& wq) {
  const double* lxg = Cm + kCmLx;
  const double lx = fmin(fmax((x > exp(lxg[0]) ? log(x) : lxg[0]), lxg[0]), lxg[kSubNc - 1]);
  int lo = 0, hi = kSubNc - 1;
  while (lo < hi - 1) {
    const int mid = (lo + hi) >> 1;
    if (lxg[mid] <= lx) lo = mid; else hi = mid;
  }
  const double tw = (lx - lxg[lo]) / (lxg[lo + 1] - lxg[lo]), om = 1.0 - tw;
  const int a = lo * kSubNq + q, b = (lo + 1) * kSubNq + q;
  const double kp = om * Cm[kCmK + a] + tw * Cm[kCmK + b];
  const double sp = om * Cm[kCmS + a] + tw * Cm[kCmS + b];
  wq = om * Cm[kCmW + a] + tw * Cm[kCmW + b];
  const double a1 = om * Cm[kCmA1 + a] + tw * Cm[kCmA1 + b];
  const double a2 = om * Cm[kCmA2 + a] + tw * Cm[kCmA2 + b];
  const double M = H[kHrM], rs = H[kHrRs], xt = H[kHrXt];
  const int m_t = (int)H[kHrNmt];
  const double* lMt = Cm + kCmLMt;
  const double lMe = log(fmax(M - sp, kSubMmin));
  double rse, kse, epe;
  if (lMe <= lMt[0]) {
    rse = exp(Cm[kCmLrs]); kse = exp(Cm[kCmLks]); epe = Cm[kCmEpt];
  } else if (lMe >= lMt[m_t - 1]) {
    rse = exp(Cm[kCmLrs + m_t - 1]); kse = exp(Cm[kCmLks + m_t - 1]); epe = Cm[kCmEpt + m_t - 1];
  } else {
    int l = 0, h = m_t - 1;
    while (l < h - 1) {
      const int mid = (l + h) >> 1;
      if (lMt[mid] <= lMe) l = mid; else h = mid;
    }
    const double f = (lMe - lMt[l]) / (lMt[l + 1] - lMt[l]);
    rse = exp(Cm[kCmLrs + l] + f * (Cm[kCmLrs + l + 1] - Cm[kCmLrs + l]));
    kse = exp(Cm[kCmLks + l] + f * (Cm[kCmLks + l + 1] - Cm[kCmLks + l]));
    epe = Cm[kCmEpt + l] + f * (Cm[kCmEpt + l + 1] - Cm[kCmEpt + l]);
  }
  if (circ) {
    epe = 0.0;
    cph = 1.0;
    sph = 0.0;
  }
  const double xe = x * rs / rse, xte = xt * rs / rse;
  // pseudo-elliptical truncated NFW at the rebuilt host (kappagamma_eps)
  const double x1 = sqrt(fmax(1.0 - epe, 0.0)) * cph * xe, x2 = sqrt(1.0 + epe) * sph * xe;
  const double xe2 = x1 * x1 + x2 * x2;
  const double xl = fmax(sqrt(xe2), 1e-12);
  const double c2 = (x1 * x1 - x2 * x2) / fmax(xe2, 1e-300);
  double kh, gh;
  if (xl < xte) {
    double k0, g0;
    trunc_kg(xl, kse, xte, k0, g0);
    if (xl < 1e-4) g0 = kse;
    const double cg = c2 * g0;
    kh = k0 + epe * cg;
    gh = sqrt(fmax(g0 * g0 + 2.0 * epe * cg * k0 + epe * epe * (k0 * k0 - cg * cg), 0.0));
  } else {
    const double mut = log1p(xte) - xte / (1.0 + xte);
    const double gpm = kse * 4.0 * mut / fmax(xe * xe, 1e-300);
    kh = 0.0;
    gh = gpm;
  }
  const double g2 = gh * gh + 2.0 * gh * a1 + a2;
  kap = kh + kp;
  gam = (g2 > 0.0) ? sqrt(g2) : 0.0;
}

int sub_hosts_launch(const double* cosmo, const double* derived, const double* slices,
                     const double* mass, const double* rows, const double* spl,
                     const double* gdense, double* host, hipStream_t st);
int sub_coarse_launch(const double* cosmo, const double* derived, const double* slices,
                      const double* spl, const double* sigtab, double* host, double* coarse,
                      double* ring, hipStream_t st);

}  // namespace sgl
