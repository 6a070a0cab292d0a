// The xi_lin (Limber) two-point term of the reference-shape path
// (SGL_FLAG_CLUSTERING on the sgl_ref_* entry points): the "bias" ingredient
// of pipeline.run with clustering="xilin", xi_closure="gauss"
// (pipeline.py:135-207).
//
//   k_perp rule   42 Gauss-Legendre nodes in ln k on [1e-4, 1e4], weights
//                 c_j = w_j k_j^2 P0(k_j) / (2 pi) (xi_kperp_grid,
//                 sgl_full.py:2483-2511);
//   deposit       per (shell, coarse bin, m, j): H = sum beta w J_{2m}(k_j R)
//                 over the bin's points, R the comoving distance of a point's
//                 lens centre, J_{2m} per distinct radius in a table; the
//                 points are sorted by (shell, bin) (stable: emission order
//                 inside a bin), runs of one lens radius summed by a segmented
//                 wave scan, each sum times its radius' table row; the exact
//                 first moments lx = sum beta w kappa J0, sum beta w gamma J2
//                 and each halo row's analytic shear tail
//                 (Cells._xi_deposit / _xi_add_grouped / xi_add_shear_tail,
//                 sgl_full.py:1081-1130, 1397-1441, 1466-1476);
//   compression   per (shell, m): A = [H; lres] diag(sqrt(c / dchi)), the
//                 eigenvectors V of A^T A by a parallel cyclic Jacobi sweep in
//                 LDS, rows kept with s_r > 1e-6 s_0 and s_r > 1e-7 of the
//                 largest s_0 of any shell (xi_end, sgl_full.py:1478-1523; the
//                 prune of Lam, sgl_full.py:2801-2812): the kept rows are
//                 A V_keep, applied here as the rotation V of each shell's
//                 tables;
//   Psi_2         sum_groups sign_m / 2 sum_r G_r^2 with G_r = sum_b rows_rb
//                 f_m(b) + res_r f_res (_Qxi_gauss, sgl_full.py:2869-2951): on
//                 the reach probes directly (k_kappa or k_gamma = 0), on the
//                 101 x 101 coarse grid k = k_max s^2 as two batched FP64 MFMA
//                 contractions (gamma bins x J_m, then kappa bins x
//                 [cos; sin]), then the not-a-knot bicubic spline in
//                 s = sqrt(k / k_max) on the full grid (_Q_interpolated,
//                 sgl_full.py:2953-3022);
//   variance      sum over the m = 0 rows of (rows . kappa_c + res)^2
//                 (Lam.sigma, sgl_full.py:3437-3446).
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "ref_xi.h"
#include "sgl_device.h"

namespace sgl {

namespace {

constexpr double kPiX = 3.141592653589793;

// np.geomspace(a, b, n)[i], endpoints exact
__device__ inline double xi_geom(double a, double b, int n, int i) {
  if (i == 0) return a;
  if (i == n - 1) return b;
  const double la = log10(a), lb = log10(b);
  return pow(10.0, la + i * ((lb - la) / (n - 1)));
}
// np.linspace(lo, hi, n)[i]
__device__ inline double xi_lin(double lo, double hi, int n, int i) {
  if (n < 2) return lo;
  return (i == n - 1) ? hi : lo + i * ((hi - lo) / (n - 1));
}

// [J0, J2, J4] of x as the reference packs them (_bessel_pack_nb,
// sgl_full.py:1132-1183): the power series below x = 2 (3, 6 or 14 terms),
// J0, J1 and the upward recurrence above; J2 = J4 = 0 without tangential shear.
__device__ inline void xi_bessel(double x, bool tang, double& J0, double& J2, double& J4) {
  if (x < 2.0) {
    const double h = 0.5 * x, h2 = h * h;
    double t0 = 1.0, s0 = 1.0, t2 = 0.5, s2 = 0.5, t4 = 1.0 / 24.0, s4 = 1.0 / 24.0;
    const int nt = (x < 0.02) ? 3 : ((x < 0.3) ? 6 : 14);
    for (int k = 1; k < nt; ++k) {
      const double f = (-h2) / k;
      t0 *= f / k;
      s0 += t0;
      t2 *= f / (k + 2);
      s2 += t2;
      t4 *= f / (k + 4);
      s4 += t4;
    }
    J0 = s0;
    J2 = tang ? s2 * (h * h) : 0.0;
    J4 = tang ? s4 * (h * h * h * h) : 0.0;
    return;
  }
  const double j0v = j0(x);
  J0 = j0v;
  if (!tang) {
    J2 = J4 = 0.0;
    return;
  }
  const double inv = 1.0 / x, j1v = j1(x);
  const double j2v = 2.0 * inv * j1v - j0v;
  const double j3v = 4.0 * inv * j2v - j1v;
  J2 = j2v;
  J4 = 6.0 * inv * j3v - j2v;
}

// J_m(x), m = 0, 1, 2 (scipy's j0, j1, jv(2, .) in _Qxi_gauss): J2 by its
// series below x = 2 (the recurrence cancels there), 2 J1 / x - J0 above
__device__ inline double xi_jm(int m, double x) {
  if (m == 0) return j0(x);
  if (m == 1) return j1(x);
  if (x < 2.0) {
    const double h = 0.5 * x, h2 = h * h;
    double t = 0.5, s = 0.5;
    for (int k = 1; k < 14; ++k) {
      t *= (-h2) / (k * (k + 2.0));
      s += t;
    }
    return s * h2;
  }
  return 2.0 * j1(x) / x - j0(x);
}

__device__ inline int xi_nz(const double* cosmo) {
  return (int)reinterpret_cast<const Cosmo*>(cosmo)[0].n_z;
}

// ------------------------------------------------------------ init ----
// k_perp rule, coarse edges and centres (one workgroup).
__global__ __launch_bounds__(256) void k_xi_init(const double* __restrict__ cosmo,
                                                 const double* __restrict__ derived,
                                                 double* __restrict__ xs) {
  const XiState L = xi_state_layout();
  const int t = threadIdx.x;
  const Cosmo c = reinterpret_cast<const Cosmo*>(cosmo)[0];
  const Derived d = reinterpret_cast<const Derived*>(derived)[0];
  // Gauss-Legendre: node i of n by Newton on P_n from cos(pi (i + 3/4) / (n + 1/2)),
  // ascending, weight 2 / ((1 - x^2) P_n'(x)^2)
  if (t < kXiNk) {
    const int pd[8] = {2, 2, 12, 8, 6, 4, 4, 4};
    int dec = 0, i = t;
    while (i >= pd[dec]) i -= pd[dec++];
    const int n = pd[dec];
    const int ii = n - 1 - i;                 // descending guess index -> ascending node
    double x = cos(kPiX * (ii + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = x;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (x * p1 - p0) / (x * x - 1.0);
      const double dx = p1 / dp;
      x -= dx;
      if (fabs(dx) < 1e-16) {
        // one more evaluation at the converged node for the weight
        p0 = 1.0;
        p1 = x;
        for (int k = 2; k <= n; ++k) {
          const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
          p0 = p1;
          p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1.0);
        break;
      }
    }
    const double w = 2.0 / ((1.0 - x * x) * dp * dp);
    const double a = log(1e-4) + dec * log(10.0);
    const double b = fmin(a + log(10.0), log(1e4));
    const double lk = 0.5 * (a + b) + 0.5 * (b - a) * x;
    const double k = exp(lk);
    const EH98Const e = eh98_const(c);
    const double T = eh98_T(e, k, lk);
    const double P0 = d.Anorm * d.Anorm * exp(c.ns * lk) * T * T;
    xs[L.kp + t] = k;
    xs[L.c + t] = 0.5 * (b - a) * w * k * k * P0 / (2.0 * kPiX);
  }
  constexpr int npos = kXiKc / 2 + 1;   // 171 positive kappa edges
  for (int i = t; i < kXiKc + 1; i += 256)
    xs[L.ekc + i] = (i < npos) ? -xi_geom(1e-7, 30.0, npos, npos - 1 - i)
                               : xi_geom(1e-7, 30.0, npos, i - npos);
  for (int i = t; i < kXiGc + 1; i += 256) xs[L.egc + i] = xi_geom(1e-9, 30.0, kXiGc + 1, i);
  __syncthreads();
  for (int i = t; i < kXiKc; i += 256) {
    const double a = xs[L.ekc + i], b = xs[L.ekc + i + 1];
    xs[L.kmc + i] = (a * b > 0.0) ? ((b > 0.0) ? 1.0 : -1.0) * sqrt(fabs(a * b)) : 0.0;
  }
  for (int i = t; i < kXiGc; i += 256) xs[L.gmc + i] = sqrt(xs[L.egc + i] * xs[L.egc + i + 1]);
  int* meta = reinterpret_cast<int*>(xs + L.meta);
  if (t == 0) meta[kXmNz] = (int)c.n_z;
  for (int i = t; i < 8; i += 256) xs[L.scal + i] = 0.0;
}

// Inverse of the not-a-knot cubic spline's collocation matrix on
// s_i = i / 100, i = 0..100 (scipy make_interp_spline(s, y, k=3): knots
// [0 x4, s_2 .. s_98, 1 x4]): coefficients = Minv y.  Gauss-Jordan with
// partial pivoting in LDS, one workgroup.
__device__ inline double xi_knot(int i) {
  constexpr int n = kXiCoarse;
  if (i < 4) return 0.0;
  if (i > n) return 1.0;
  return xi_lin(0.0, 1.0, n, i - 2);
}
// the four nonzero cubic B-splines at s in [0, 1]: index of the first and
// their values (Cox-de Boor on the span mu, t[mu] <= s < t[mu + 1], the last
// span closed at s = 1)
__device__ inline int xi_basis(double s, double* N) {
  constexpr int n = kXiCoarse;
  int lo = 3, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (xi_knot(mid) <= s) lo = mid; else hi = mid;
  }
  const int mu = lo;
  double left[4], right[4];
  N[0] = 1.0;
  for (int j = 1; j <= 3; ++j) {
    left[j] = s - xi_knot(mu + 1 - j);
    right[j] = xi_knot(mu + j) - s;
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      const double den = right[r + 1] + left[j - r];
      const double tmp = (den != 0.0) ? N[r] / den : 0.0;
      N[r] = saved + right[r + 1] * tmp;
      saved = left[j - r] * tmp;
    }
    N[j] = saved;
  }
  return mu - 3;
}

__global__ __launch_bounds__(256) void k_xi_minv(double* __restrict__ xs) {
  constexpr int n = kXiCoarse;
  __shared__ double A[n][n];
  __shared__ int piv;
  __shared__ double pv;
  const XiState L = xi_state_layout();
  const int t = threadIdx.x;
  for (int q = t; q < n * n; q += 256) A[q / n][q % n] = 0.0;
  __syncthreads();
  if (t < n) {
    double N[4];
    const int f = xi_basis(xi_lin(0.0, 1.0, n, t), N);
    for (int j = 0; j < 4; ++j)
      if (f + j >= 0 && f + j < n) A[t][f + j] = N[j];
  }
  __syncthreads();
  // in-place Gauss-Jordan inversion with row pivoting; column swaps undone last
  __shared__ int perm[n];
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      int p = k;
      double best = fabs(A[k][k]);
      for (int i = k + 1; i < n; ++i)
        if (fabs(A[i][k]) > best) { best = fabs(A[i][k]); p = i; }
      piv = p;
      perm[k] = p;
    }
    __syncthreads();
    const int p = piv;
    if (p != k)
      for (int j = t; j < n; j += 256) {
        const double a = A[k][j];
        A[k][j] = A[p][j];
        A[p][j] = a;
      }
    __syncthreads();
    if (t == 0) pv = A[k][k];
    __syncthreads();
    const double inv = 1.0 / pv;
    for (int j = t; j < n; j += 256) A[k][j] = (j == k) ? inv : A[k][j] * inv;
    __syncthreads();
    for (int q = t; q < n * n; q += 256) {
      const int i = q / n, j = q % n;
      if (i == k) continue;
      const double f = A[i][k];
      if (j == k) continue;
      A[i][j] -= f * A[k][j];
    }
    __syncthreads();
    for (int i = t; i < n; i += 256)
      if (i != k) A[i][k] = -A[i][k] * A[k][k];
    __syncthreads();
  }
  // undo the row interchanges as column interchanges, last first
  for (int k = n - 1; k >= 0; --k) {
    const int p = perm[k];
    if (p != k)
      for (int i = t; i < n; i += 256) {
        const double a = A[i][k];
        A[i][k] = A[i][p];
        A[i][p] = a;
      }
    __syncthreads();
  }
  for (int q = t; q < n * n; q += 256) xs[L.minv + q] = A[q / n][q % n];
}

// ------------------------------------------------------- J table ----
// [J0, J2, J4](k_j |R|) of every lens radius, [n_cells][126].
__global__ void k_xi_jtab(long long n_cells, const double* __restrict__ cells,
                          const double* __restrict__ xs, double* __restrict__ JT) {
  const XiState L = xi_state_layout();
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_cells * kXiNk) return;
  const long long c = q / kXiNk;
  const int j = (int)(q - c * kXiNk);
  const double R = cells[2 * c];
  double J0, J2, J4;
  xi_bessel(xs[L.kp + j] * fabs(R), R > 0.0, J0, J2, J4);
  double* o = JT + (size_t)c * kXiVals;
  o[j] = J0;
  o[kXiNk + j] = J2;
  o[2 * kXiNk + j] = J4;
}

__global__ void k_xi_copy_int(const int* __restrict__ a, int* __restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void k_xi_iota(int n, int* __restrict__ v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = i;
}

// runs of the sorted keys: parts per run (none for the outside key), the
// multi-part runs' part slots
__global__ void k_xi_nparts(const int* __restrict__ ukey, const int* __restrict__ rlen,
                            const int* __restrict__ nruns, int nr_cap, int* __restrict__ nparts,
                            int* __restrict__ nmulti) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nr_cap) return;
  int np = 0;
  if (r < *nruns && ukey[r] != kXiKeyInvalid) np = (rlen[r] + kXiPart - 1) / kXiPart;
  nparts[r] = np;
  nmulti[r] = (np > 1) ? np : 0;
}

__global__ void k_xi_meta(const int* __restrict__ ukey, const int* __restrict__ nruns,
                          const int* __restrict__ nparts, const long long* __restrict__ poff,
                          double* __restrict__ xs) {
  const XiState L = xi_state_layout();
  int* meta = reinterpret_cast<int*>(xs + L.meta);
  const int nr = *nruns;
  const int valid = (nr > 0 && ukey[nr - 1] == kXiKeyInvalid) ? nr - 1 : nr;
  meta[kXmRuns] = valid;
  meta[kXmParts] = (nr > 0) ? (int)(poff[nr - 1] + nparts[nr - 1]) : 0;
}

// ---------------------------------------------------------- walk ----
// One wave per part of a run (<= kXiPart points of one (shell, bin) in
// emission order).  64 points at a time: segmented sums of (beta w,
// beta w kappa, beta w gamma) over runs of one lens radius (fixed shuffle
// tree), then per segment the radius' J row times the sums; lane l holds
// values l and 64 + l of the [J0 | J2 | J4] row.
__global__ __launch_bounds__(256) void k_xi_walk(
    const int* __restrict__ rlen, int n_runs, const long long* __restrict__ rstart,
    const int* __restrict__ nparts, const long long* __restrict__ poff,
    const long long* __restrict__ hoff, int total_parts, const int* __restrict__ sval,
    const double* __restrict__ K, const double* __restrict__ G, const double* __restrict__ W,
    const int* __restrict__ pcid, const double* __restrict__ cells, const double* __restrict__ JT,
    double* __restrict__ Hx, double* __restrict__ Hp, double* __restrict__ Lp,
    double* __restrict__ Wp) {
  const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wv >= total_parts) return;
  int lo = 0, hi = n_runs - 1;                  // last run with poff <= wv
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (poff[mid] <= wv) lo = mid; else hi = mid - 1;
  }
  const int r = lo;
  const long long p = wv - poff[r];
  const long long a = rstart[r] + p * kXiPart;
  const long long b = min(rstart[r] + (long long)rlen[r], a + kXiPart);
  double h0 = 0.0, h1 = 0.0, lk = 0.0, lg = 0.0, wsum = 0.0;
  for (long long base = a; base < b; base += 64) {
    const long long q = base + lane;
    const bool ok = q < b;
    int cid = -1;
    double vs = 0.0, vk = 0.0, vg = 0.0;
    if (ok) {
      const int idx = sval[q];
      cid = pcid[idx];
      wsum += W[idx];
      const double bw = W[idx] * cells[2 * (size_t)cid + 1];
      vs = bw;
      vk = bw * K[idx];
      vg = bw * G[idx];
    }
    const int up = __shfl_up(cid, 1);
    const bool head = ok && (lane == 0 || up != cid);
    // inclusive segmented scan (left + right at every step)
    bool f = head;
    for (int off = 1; off < 64; off <<= 1) {
      const double us = __shfl_up(vs, off), uk = __shfl_up(vk, off), ug = __shfl_up(vg, off);
      const int uf = __shfl_up((int)f, off);
      if (lane >= off) {
        if (!f) {
          vs = us + vs;
          vk = uk + vk;
          vg = ug + vg;
        }
        f = f || uf;
      }
    }
    const int dn = __shfl_down(cid, 1);
    const bool tail = ok && (lane == 63 || base + lane + 1 >= b || dn != cid);
    unsigned long long tm = __ballot(tail);
    while (tm) {
      const int lt = __ffsll((long long)tm) - 1;
      tm &= tm - 1;
      const int c = __shfl(cid, lt);
      const double S = __shfl(vs, lt), Sk = __shfl(vk, lt), Sg = __shfl(vg, lt);
      const double* J = JT + (size_t)c * kXiVals;
      const double v0 = J[lane];
      const double v1 = (lane < kXiVals - 64) ? J[64 + lane] : 0.0;
      h0 += S * v0;
      h1 += S * v1;
      lk += Sk * v0;
      lg += Sg * ((lane >= kXiNk) ? v0 : v1);
    }
  }
  // the part's lens rate (the points' weights, lane-strided then a fixed tree)
  for (int off = 32; off > 0; off >>= 1) wsum += __shfl_xor(wsum, off);
  if (lane == 0) Wp[wv] = wsum;
  double* H = (nparts[r] > 1) ? Hp + (size_t)(hoff[r] + p) * kXiVals : Hx + (size_t)r * kXiVals;
  H[lane] = h0;
  if (lane < kXiVals - 64) H[64 + lane] = h1;
  double* Lo = Lp + (size_t)wv * (2 * kXiNk);
  if (lane < kXiNk) Lo[lane] = lk;
  else Lo[lane] = lg;                           // 42 + j, j = lane - 42
  if (lane < 2 * kXiNk - 64) Lo[64 + lane] = lg;  // 42 + j, j = 22 + lane
}

// runs of more than one part: their parts added in order
__global__ void k_xi_combine(int n_runs, const int* __restrict__ nparts,
                             const long long* __restrict__ hoff, const double* __restrict__ Hp,
                             double* __restrict__ Hx) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= (long long)n_runs * kXiVals) return;
  const int r = (int)(q / kXiVals), v = (int)(q - (long long)r * kXiVals);
  const int np = nparts[r];
  if (np < 2) return;
  double s = 0.0;
  for (int p = 0; p < np; ++p) s += Hp[(size_t)(hoff[r] + p) * kXiVals + v];
  Hx[(size_t)r * kXiVals + v] = s;
}

// lens rate per run (the positive closure's W_b of a (shell, coarse bin)):
// its parts in order
__global__ void k_xi_wrun(int n_runs, const int* __restrict__ nparts,
                          const long long* __restrict__ poff, const double* __restrict__ Wp,
                          double* __restrict__ wrun) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_runs) return;
  double s = 0.0;
  for (long long p = poff[r]; p < poff[r] + nparts[r]; ++p) s += Wp[p];
  wrun[r] = s;
}

// occupied coarse kappa and gamma bins over every shell
__global__ void k_xi_occ(int n_runs, const int* __restrict__ ukey, int* __restrict__ occ) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_runs) return;
  const int b = ukey[r] % kXiBins;
  occ[b / kXiGc] = 1;
  occ[kXiKc + b % kXiGc] = 1;
}

// compact maps of the occupied bins and the first run of every shell
__global__ __launch_bounds__(256) void k_xi_maps(int n_runs, const int* __restrict__ ukey,
                                                 const int* __restrict__ occ,
                                                 double* __restrict__ xs) {
  const XiState L = xi_state_layout();
  int* meta = reinterpret_cast<int*>(xs + L.meta);
  const int t = threadIdx.x;
  if (t == 0) {
    int np = 0, nq = 0;
    for (int p = 0; p < kXiKc; ++p) {
      meta[kXmPmap + p] = occ[p] ? np : -1;
      if (occ[p]) meta[kXmPlist + np++] = p;
    }
    for (int q = 0; q < kXiGc; ++q) {
      meta[kXmQmap + q] = occ[kXiKc + q] ? nq : -1;
      if (occ[kXiKc + q]) meta[kXmQlist + nq++] = q;
    }
    meta[kXmNP] = np;
    meta[kXmNQ] = nq;
  }
  for (int s = t; s <= kRefMaxShells; s += 256) {
    const int key = s * kXiBins;
    int lo = 0, hi = n_runs;                    // first run with ukey >= key
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (ukey[mid] < key) lo = mid + 1; else hi = mid;
    }
    meta[kXmRfirst + s] = lo;
  }
}

// lx per shell: the parts' first moments in order, in 64 chunks per shell,
// then the chunks in order and every halo row's shear tail (sgl_full.py:
// 1466-1476): beta n_area 2 pi Gam J1(k R_m) / (k R_m), Gam = k_s 4 mu(x_t)
// (r_s (1 + z))^2, R_m = x_m r_s (1 + z), n_area = Nbar / (pi R_m^2).
constexpr int kXiLxChunks = 64;
__global__ __launch_bounds__(128) void k_xi_lx_chunks(const long long* __restrict__ poff,
                                                      const int* __restrict__ nparts,
                                                      const double* __restrict__ xs,
                                                      const double* __restrict__ Lp,
                                                      double* __restrict__ Lc) {
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const int s = blockIdx.x, ch = blockIdx.y, v = threadIdx.x;
  if (v >= 2 * kXiNk) return;
  const int r0 = meta[kXmRfirst + s], r1 = meta[kXmRfirst + s + 1];
  double acc = 0.0;
  if (r1 > r0) {
    const long long p0 = poff[r0], p1 = poff[r1 - 1] + nparts[r1 - 1];
    const long long per = (p1 - p0 + kXiLxChunks - 1) / kXiLxChunks;
    const long long a = p0 + ch * per, b = min(p1, a + per);
    for (long long p = a; p < b; ++p) acc += Lp[(size_t)p * (2 * kXiNk) + v];
  }
  Lc[((size_t)s * kXiLxChunks + ch) * (2 * kXiNk) + v] = acc;
}

__global__ __launch_bounds__(128) void k_xi_lx(const double* __restrict__ cosmo,
                                               const double* __restrict__ slices,
                                               const double* __restrict__ rows,
                                               const double* __restrict__ Lc,
                                               double* __restrict__ xs) {
  const XiState L = xi_state_layout();
  const int s = blockIdx.x, v = threadIdx.x;
  if (v >= 2 * kXiNk) return;
  double acc = 0.0;
  for (int ch = 0; ch < kXiLxChunks; ++ch) acc += Lc[((size_t)s * kXiLxChunks + ch) * (2 * kXiNk) + v];
  if (v >= kXiNk && s < xi_nz(cosmo)) {
    const double k = xs[L.kp + v - kXiNk];
    const double zp1 = 1.0 + slices[(size_t)s * kSliceStride + kSlZ];
    for (int jm = 0; jm < kRefMasses; ++jm) {
      const double* R = rows + ((size_t)s * kRefMasses + jm) * kRowStride;
      const double ks = R[kRowKs], rs = R[kRowRs], xt = R[kRowXt];
      if (!(rs > 0.0)) continue;
      const double mut4 = ks * 4.0 * (log(1.0 + xt) - xt / (1.0 + xt));
      const double gxt = mut4 / (xt * xt);
      const double xm = (gxt > 1e-8) ? xt * sqrt(gxt / 1e-8) : xt;
      const double Nbar = kPiX * R[kRowW] * (xm * rs) * (xm * rs);
      const double a = rs * zp1;
      const double Gam = mut4 * a * a, Rm = xm * a;
      const double n_area = Nbar / (kPiX * Rm * Rm);
      const double arg = k * Rm;
      acc += R[kRowBeta] * n_area * 2.0 * kPiX * Gam * j1(arg) / arg;
    }
  }
  xs[L.lx + (size_t)s * 2 * kXiNk + v] = acc;
}

// ------------------------------------------------- per-group tables ----
// One workgroup per (shell, m): lres, the Gram matrix of A = [H; lres] sw, its
// eigen-decomposition (parallel cyclic Jacobi), and the sums over the shell's
// bins that the reach probes and the variance need (before the rotation).
constexpr int kJacSweeps = 16;
constexpr int kGp = kXiNk + 1;                 // LDS row pitch

__global__ __launch_bounds__(256) void k_xi_group(const double* __restrict__ cosmo,
                                                  const double* __restrict__ slices,
                                                  const int* __restrict__ ukey,
                                                  const double* __restrict__ Hx,
                                                  double* __restrict__ xs) {
  __shared__ double Gm[kXiNk][kGp];
  __shared__ double Vm[kXiNk][kGp];
  __shared__ double tile[64][kGp];
  __shared__ double sw[kXiNk], lres[kXiNk], cs[kXiNk / 2], sn[kXiNk / 2];
  __shared__ int pp[kXiNk / 2], qq[kXiNk / 2], pl[kXiNk];
  __shared__ double red[256];
  __shared__ int conv;
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const int g = blockIdx.x, s = g / 3, m = g - 3 * s, t = threadIdx.x;
  const bool live = s < xi_nz(cosmo);
  const int r0 = live ? meta[kXmRfirst + s] : 0, r1 = live ? meta[kXmRfirst + s + 1] : 0;
  const double* kmc = xs + L.kmc;
  const double* gmc = xs + L.gmc;
  if (t < kXiNk) {
    sw[t] = live ? sqrt(xs[L.c + t] / slices[(size_t)s * kSliceStride + kSlDchi]) : 0.0;
    // lres = lx - sum_b centre_b H_b (m = 0: kappa centres, m = 1: gamma
    // centres), none for m = 2
    double l = 0.0;
    if (live && m < 2) {
      double cen = 0.0;
      for (int r = r0; r < r1; ++r) {
        const int b = ukey[r] - s * kXiBins;
        const double cb = (m == 0) ? kmc[b / kXiGc] : gmc[b % kXiGc];
        cen += cb * Hx[(size_t)r * kXiVals + m * kXiNk + t];
      }
      l = xs[L.lx + (size_t)s * 2 * kXiNk + m * kXiNk + t] - cen;
    }
    lres[t] = l;
  }
  __syncthreads();
  // Gram = sum over bins (in run order) + the lres row
  double acc[7];
  for (int u = 0; u < 7; ++u) acc[u] = 0.0;
  for (int c0 = r0; c0 < r1; c0 += 64) {
    const int nc = min(64, r1 - c0);
    __syncthreads();
    for (int q = t; q < nc * kXiNk; q += 256) {
      const int rr = q / kXiNk, j = q - rr * kXiNk;
      tile[rr][j] = Hx[(size_t)(c0 + rr) * kXiVals + m * kXiNk + j] * sw[j];
    }
    __syncthreads();
    for (int u = 0; u < 7; ++u) {
      const int e = t + 256 * u;
      if (e >= kXiNk * kXiNk) break;
      const int i = e / kXiNk, j = e - i * kXiNk;
      double sacc = acc[u];
      for (int rr = 0; rr < nc; ++rr) sacc += tile[rr][i] * tile[rr][j];
      acc[u] = sacc;
    }
  }
  __syncthreads();
  for (int u = 0; u < 7; ++u) {
    const int e = t + 256 * u;
    if (e >= kXiNk * kXiNk) break;
    const int i = e / kXiNk, j = e - i * kXiNk;
    // symmetric by construction: the (i, j) and (j, i) sums run in the same order
    Gm[i][j] = acc[u] + (lres[i] * sw[i]) * (lres[j] * sw[j]);
    Vm[i][j] = (i == j) ? 1.0 : 0.0;
  }
  if (t < kXiNk) pl[t] = t;
  __syncthreads();
  // parallel cyclic Jacobi: 21 disjoint pairs per step (round robin), 41
  // steps per sweep
  for (int sweep = 0; sweep < kJacSweeps; ++sweep) {
    // convergence: off-diagonal mass against the diagonal
    double off = 0.0, dia = 0.0;
    for (int e = t; e < kXiNk * kXiNk; e += 256) {
      const int i = e / kXiNk, j = e - i * kXiNk;
      if (i != j) off += Gm[i][j] * Gm[i][j];
      else dia += Gm[i][j] * Gm[i][j];
    }
    red[t] = off;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    const double offs = red[0];
    __syncthreads();
    red[t] = dia;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    if (t == 0) conv = (offs <= 1e-36 * red[0]) ? 1 : 0;
    __syncthreads();
    if (conv) break;
    for (int step = 0; step < kXiNk - 1; ++step) {
      if (t < kXiNk / 2) {
        const int a = pl[t], b = pl[kXiNk - 1 - t];
        const int p = min(a, b), q = max(a, b);
        pp[t] = p;
        qq[t] = q;
        const double apq = Gm[p][q];
        double c = 1.0, sv = 0.0;
        if (apq != 0.0) {
          const double tau = (Gm[q][q] - Gm[p][p]) / (2.0 * apq);
          const double tt = (tau >= 0.0) ? 1.0 / (tau + sqrt(1.0 + tau * tau))
                                         : -1.0 / (-tau + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + tt * tt);
          sv = tt * c;
        }
        cs[t] = c;
        sn[t] = sv;
      }
      __syncthreads();
      // columns of G and V: (x_p, x_q) <- (c x_p - s x_q, s x_p + c x_q)
      for (int e = t; e < kXiNk * (kXiNk / 2); e += 256) {
        const int k = e / kXiNk, i = e - k * kXiNk;
        const int p = pp[k], q = qq[k];
        const double c = cs[k], sv = sn[k];
        const double gp = Gm[i][p], gq = Gm[i][q];
        Gm[i][p] = c * gp - sv * gq;
        Gm[i][q] = sv * gp + c * gq;
        const double vp = Vm[i][p], vq = Vm[i][q];
        Vm[i][p] = c * vp - sv * vq;
        Vm[i][q] = sv * vp + c * vq;
      }
      __syncthreads();
      // rows of G
      for (int e = t; e < kXiNk * (kXiNk / 2); e += 256) {
        const int k = e / kXiNk, j = e - k * kXiNk;
        const int p = pp[k], q = qq[k];
        const double c = cs[k], sv = sn[k];
        const double gp = Gm[p][j], gq = Gm[q][j];
        Gm[p][j] = c * gp - sv * gq;
        Gm[q][j] = sv * gp + c * gq;
      }
      __syncthreads();
      if (t < kXiNk / 2) {
        Gm[pp[t]][qq[t]] = 0.0;
        Gm[qq[t]][pp[t]] = 0.0;
      }
      // round robin: slot 0 fixed, the others rotate by one
      int nxt = 0;
      if (t < kXiNk) nxt = (t == 0) ? pl[0] : ((t == 1) ? pl[kXiNk - 1] : pl[t - 1]);
      __syncthreads();
      if (t < kXiNk) pl[t] = nxt;
      __syncthreads();
    }
  }
  // eigenvalues descending (ties by index), eigenvectors as columns
  double* lam = xs + L.lam + (size_t)g * kXiNk;
  double* V = xs + L.V + (size_t)g * kXiNk * kXiNk;
  if (t < kXiNk) {
    const double li = Gm[t][t];
    int rank = 0;
    for (int j = 0; j < kXiNk; ++j) {
      const double lj = Gm[j][j];
      rank += (lj > li || (lj == li && j < t)) ? 1 : 0;
    }
    lam[rank] = li;
    for (int i = 0; i < kXiNk; ++i) V[i * kXiNk + rank] = Vm[i][t];
    xs[L.res + (size_t)g * kXiNk + t] = lres[t] * sw[t];
    if (m == 0) xs[L.sw + (size_t)s * kXiNk + t] = sw[t];
  }
}

// Loadings of the positive closure (pos_closure.population,
// pos_closure.py:405-421), one workgroup per shell, on the m = 0 group before
// the global prune: the kept rows (s_r > 1e-6 s_0) are (A V)_r, so the
// 'kappa' mode m_r = rows_r . kappa_c + res_r is V^T (sum_b A_b kappa_b + A_res)
// and r0_b = (m / |m|) . rows_b = A_b . (V m / |m|); phic_b = sgn r0_b / W_b
// (0 where W_b <= 0), sgn the sign of r0 . W; v0 = |V[:, 0]|, the leading
// k_perp profile (sgl_full.py:1520-1521).
__global__ __launch_bounds__(256) void k_xi_loading(const double* __restrict__ cosmo,
                                                    const double* __restrict__ slices,
                                                    const int* __restrict__ ukey,
                                                    const double* __restrict__ Hx,
                                                    const double* __restrict__ wrun,
                                                    const double* __restrict__ xs,
                                                    double* __restrict__ phic,
                                                    double* __restrict__ v0) {
  __shared__ double sw[kXiNk], vec[kXiNk], mm[kXiNk], y[kXiNk], red[256];
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const int s = blockIdx.x, t = threadIdx.x, g = 3 * s;
  if (s >= xi_nz(cosmo)) return;
  const int r0 = meta[kXmRfirst + s], r1 = meta[kXmRfirst + s + 1];
  const double* lam = xs + L.lam + (size_t)g * kXiNk;
  const double* V = xs + L.V + (size_t)g * kXiNk * kXiNk;
  const double s0 = sqrt(fmax(lam[0], 0.0));
  if (t < kXiNk) {
    sw[t] = sqrt(xs[L.c + t] / slices[(size_t)s * kSliceStride + kSlDchi]);
    double a = 0.0;
    for (int r = r0; r < r1; ++r) {
      const int b = ukey[r] - s * kXiBins;
      a += Hx[(size_t)r * kXiVals + t] * sw[t] * xs[L.kmc + b / kXiGc];
    }
    vec[t] = a + xs[L.res + (size_t)g * kXiNk + t];
    v0[(size_t)s * kXiNk + t] = fabs(V[t * kXiNk]);
  }
  __syncthreads();
  if (t < kXiNk) {
    const double sr = sqrt(fmax(lam[t], 0.0));
    double a = 0.0;
    if (s0 > 0.0 && sr > kXiSvdTol * s0)
      for (int j = 0; j < kXiNk; ++j) a += vec[j] * V[j * kXiNk + t];
    mm[t] = a;
  }
  __syncthreads();
  if (t < kXiNk) {
    double n2 = 0.0;
    for (int r = 0; r < kXiNk; ++r) n2 += mm[r] * mm[r];
    const double inv = 1.0 / sqrt(n2);
    double a = 0.0;
    for (int r = 0; r < kXiNk; ++r) a += V[t * kXiNk + r] * (mm[r] * inv);
    y[t] = a;
  }
  __syncthreads();
  double part = 0.0;
  for (int r = r0 + t; r < r1; r += 256) {
    double a = 0.0;
    for (int j = 0; j < kXiNk; ++j) a += Hx[(size_t)r * kXiVals + j] * sw[j] * y[j];
    phic[r] = a;
    part += a * wrun[r];
  }
  red[t] = part;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double sgn = (red[0] >= 0.0) ? 1.0 : -1.0;
  for (int r = r0 + t; r < r1; r += 256) {
    const double wb = wrun[r];
    phic[r] = (wb > 0.0) ? sgn * phic[r] / wb : 0.0;
  }
}

// keep: s_r > 1e-6 s_0 of the group and s_r > 1e-7 of the largest s_0 of any
// group; the dropped columns of V are zeroed (one workgroup)
__global__ __launch_bounds__(256) void k_xi_keep(double* __restrict__ xs) {
  __shared__ double red[256];
  const XiState L = xi_state_layout();
  const int t = threadIdx.x;
  double mx = 0.0;
  for (int g = t; g < kXiGroups; g += 256) mx = fmax(mx, sqrt(fmax(xs[L.lam + (size_t)g * kXiNk], 0.0)));
  red[t] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  const double s0max = red[0];
  if (t == 0) xs[L.scal + 1] = s0max;
  for (int e = t; e < kXiGroups * kXiNk; e += 256) {
    const int g = e / kXiNk, r = e - g * kXiNk;
    const double* lam = xs + L.lam + (size_t)g * kXiNk;
    const double s0 = sqrt(fmax(lam[0], 0.0)), sr = sqrt(fmax(lam[r], 0.0));
    const bool keep = s0 > 0.0 && sr > kXiSvdTol * s0 && sr > kXiPrune * s0max;
    if (!keep) {
      double* V = xs + L.V + (size_t)g * kXiNk * kXiNk;
      for (int i = 0; i < kXiNk; ++i) V[i * kXiNk + r] = 0.0;
    }
  }
}

// The rotated tables of every group: res' = V^T res, srow' = V^T sum_b A_b,
// W0' (m = 0, per kappa bin) and U' (per gamma bin), and the variance
// contribution sum_r (V^T (sum_b A_b kappa_b + res))_r^2 of the m = 0 groups.
__global__ __launch_bounds__(256) void k_xi_tables(const double* __restrict__ cosmo,
                                                   const double* __restrict__ slices,
                                                   const int* __restrict__ ukey,
                                                   const double* __restrict__ Hx,
                                                   double* __restrict__ xs,
                                                   double* __restrict__ varpart) {
  __shared__ double Vs[kXiNk][kGp];
  __shared__ double Uq[kXiGc][kGp];
  __shared__ double vec[3][kXiNk];
  __shared__ double sw[kXiNk];
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const int g = blockIdx.x, s = g / 3, m = g - 3 * s, t = threadIdx.x;
  const bool live = s < xi_nz(cosmo);
  const int r0 = live ? meta[kXmRfirst + s] : 0, r1 = live ? meta[kXmRfirst + s + 1] : 0;
  const int nP = meta[kXmNP], nQ = meta[kXmNQ];
  const double* V = xs + L.V + (size_t)g * kXiNk * kXiNk;
  for (int e = t; e < kXiNk * kXiNk; e += 256) Vs[e / kXiNk][e % kXiNk] = V[e];
  for (int e = t; e < kXiGc * kXiNk; e += 256) Uq[e / kXiNk][e % kXiNk] = 0.0;
  if (t < kXiNk)
    sw[t] = live ? sqrt(xs[L.c + t] / slices[(size_t)s * kSliceStride + kSlDchi]) : 0.0;
  __syncthreads();
  double* W0 = xs + L.W0 + (size_t)s * kXiKc * kXiNk;   // [p'][r], m = 0 only
  if (t < kXiNk) {
    const int j = t;
    double srow = 0.0, kmom = 0.0, gmom = 0.0, wp = 0.0;
    int pcur = -1;
    for (int r = r0; r < r1; ++r) {
      const int b = ukey[r] - s * kXiBins;
      const int p = b / kXiGc, q = b - p * kXiGc;
      const double A = Hx[(size_t)r * kXiVals + m * kXiNk + j] * sw[j];
      srow += A;
      kmom += A * xs[L.kmc + p];
      gmom += A * xs[L.gmc + q];
      Uq[meta[kXmQmap + q]][j] += A;
      if (m == 0) {
        if (p != pcur) {
          if (pcur >= 0) W0[(size_t)meta[kXmPmap + pcur] * kXiNk + j] = wp;
          pcur = p;
          wp = 0.0;
        }
        wp += A;
      }
    }
    if (m == 0 && pcur >= 0) W0[(size_t)meta[kXmPmap + pcur] * kXiNk + j] = wp;
    vec[0][j] = srow;
    vec[1][j] = xs[L.res + (size_t)g * kXiNk + j];
    // m = 0: the kappa moment (the variance); m = 1: the gamma moment (the
    // positive closure's shear target, pos_closure.py:443-446)
    vec[2][j] = ((m == 1) ? gmom : kmom) + vec[1][j];
  }
  __syncthreads();
  // rotate: out_r = sum_j x_j V[j][r]
  __shared__ double rot[3][kXiNk];
  if (t < 3 * kXiNk) {
    const int w = t / kXiNk, r = t - w * kXiNk;
    double a = 0.0;
    for (int j = 0; j < kXiNk; ++j) a += vec[w][j] * Vs[j][r];
    rot[w][r] = a;
  }
  __syncthreads();
  if (t < kXiNk) {
    xs[L.srow + (size_t)g * kXiNk + t] = rot[0][t];
    xs[L.res + (size_t)g * kXiNk + t] = rot[1][t];
  }
  if (t == 0) {
    double v = 0.0;
    if (m < 2)
      for (int r = 0; r < kXiNk; ++r) v += rot[2][r] * rot[2][r];
    varpart[g] = v;
  }
  // U' [q'][r] for every group, W0' [p'][r] (in place) for m = 0
  double* U = xs + L.U + (size_t)g * kXiGc * kXiNk;
  for (int e = t; e < nQ * kXiNk; e += 256) {
    const int q = e / kXiNk, r = e - q * kXiNk;
    double a = 0.0;
    for (int j = 0; j < kXiNk; ++j) a += Uq[q][j] * Vs[j][r];
    U[e] = a;
  }
  if (m == 0) {
    __syncthreads();
    for (int p0 = 0; p0 < nP; p0 += kXiGc) {
      const int np = min(kXiGc, nP - p0);
      __syncthreads();
      for (int e = t; e < np * kXiNk; e += 256) Uq[e / kXiNk][e % kXiNk] = W0[(size_t)p0 * kXiNk + e];
      __syncthreads();
      for (int e = t; e < np * kXiNk; e += 256) {
        const int q = e / kXiNk, r = e - q * kXiNk;
        double a = 0.0;
        for (int j = 0; j < kXiNk; ++j) a += Uq[q][j] * Vs[j][r];
        W0[(size_t)p0 * kXiNk + e] = a;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_xi_var(const double* __restrict__ varpart,
                                                double* __restrict__ xs) {
  const XiState L = xi_state_layout();
  if (threadIdx.x != 0) return;
  double v = 0.0, vg = 0.0;
  for (int g = 0; g < kXiGroups; g += 3) v += varpart[g];
  for (int g = 1; g < kXiGroups; g += 3) vg += varpart[g];
  xs[L.scal] = v;
  xs[L.scal + 4] = vg;
}

// zero W0 for shells without rows of their own kappa bins (set once per PDF)
__global__ void k_xi_zero(double* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ------------------------------------------------------ axis probes ----
// out[i] += Psi_2(k_i, 0) (axis 0) or Psi_2(0, k_i) (axis 1): one workgroup
// per probe, threads over (group, row), a fixed-order block sum.
__global__ __launch_bounds__(256) void k_xi_axis(const double* __restrict__ cosmo,
                                                 const double* __restrict__ xs,
                                                 const double* __restrict__ kp, int axis,
                                                 double* __restrict__ out) {
  __shared__ double red[2][256];
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const int i = blockIdx.x, t = threadIdx.x;
  const double k = kp[i];
  const int nz = xi_nz(cosmo), nP = meta[kXmNP], nQ = meta[kXmNQ];
  double qr = 0.0, qi = 0.0;
  for (int e = t; e < 3 * nz * kXiNk; e += 256) {
    const int g = e / kXiNk, r = e - g * kXiNk, s = g / 3, m = g - 3 * s;
    const double res = xs[L.res + (size_t)g * kXiNk + r];
    const double srow = xs[L.srow + (size_t)g * kXiNk + r];
    double Gr = 0.0, Gi = 0.0;
    double sign = (m == 0) ? 1.0 : 2.0 * ((m & 1) ? -1.0 : 1.0);
    if (axis == 0) {
      if (m != 0) continue;
      const double* W0 = xs + L.W0 + (size_t)s * kXiKc * kXiNk;
      for (int pq = 0; pq < nP; ++pq) {
        const double kap = xs[L.kmc + meta[kXmPlist + pq]];
        double sp, cp;
        sincos(k * kap, &sp, &cp);
        const double w = W0[(size_t)pq * kXiNk + r];
        Gr += cp * w;
        Gi += sp * w;
      }
      Gr -= srow;
      Gi += k * res;
    } else {
      const double* U = xs + L.U + (size_t)g * kXiGc * kXiNk;
      for (int qq = 0; qq < nQ; ++qq)
        Gr += xi_jm(m, k * xs[L.gmc + meta[kXmQlist + qq]]) * U[(size_t)qq * kXiNk + r];
      if (m == 0) Gr -= srow;
      else if (m == 1) Gr += 0.5 * k * res;
    }
    qr += 0.5 * sign * (Gr * Gr - Gi * Gi);
    qi += 0.5 * sign * 2.0 * Gr * Gi;
  }
  red[0][t] = qr;
  red[1][t] = qi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] += red[0][t + w];
      red[1][t] += red[1][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[2 * i] += red[0][0];
    out[2 * i + 1] += red[1][0];
  }
}

// ------------------------------------------------------- coarse grid ----
// boxes B[g][q'][p'][j] = A_bj = H_bj sw_j for the shell's bins, zero
// elsewhere (thread per (run, m, j), j fastest)
__global__ void k_xi_box(int n_runs, const int* __restrict__ ukey, const double* __restrict__ Hx,
                         const double* __restrict__ xs, int nP, int nQ, double* __restrict__ box) {
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n_runs * kXiVals) return;
  const int run = (int)(e / kXiVals);
  const int mj = (int)(e - (long long)run * kXiVals), m = mj / kXiNk, j = mj - m * kXiNk;
  const int key = ukey[run], s = key / kXiBins, b = key - s * kXiBins;
  const int p = meta[kXmPmap + b / kXiGc], q = meta[kXmQmap + b % kXiGc];
  const int g = 3 * s + m;
  box[(((size_t)g * nQ + q) * nP + p) * kXiNk + j] =
      Hx[(size_t)run * kXiVals + mj] * xs[L.sw + (size_t)s * kXiNk + j];
}

// J_m(kg_c gamma_q) [3][nQ][101] and [cos; sin](kk_c kappa_p) [202][nP]
__global__ void k_xi_jq(const double* __restrict__ xs, int nP, int nQ, double kk_max,
                        double kg_max, double* __restrict__ Jq, double* __restrict__ E) {
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int nc = kXiCoarse;
  if (e < 3 * nQ * nc) {
    const int m = e / (nQ * nc), rem = e - m * nQ * nc, q = rem / nc, c = rem - q * nc;
    const double s = xi_lin(0.0, 1.0, nc, c);
    Jq[e] = xi_jm(m, kg_max * s * s * xs[L.gmc + meta[kXmQlist + q]]);
  }
  const int e2 = e - 3 * nQ * nc;
  if (e2 >= 0 && e2 < nc * nP) {
    const int c = e2 / nP, p = e2 - c * nP;
    const double s = xi_lin(0.0, 1.0, nc, c);
    double sp, cp;
    sincos(kk_max * s * s * xs[L.kmc + meta[kXmPlist + p]], &sp, &cp);
    E[(size_t)c * nP + p] = cp;
    E[(size_t)(nc + c) * nP + p] = sp;
  }
}

// Psi_2 on the coarse grid from G[g][202][42 * 101] (rows: cos part, sin
// part): partial sums over chunks of groups, then the chunks in order
constexpr int kXiQChunks = 16;
__global__ void k_xi_qpart(const double* __restrict__ xs, const double* __restrict__ Gm,
                           int n_groups, double kk_max, double kg_max,
                           double* __restrict__ qpart) {
  const XiState L = xi_state_layout();
  constexpr int nc = kXiCoarse;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int ch = blockIdx.y;
  if (e >= nc * nc) return;
  const int ik = e / nc, ig = e - ik * nc;
  const double sk = xi_lin(0.0, 1.0, nc, ik), sg = xi_lin(0.0, 1.0, nc, ig);
  const double kk = kk_max * sk * sk, kg = kg_max * sg * sg;
  const int per = (n_groups + kXiQChunks - 1) / kXiQChunks;
  const int g0 = ch * per, g1 = min(n_groups, g0 + per);
  double qr = 0.0, qi = 0.0;
  for (int g = g0; g < g1; ++g) {
    const int m = g % 3;
    const double sign = (m == 0) ? 1.0 : 2.0 * ((m & 1) ? -1.0 : 1.0);
    const double* G = Gm + (size_t)g * (2 * nc) * (kXiNk * nc);
    const double* res = xs + L.res + (size_t)g * kXiNk;
    const double* srow = xs + L.srow + (size_t)g * kXiNk;
    double ar = 0.0, ai = 0.0;
    for (int r = 0; r < kXiNk; ++r) {
      double Gr = G[(size_t)ik * (kXiNk * nc) + r * nc + ig];
      double Gi = G[(size_t)(nc + ik) * (kXiNk * nc) + r * nc + ig];
      if (m == 0) {
        Gr -= srow[r];
        Gi += kk * res[r];
      } else if (m == 1) {
        Gr += 0.5 * kg * res[r];
      }
      ar += Gr * Gr - Gi * Gi;
      ai += 2.0 * Gr * Gi;
    }
    qr += 0.5 * sign * ar;
    qi += 0.5 * sign * ai;
  }
  qpart[((size_t)ch * nc * nc + e) * 2] = qr;
  qpart[((size_t)ch * nc * nc + e) * 2 + 1] = qi;
}

// Q coarse = sum of the chunks; spline coefficients C = Minv Q Minv^T
__global__ void k_xi_qsum(const double* __restrict__ qpart, double* __restrict__ xs) {
  const XiState L = xi_state_layout();
  constexpr int nc = kXiCoarse;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nc * nc) return;
  double qr = 0.0, qi = 0.0;
  for (int ch = 0; ch < kXiQChunks; ++ch) {
    qr += qpart[((size_t)ch * nc * nc + e) * 2];
    qi += qpart[((size_t)ch * nc * nc + e) * 2 + 1];
  }
  xs[L.qc + e] = qr;
  xs[L.qc + nc * nc + e] = qi;
}
// tmp[i][b] = sum_l Q[i][l] Minv[b][l]  (re and im), then C[a][b] = sum_i Minv[a][i] tmp[i][b]
__global__ void k_xi_spl1(const double* __restrict__ xs, double* __restrict__ tmp) {
  const XiState L = xi_state_layout();
  constexpr int nc = kXiCoarse;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * nc * nc) return;
  const int part = e / (nc * nc), rem = e - part * nc * nc, i = rem / nc, b = rem - i * nc;
  const double* Q = xs + L.qc + (size_t)part * nc * nc;
  const double* M = xs + L.minv;
  double a = 0.0;
  for (int l = 0; l < nc; ++l) a += Q[i * nc + l] * M[b * nc + l];
  tmp[e] = a;
}
__global__ void k_xi_spl2(const double* __restrict__ tmp, double* __restrict__ xs) {
  const XiState L = xi_state_layout();
  constexpr int nc = kXiCoarse;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * nc * nc) return;
  const int part = e / (nc * nc), rem = e - part * nc * nc, a = rem / nc, b = rem - a * nc;
  const double* M = xs + L.minv;
  const double* T = tmp + (size_t)part * nc * nc;
  double v = 0.0;
  for (int i = 0; i < nc; ++i) v += M[a * nc + i] * T[i * nc + b];
  xs[L.cs + (size_t)part * nc * nc + rem] = v;
}

// psi[i][l] = the bicubic spline at s = sqrt(k / k_max) on the full grid
// k_k = linspace(0, kk_max, nk), k_g = linspace(0, kg_max, ng)
__global__ void k_xi_psi(const double* __restrict__ xs, int nk, int ng, double kk_max,
                         double kg_max, double* __restrict__ psi) {
  const XiState L = xi_state_layout();
  constexpr int nc = kXiCoarse;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nk * ng) return;
  const int i = (int)(e / ng), l = (int)(e - (long long)i * ng);
  const double kk = xi_lin(0.0, kk_max, nk, i), kg = xi_lin(0.0, kg_max, ng, l);
  const double sk = sqrt(fmin(fmax(kk / kk_max, 0.0), 1.0));
  const double sg = sqrt(fmin(fmax(kg / kg_max, 0.0), 1.0));
  double Nk[4], Ng[4];
  const int fk = xi_basis(sk, Nk), fg = xi_basis(sg, Ng);
  const double* Cr = xs + L.cs;
  const double* Ci = Cr + nc * nc;
  double vr = 0.0, vi = 0.0;
  for (int a = 0; a < 4; ++a) {
    double rr = 0.0, ri = 0.0;
    for (int b = 0; b < 4; ++b) {
      rr += Cr[(fk + a) * nc + fg + b] * Ng[b];
      ri += Ci[(fk + a) * nc + fg + b] * Ng[b];
    }
    vr += Nk[a] * rr;
    vi += Nk[a] * ri;
  }
  psi[2 * e] = vr;
  psi[2 * e + 1] = vi;
}

// [n_runs, n_parts, n_z, nP, nQ, variance, largest singular value] as doubles
__global__ void k_xi_info(const double* __restrict__ xs, double* __restrict__ out) {
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const int t = threadIdx.x;
  if (t < 5) out[t] = (double)meta[t];
  if (t == 5) out[5] = xs[L.scal];
  if (t == 6) out[6] = xs[L.scal + 1];
  if (t == 7) out[7] = xs[L.scal + 4];
}

inline size_t al256(size_t n) { return (n + 255) & ~size_t(255); }
hipStream_t xst(void* s) { return reinterpret_cast<hipStream_t>(s); }
unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// deposit workspace: J table, sorted keys / indices, runs, parts
struct XiSortWs {
  size_t jt, skey, vin, vout, ukey, rlen, nruns, rstart, nparts, poff, nmulti, hoff, tmp, total;
  size_t tmp_bytes;
};
constexpr long long kXiRunCap = (long long)kXiKeyInvalid + 1;
XiSortWs xi_sort_ws(long long n, long long n_cells) {
  XiSortWs W;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += al256(b); return r; };
  size_t sort_b = 0, rle_b = 0, scan_b = 0;
  (void)rocprim::radix_sort_pairs((void*)nullptr, sort_b, (int*)nullptr, (int*)nullptr,
                                  (int*)nullptr, (int*)nullptr, (size_t)n, 0, kXiKeyBits);
  (void)rocprim::run_length_encode((void*)nullptr, rle_b, (int*)nullptr, (size_t)n, (int*)nullptr,
                                   (int*)nullptr, (int*)nullptr);
  (void)rocprim::exclusive_scan((void*)nullptr, scan_b, (int*)nullptr, (long long*)nullptr, 0LL,
                                (size_t)kXiRunCap, rocprim::plus<long long>());
  W.tmp_bytes = std::max(sort_b, std::max(rle_b, scan_b));
  const size_t nr = (size_t)std::min<long long>(kXiRunCap, n + 1);
  W.jt = take((size_t)n_cells * kXiVals * sizeof(double));
  W.skey = take((size_t)n * sizeof(int));
  W.vin = take((size_t)n * sizeof(int));
  W.vout = take((size_t)n * sizeof(int));
  W.ukey = take(nr * sizeof(int));
  W.rlen = take(nr * sizeof(int));
  W.nruns = take(sizeof(int) * 64);
  W.rstart = take(nr * sizeof(long long));
  W.nparts = take(nr * sizeof(int));
  W.poff = take(nr * sizeof(long long));
  W.nmulti = take(nr * sizeof(int));
  W.hoff = take(nr * sizeof(long long));
  W.tmp = take(W.tmp_bytes);
  W.total = o;
  return W;
}
// rows workspace: the run keys and Hx per run (read again by the grid stage),
// the multi-part partials, first-moment parts
struct XiRowsWs {
  size_t ukey, hx, hp, lp, lc, occ, varpart, wp, wrun, phic, v0, total;
};
XiRowsWs xi_rows_ws(long long n_runs, long long n_parts) {
  XiRowsWs W;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += al256(b); return r; };
  W.ukey = take((size_t)std::max<long long>(n_runs, 1) * sizeof(int));
  W.hx = take((size_t)std::max<long long>(n_runs, 1) * kXiVals * sizeof(double));
  W.hp = take((size_t)std::max<long long>(n_parts, 1) * kXiVals * sizeof(double));
  W.lp = take((size_t)std::max<long long>(n_parts, 1) * 2 * kXiNk * sizeof(double));
  W.lc = take((size_t)(kRefMaxShells + 1) * kXiLxChunks * 2 * kXiNk * sizeof(double));
  W.occ = take((kXiKc + kXiGc) * sizeof(int));
  W.varpart = take(kXiGroups * sizeof(double));
  W.wp = take((size_t)std::max<long long>(n_parts, 1) * sizeof(double));
  W.wrun = take((size_t)std::max<long long>(n_runs, 1) * sizeof(double));
  W.phic = take((size_t)std::max<long long>(n_runs, 1) * sizeof(double));
  W.v0 = take((size_t)kRefMaxShells * kXiNk * sizeof(double));
  W.total = o;
  return W;
}
// grid workspace: boxes (as deposited and rotated), J_m and E tables, the
// two contraction outputs
struct XiGridWs {
  size_t box, rot, jq, e, T, G, qpart, tmp, total;
};
XiGridWs xi_grid_ws(int n_z, int nP, int nQ) {
  XiGridWs W;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += al256(b); return r; };
  const size_t ng = (size_t)3 * n_z;
  constexpr size_t nc = kXiCoarse;
  W.box = take(ng * nP * kXiNk * nQ * sizeof(double));
  W.rot = take(ng * nP * kXiNk * nQ * sizeof(double));
  W.jq = take((size_t)3 * nQ * nc * sizeof(double));
  W.e = take((size_t)2 * nc * nP * sizeof(double));
  W.T = take(ng * nP * kXiNk * nc * sizeof(double));
  W.G = take(ng * 2 * nc * kXiNk * nc * sizeof(double));
  W.qpart = take((size_t)kXiQChunks * nc * nc * 2 * sizeof(double));
  W.tmp = take((size_t)2 * nc * nc * sizeof(double));
  W.total = o;
  return W;
}

}  // namespace

XiRowsView xi_rows_view(const void* rows_ws, long long n_runs, long long n_parts) {
  const XiRowsWs W = xi_rows_ws(n_runs, n_parts);
  const char* b = static_cast<const char*>(rows_ws);
  return XiRowsView{(const int*)(b + W.ukey), (const double*)(b + W.phic),
                    (const double*)(b + W.wrun), (const double*)(b + W.v0)};
}
}  // namespace sgl

using namespace sgl;

namespace {
struct RefTabs {
  const double *cosmo, *derived, *slices, *rows;
};
RefTabs ref_tabs(const void* tables, int32_t flags) {
  const RefTablesOffsets o = sgl::ref_tables_offsets(flags);
  const double* T = static_cast<const double*>(tables);
  return RefTabs{T + o.cosmo, T + o.derived, T + o.slices, T + o.rows};
}
bool hip_ok() { return hipGetLastError() == hipSuccess; }
}  // namespace

extern "C" {

size_t sgl_ref_xi_state_bytes(void) { return xi_state_layout().total * sizeof(double); }

sgl_status sgl_ref_xi_init(const void* tables, int32_t flags, void* xi_state, void* stream) {
  if (!tables || !xi_state || !(flags & SGL_FLAG_CLUSTERING) || !sgl_ref_tasks(flags))
    return SGL_ERR_INVALID;
  const RefTabs T = ref_tabs(tables, flags);
  double* xs = static_cast<double*>(xi_state);
  hipStream_t st = xst(stream);
  const XiState L = xi_state_layout();
  hipLaunchKernelGGL(k_xi_zero, dim3(nblk((long long)L.total)), dim3(256), 0, st, xs,
                     (long long)L.total);
  hipLaunchKernelGGL(k_xi_init, dim3(1), dim3(256), 0, st, T.cosmo, T.derived, xs);
  hipLaunchKernelGGL(k_xi_minv, dim3(1), dim3(256), 0, st, xs);
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_xi_info(const void* xi_state, double* out, void* stream) {
  if (!xi_state || !out) return SGL_ERR_INVALID;
  hipLaunchKernelGGL(k_xi_info, dim3(1), dim3(64), 0, xst(stream),
                     static_cast<const double*>(xi_state), out);
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_xi_sort_bytes(int64_t n_points, int64_t n_cells) {
  if (n_points < 1 || n_points > ((int64_t)1 << 31) - 1 || n_cells < 1) return 0;
  return xi_sort_ws(n_points, n_cells).total;
}

sgl_status sgl_ref_xi_sort(const void* tables, int32_t flags, int64_t n_points,
                           const int32_t* xi_pts, int64_t n_cells, const double* xi_cells,
                           void* ws, void* xi_state, void* stream) {
  if (!tables || !xi_pts || !xi_cells || !ws || !xi_state || n_points < 1 ||
      n_points > ((int64_t)1 << 31) - 1 || n_cells < 1 || !(flags & SGL_FLAG_CLUSTERING))
    return SGL_ERR_INVALID;
  const XiSortWs W = xi_sort_ws(n_points, n_cells);
  char* base = static_cast<char*>(ws);
  double* xs = static_cast<double*>(xi_state);
  const XiState L = xi_state_layout();
  hipStream_t st = xst(stream);
  double* JT = (double*)(base + W.jt);
  int* skey = (int*)(base + W.skey);
  int* vin = (int*)(base + W.vin);
  int* vout = (int*)(base + W.vout);
  int* ukey = (int*)(base + W.ukey);
  int* rlen = (int*)(base + W.rlen);
  int* nruns = (int*)(base + W.nruns);
  long long* rstart = (long long*)(base + W.rstart);
  int* nparts = (int*)(base + W.nparts);
  long long* poff = (long long*)(base + W.poff);
  int* nmulti = (int*)(base + W.nmulti);
  long long* hoff = (long long*)(base + W.hoff);
  void* tmp = base + W.tmp;
  const long long n = n_points;
  const int nr = (int)std::min<long long>(kXiRunCap, n + 1);
  hipLaunchKernelGGL(k_xi_jtab, dim3(nblk(n_cells * kXiNk)), dim3(256), 0, st, (long long)n_cells,
                     xi_cells, xs, JT);
  hipLaunchKernelGGL(k_xi_iota, dim3(nblk(n)), dim3(256), 0, st, (int)n, vin);
  size_t tb = W.tmp_bytes;
  // stable: a bin's points stay in emission order
  if (rocprim::radix_sort_pairs(tmp, tb, xi_pts + n, skey, vin, vout, (size_t)n, 0, kXiKeyBits,
                                st) != hipSuccess)
    return SGL_ERR_HIP;
  tb = W.tmp_bytes;
  if (rocprim::run_length_encode(tmp, tb, skey, (size_t)n, ukey, rlen, nruns, st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_xi_nparts, dim3(nblk(nr)), dim3(256), 0, st, ukey, rlen, nruns, nr, nparts,
                     nmulti);
  // run starts: lengths past the run count are never read, so the scan
  // covers the run slots that exist; the parts and part slots likewise
  tb = W.tmp_bytes;
  if (rocprim::exclusive_scan(tmp, tb, rlen, rstart, 0LL, (size_t)nr, rocprim::plus<long long>(),
                              st) != hipSuccess)
    return SGL_ERR_HIP;
  tb = W.tmp_bytes;
  if (rocprim::exclusive_scan(tmp, tb, nparts, poff, 0LL, (size_t)nr, rocprim::plus<long long>(),
                              st) != hipSuccess)
    return SGL_ERR_HIP;
  tb = W.tmp_bytes;
  if (rocprim::exclusive_scan(tmp, tb, nmulti, hoff, 0LL, (size_t)nr, rocprim::plus<long long>(),
                              st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_xi_meta, dim3(1), dim3(1), 0, st, ukey, nruns, nparts, poff, xs);
  (void)L;
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_xi_rows_bytes(int64_t n_runs, int64_t n_parts) {
  if (n_runs < 0 || n_parts < 0) return 0;
  return xi_rows_ws(n_runs, n_parts).total;
}

sgl_status sgl_ref_xi_rows(const void* tables, int32_t flags, int64_t n_points, const double* kap,
                           const double* gam, const double* w, const int32_t* xi_pts,
                           int64_t n_cells, const double* xi_cells, const void* sort_ws,
                           int64_t n_runs, int64_t n_parts, void* rows_ws, void* xi_state,
                           void* stream) {
  if (!tables || !kap || !gam || !w || !xi_pts || !xi_cells || !sort_ws || !rows_ws ||
      !xi_state || n_points < 1 || n_points > ((int64_t)1 << 31) - 1 || n_cells < 1 ||
      n_runs < 0 || n_parts < 0 || n_runs > kXiRunCap || !(flags & SGL_FLAG_CLUSTERING))
    return SGL_ERR_INVALID;
  const RefTabs T = ref_tabs(tables, flags);
  const XiSortWs S = xi_sort_ws(n_points, n_cells);
  const XiRowsWs W = xi_rows_ws(n_runs, n_parts);
  const char* sb = static_cast<const char*>(sort_ws);
  char* rb = static_cast<char*>(rows_ws);
  double* xs = static_cast<double*>(xi_state);
  const XiState L = xi_state_layout();
  hipStream_t st = xst(stream);
  const double* JT = (const double*)(sb + S.jt);
  const int* vout = (const int*)(sb + S.vout);
  const int* ukey = (const int*)(sb + S.ukey);
  const int* rlen = (const int*)(sb + S.rlen);
  const long long* rstart = (const long long*)(sb + S.rstart);
  const int* nparts = (const int*)(sb + S.nparts);
  const long long* poff = (const long long*)(sb + S.poff);
  const long long* hoff = (const long long*)(sb + S.hoff);
  double* Hx = (double*)(rb + W.hx);
  double* Hp = (double*)(rb + W.hp);
  double* Lp = (double*)(rb + W.lp);
  double* Lc = (double*)(rb + W.lc);
  int* occ = (int*)(rb + W.occ);
  double* varpart = (double*)(rb + W.varpart);
  double* Wp = (double*)(rb + W.wp);
  double* wrun = (double*)(rb + W.wrun);
  double* phic = (double*)(rb + W.phic);
  double* v0 = (double*)(rb + W.v0);
  int* ukey_keep = (int*)(rb + W.ukey);
  const int nr = (int)n_runs;
  if (nr > 0)
    hipLaunchKernelGGL(k_xi_copy_int, dim3(nblk(nr)), dim3(256), 0, st, ukey, ukey_keep, nr);
  if (n_parts > 0)
    hipLaunchKernelGGL(k_xi_walk, dim3(nblk(n_parts * 64)), dim3(256), 0, st, rlen, nr, rstart,
                       nparts, poff, hoff, (int)n_parts, vout, kap, gam, w, xi_pts, xi_cells, JT,
                       Hx, Hp, Lp, Wp);
  if (nr > 0) {
    hipLaunchKernelGGL(k_xi_combine, dim3(nblk((long long)nr * kXiVals)), dim3(256), 0, st, nr,
                       nparts, hoff, Hp, Hx);
    hipLaunchKernelGGL(k_xi_wrun, dim3(nblk(nr)), dim3(256), 0, st, nr, nparts, poff, Wp, wrun);
  }
  hipLaunchKernelGGL(k_xi_zero, dim3(1), dim3(256), 0, st, (double*)occ,
                     (long long)((kXiKc + kXiGc) / 2 + 1));
  if (nr > 0) hipLaunchKernelGGL(k_xi_occ, dim3(nblk(nr)), dim3(256), 0, st, nr, ukey, occ);
  hipLaunchKernelGGL(k_xi_maps, dim3(1), dim3(256), 0, st, nr, ukey, occ, xs);
  hipLaunchKernelGGL(k_xi_lx_chunks, dim3(kRefMaxShells, kXiLxChunks), dim3(128), 0, st, poff,
                     nparts, xs, Lp, Lc);
  hipLaunchKernelGGL(k_xi_lx, dim3(kRefMaxShells), dim3(128), 0, st, T.cosmo, T.slices, T.rows, Lc,
                     xs);
  hipLaunchKernelGGL(k_xi_group, dim3(kXiGroups), dim3(256), 0, st, T.cosmo, T.slices, ukey, Hx,
                     xs);
  hipLaunchKernelGGL(k_xi_loading, dim3(kRefMaxShells), dim3(256), 0, st, T.cosmo, T.slices, ukey,
                     Hx, wrun, xs, phic, v0);
  hipLaunchKernelGGL(k_xi_keep, dim3(1), dim3(256), 0, st, xs);
  hipLaunchKernelGGL(k_xi_zero, dim3(nblk((long long)kRefMaxShells * kXiKc * kXiNk)), dim3(256), 0,
                     st, xs + L.W0, (long long)kRefMaxShells * kXiKc * kXiNk);
  hipLaunchKernelGGL(k_xi_tables, dim3(kXiGroups), dim3(256), 0, st, T.cosmo, T.slices, ukey, Hx,
                     xs, varpart);
  hipLaunchKernelGGL(k_xi_var, dim3(1), dim3(64), 0, st, varpart, xs);
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_xi_axis(const void* tables, int32_t flags, const void* xi_state,
                           const double* k, int32_t n, int32_t axis, double* out, void* stream) {
  if (!tables || !xi_state || !k || !out || n < 1 || (axis != 0 && axis != 1) ||
      !(flags & SGL_FLAG_CLUSTERING))
    return SGL_ERR_INVALID;
  const RefTabs T = ref_tabs(tables, flags);
  hipLaunchKernelGGL(k_xi_axis, dim3(n), dim3(256), 0, xst(stream), T.cosmo,
                     static_cast<const double*>(xi_state), k, axis, out);
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_xi_grid_bytes(int32_t n_z, int32_t nP, int32_t nQ) {
  if (n_z < 1 || n_z > kRefMaxShells || nP < 1 || nP > kXiKc || nQ < 1 || nQ > kXiGc) return 0;
  return xi_grid_ws(n_z, nP, nQ).total;
}

sgl_status sgl_ref_xi_grid(const void* tables, int32_t flags, const void* xi_state,
                           const void* rows_ws, int64_t n_runs, int32_t n_z, int32_t nP,
                           int32_t nQ, void* grid_ws, double kk_max, double kg_max, int32_t nk,
                           int32_t ng, double* psi, void* stream) {
  if (!tables || !xi_state || !rows_ws || !grid_ws || !psi || n_runs < 1 ||
      n_runs > kXiRunCap || n_z < 1 || n_z > kRefMaxShells || nP < 1 || nP > kXiKc || nQ < 1 ||
      nQ > kXiGc || nk < 2 || ng < 2 || !(kk_max > 0.0) || !(kg_max > 0.0) ||
      !(flags & SGL_FLAG_CLUSTERING))
    return SGL_ERR_INVALID;
  const XiRowsWs RW = xi_rows_ws(n_runs, 0);
  const XiGridWs W = xi_grid_ws(n_z, nP, nQ);
  const char* rb = static_cast<const char*>(rows_ws);
  const int* ukey = (const int*)(rb + RW.ukey);
  const double* Hx = (const double*)(rb + RW.hx);
  char* gb = static_cast<char*>(grid_ws);
  double* box = (double*)(gb + W.box);
  double* rot = (double*)(gb + W.rot);
  double* Jq = (double*)(gb + W.jq);
  double* E = (double*)(gb + W.e);
  double* Tm = (double*)(gb + W.T);
  double* Gm = (double*)(gb + W.G);
  double* qpart = (double*)(gb + W.qpart);
  double* tmp = (double*)(gb + W.tmp);
  double* xs = const_cast<double*>(static_cast<const double*>(xi_state));
  const XiState L = xi_state_layout();
  hipStream_t st = xst(stream);
  constexpr int nc = kXiCoarse;
  const int ngr = 3 * n_z;
  const long long per_g = (long long)nQ * nP * kXiNk;
  hipLaunchKernelGGL(k_xi_zero, dim3(nblk(ngr * per_g)), dim3(256), 0, st, box, ngr * per_g);
  hipLaunchKernelGGL(k_xi_box, dim3(nblk(n_runs * kXiVals)), dim3(256), 0, st, (int)n_runs, ukey,
                     Hx, xs, nP, nQ, box);
  // rotation by the kept eigenvectors: rot_g [(q', p')][r] = box_g [(q', p')][j] V_g [j][r]
  int rc = xi_gemm_batched(nQ * nP, kXiNk, kXiNk, box, kXiNk, per_g, false, xs + L.V, kXiNk,
                           (long)kXiNk * kXiNk, rot, kXiNk, per_g, ngr, st);
  if (rc != SGL_OK) return (sgl_status)rc;
  hipLaunchKernelGGL(k_xi_jq, dim3(nblk(3 * nQ * nc + nc * nP)), dim3(256), 0, st, xs, nP, nQ,
                     kk_max, kg_max, Jq, E);
  // gamma bins: T_g [(p', r)][c] = sum_q' rot_g [q'][(p', r)] J_m [q'][c]
  for (int m = 0; m < 3; ++m) {
    rc = xi_gemm_batched(nP * kXiNk, nc, nQ, rot + m * per_g, (long)nP * kXiNk, 3 * per_g, true,
                         Jq + (size_t)m * nQ * nc, nc, 0, Tm + (size_t)m * nP * kXiNk * nc, nc,
                         3L * nP * kXiNk * nc, n_z, st);
    if (rc != SGL_OK) return (sgl_status)rc;
  }
  // kappa bins: G_g [202][(r, c)] = [cos; sin] [202][p'] T_g [p'][(r, c)]
  rc = xi_gemm_batched(2 * nc, kXiNk * nc, nP, E, nP, 0, false, Tm, kXiNk * nc,
                       (long)nP * kXiNk * nc, Gm, kXiNk * nc, 2L * nc * kXiNk * nc, ngr, st);
  if (rc != SGL_OK) return (sgl_status)rc;
  hipLaunchKernelGGL(k_xi_qpart, dim3(nblk(nc * nc), kXiQChunks), dim3(256), 0, st, xs, Gm, ngr,
                     kk_max, kg_max, qpart);
  hipLaunchKernelGGL(k_xi_qsum, dim3(nblk(nc * nc)), dim3(256), 0, st, qpart, xs);
  hipLaunchKernelGGL(k_xi_spl1, dim3(nblk(2 * nc * nc)), dim3(256), 0, st, xs, tmp);
  hipLaunchKernelGGL(k_xi_spl2, dim3(nblk(2 * nc * nc)), dim3(256), 0, st, tmp, xs);
  hipLaunchKernelGGL(k_xi_psi, dim3(nblk((long long)nk * ng)), dim3(256), 0, st, xs, nk, ng, kk_max,
                     kg_max, psi);
  (void)tables;
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

}  // extern "C"
