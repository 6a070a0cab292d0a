// Reference-shape path: the reference's own 2-D (k_kappa, k_gamma) chain.
//
//   population  per (shell, mass) row of the reference grid (SGL_FLAG_REF_GRID
//               tables): the adaptive impact-parameter cells of _truncated_host
//               (sgl_full.py:2025-2138) with the critical-curve refinement of
//               _refine_edges / _refine_edges_eps_batch (sgl_full.py:2342-2372,
//               2279-2339), the (pseudo-elliptical) truncated NFW and the monopole
//               tail, and the filament cells (sgl_full.py:1984-2007); every cell
//               keyed by its bin of the 961 x 300 fine (kappa, gamma) lattice
//               (Cells.add / _add_direct_kernel, sgl_full.py:1314-1383, 971-1046);
//   binning     a stable radix sort of the keys, then each bin's weight and
//               centroid sums in the cells' emission order (fixed order: the
//               result does not depend on scheduling);
//   Lambda      Lambda(k_k, k_g) = sum_b w_b (e^{i k_k kap_b} - 1 - i k_k kap_b)
//               + sum_b w_b e^{i k_k kap_b} (J0(k_g gam_b) - 1) (the split of
//               Lam._sep_raw, sgl_full.py:3040-3086, with the dense product of
//               sgl_direct.direct_real, sgl_direct.py:40-51): the first sum per
//               k_k with the compensated cos - 1 and sin - x (sgl_full.py:2598-2611),
//               the second an FP64 MFMA contraction [2 n_k x N_bins] x
//               [N_bins x n_g] whose epilogue exponentiates to Phi;
//   inversion   Hankel then Fourier as two FP64 MFMA contractions
//               (sgl.P_vector_joint, sgl.py:1127-1156), the Euler-Maclaurin
//               endpoint terms (sgl_full.py:277-298) in the Fourier epilogue, and
//               the constant-mu arc integral on the bilinear interpolant
//               (invert.P_i, sgl_full.py:299-360), then the normalisation and
//               moments (sgl_full.py:362-371).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "ref_sub.h"
#include "ref_xi.h"

namespace sgl {

constexpr double kKapFloor = 1e-8;     // pipeline.run kap_floor (pipeline.py:117)
constexpr int kRefNx = 260;            // build_population nx (sgl_full.py:1536)
constexpr int kXgN = 900;              // sgl._XG = logspace(-6, 3.5, 900) (sgl.py:754)
constexpr int kBisect = 45;            // sgl._BISECT_ITERS (sgl.py:757)
constexpr int kEdgeOff = 22;           // logspace(-11, -1.2, 22) offsets (sgl_full.py:2362)
constexpr int kEllScan = 1200;         // elliptical critical-curve scan (sgl_full.py:2294)
constexpr int kEllScreen = 50;         // its screening points (sgl_full.py:2288)
constexpr int kFineKRef = 961, kFineGRef = 300;
constexpr int kRefBins = kFineKRef * kFineGRef;
constexpr int kKeyInvalid = kRefBins;  // outside the lattice: sorts after every bin
constexpr int kMaxCRoots = 4, kMaxERoots = 8;
constexpr int kRootStride = 128;       // doubles per row: [0] n, [1..4] circular roots,
                                       // [8 + 9 q] n, [9 + 9 q ...] roots of orientation q
constexpr int kMaxEdges = 1024;        // edges of one (row, orientation) after merging
constexpr int kRefFil = kFilU * kFilPhi;
constexpr int kStatusRefRoots = 16;    // more critical-curve roots than the slots
constexpr int kStatusRefEdges = 32;    // more edges than kMaxEdges

// ------------------------------------------------------------ row params --
struct RefRow {
  double ks, rs, xt, xm, Nbar, eps, x_ell, x_in, mut4;
  int n_in, nph;
  bool ok;
};

__device__ inline RefRow ref_row(const double* R, int ell) {
  RefRow p;
  p.ks = R[kRowKs];
  p.rs = R[kRowRs];
  p.xt = R[kRowXt];
  p.eps = ell ? R[6] : 0.0;
  p.nph = ell ? kEllPhi : 1;
  const double wrow = R[kRowW];
  p.ok = (p.rs > 0.0);
  const double xt = p.xt;
  const double mut = log(1.0 + xt) - xt / (1.0 + xt);
  p.mut4 = p.ks * 4.0 * mut;
  const double g_xt = p.mut4 / (xt * xt);
  p.xm = (g_xt > kKapFloor) ? xt * sqrt(g_xt / kKapFloor) : xt;
  p.Nbar = kPi * wrow * (p.xm * p.rs) * (p.xm * p.rs);
  p.x_ell = (p.eps > 0.0) ? fmin(xt / sqrt(1.0 - p.eps), p.xm) : fmin(xt, p.xm);
  p.x_in = 1e-6 * xt;
  const int nx_eff = (p.ks > 0.05) ? kRefNx : max(80, kRefNx / 3);
  p.n_in = max(nx_eff, (int)ceil(nx_eff / 6.0 * log10(p.x_ell / p.x_in)));
  return p;
}

// np.geomspace(a, b, n)[i]: 10^(log10 a + i (log10 b - log10 a)/(n-1)), endpoints exact
__device__ inline double geom(double a, double b, int n, int i) {
  if (i == 0) return a;
  if (i == n - 1) return b;
  const double la = log10(a), lb = log10(b);
  return pow(10.0, la + i * ((lb - la) / (n - 1)));
}
// np.logspace(-6, 3.5, 900)[p]
__device__ inline double xg_at(int p) {
  return pow(10.0, (p == kXgN - 1) ? 3.5 : -6.0 + p * (9.5 / (kXgN - 1)));
}
// np.logspace(-11, -1.2, 22)[q]
__device__ inline double edge_off(int q) {
  return pow(10.0, (q == kEdgeOff - 1) ? -1.2 : -11.0 + q * (9.8 / (kEdgeOff - 1)));
}

__device__ inline int sgn(double d) { return (d > 0.0) ? 1 : ((d < 0.0) ? -1 : 0); }

// untruncated NFW detA sign (sgl.kappa_gamma, sgl.py:722-740)
__device__ inline int det_sign_circ(double x, double ks) {
  double F, h;
  nfw_fh(x, F, h);
  const double k = 2.0 * ks * F, g = 4.0 * ks * h / (x * x) - k;
  return sgn((1.0 - k) * (1.0 - k) - g * g);
}

// pseudo-elliptical truncated NFW at (x, phi) (kappagamma_eps, sgl_full.py:582-613)
__device__ inline void kg_eps_ref(double x, double cph, double sph, double eps, double ks,
                                  double xt, double mut4, double& kap, double& gam) {
  const double x1 = sqrt(1.0 - eps) * cph * x, x2 = sqrt(1.0 + eps) * sph * x;
  const double xe = fmax(sqrt(x1 * x1 + x2 * x2), 1e-12);
  const double c2 = (x1 * x1 - x2 * x2) / fmax(xe * xe, 1e-300);
  double k0, g0;
  trunc_kg(xe, ks, xt, k0, g0);
  if (xe < 1e-4) g0 = ks;
  double k = k0 + eps * c2 * g0;
  double g2 = g0 * g0 + 2.0 * eps * c2 * g0 * k0 + eps * eps * (k0 * k0 - (c2 * g0) * (c2 * g0));
  if (xe >= xt) {
    const double xx = fmax(x, 1e-300);
    const double gpm = mut4 / (xx * xx);
    k = 0.0;
    g2 = gpm * gpm;
  }
  kap = k;
  gam = sqrt(fmax(g2, 0.0));
}

__device__ inline void ell_phi(int q, double& c, double& s) {
  const double ph = (q + 0.5) * (0.5 * kPi / kEllPhi);
  sincos(ph, &s, &c);
}

// ------------------------------------------------------- roots (per row) --
// One workgroup per (shell, mass) row: the circular critical radii (sign
// changes of detA of the untruncated profile on sgl._XG, 45 bisections each)
// and, for elliptical halos, the 50-point screen and the per-orientation
// roots on 1200 log points.
__global__ __launch_bounds__(256) void k_ref_roots(const double* __restrict__ cosmo,
                                                   const double* __restrict__ rows, int n_z_max,
                                                   int ell, double* __restrict__ roots,
                                                   int* __restrict__ status) {
  __shared__ signed char sg[kEllPhi * kEllScan];
  __shared__ double ra[kEllPhi * kMaxERoots], rb[kEllPhi * kMaxERoots];
  __shared__ signed char r0[kEllPhi * kMaxERoots];
  __shared__ int cnt[kEllPhi];
  __shared__ int fail;
  const int j = blockIdx.x, i = blockIdx.y, t = threadIdx.x;
  const int row = i * kRefMasses + j;
  double* out = roots + (size_t)row * kRootStride;
  for (int q = t; q < kRootStride; q += blockDim.x) out[q] = 0.0;
  const Cosmo c = reinterpret_cast<const Cosmo*>(cosmo)[0];
  if (i >= (int)c.n_z) return;
  const RefRow p = ref_row(rows + (size_t)row * kRowStride, ell);
  if (!p.ok || p.ks < 0.05) return;
  // circular roots (sgl_full.py:2349-2372)
  for (int q = t; q < kXgN; q += blockDim.x) sg[q] = (signed char)det_sign_circ(xg_at(q), p.ks);
  __syncthreads();
  if (t == 0) {
    int n = 0;
    for (int q = 0; q + 1 < kXgN; ++q)
      if (sg[q] * sg[q + 1] < 0) {
        if (n < kMaxCRoots) {
          ra[n] = xg_at(q);
          rb[n] = xg_at(q + 1);
          r0[n] = sg[q];
        }
        ++n;
      }
    if (n > kMaxCRoots) atomicOr(status, kStatusRefRoots);
    cnt[0] = min(n, kMaxCRoots);
  }
  __syncthreads();
  const int nc = cnt[0];
  if (t < nc) {
    double a = ra[t], b = rb[t];
    const int s0 = r0[t];
    for (int it = 0; it < kBisect; ++it) {
      const double m = 0.5 * (a + b);
      if (det_sign_circ(m, p.ks) == s0) a = m; else b = m;
    }
    out[1 + t] = 0.5 * (a + b);
  }
  if (t == 0) out[0] = nc;
  __syncthreads();
  if (!ell) return;
  // elliptical screen: detA > 0.02 at 50 points for phi = 0 and pi/2 means no root
  const double xa = fmax(p.x_in, 1e-7);
  if (t == 0) fail = 0;
  __syncthreads();
  if (t < 2 * kEllScreen) {
    const int q = t % kEllScreen, side = t / kEllScreen;
    double k, g, sph, cph;
    sincos(side ? 0.5 * kPi : 0.0, &sph, &cph);
    kg_eps_ref(geom(xa, p.x_ell, kEllScreen, q), cph, sph, p.eps, p.ks, p.xt, p.mut4, k, g);
    if (!((1.0 - k) * (1.0 - k) - g * g > 0.02)) atomicOr(&fail, 1);
  }
  __syncthreads();
  if (!fail) return;
  for (int idx = t; idx < kEllPhi * kEllScan; idx += blockDim.x) {
    const int q = idx / kEllScan, m = idx - q * kEllScan;
    double cph, sph, k, g;
    ell_phi(q, cph, sph);
    kg_eps_ref(geom(xa, p.x_ell, kEllScan, m), cph, sph, p.eps, p.ks, p.xt, p.mut4, k, g);
    sg[idx] = (signed char)sgn((1.0 - k) * (1.0 - k) - g * g);
  }
  __syncthreads();
  if (t < kEllPhi) {
    int n = 0;
    const signed char* S = sg + t * kEllScan;
    for (int m = 0; m + 1 < kEllScan; ++m)
      if (S[m] * S[m + 1] < 0) {
        if (n < kMaxERoots) {
          ra[t * kMaxERoots + n] = geom(xa, p.x_ell, kEllScan, m);
          rb[t * kMaxERoots + n] = geom(xa, p.x_ell, kEllScan, m + 1);
          r0[t * kMaxERoots + n] = S[m];
        }
        ++n;
      }
    if (n > kMaxERoots) atomicOr(status, kStatusRefRoots);
    cnt[t] = min(n, kMaxERoots);
  }
  __syncthreads();
  if (t < kEllPhi * kMaxERoots) {
    const int q = t / kMaxERoots, r = t % kMaxERoots;
    if (r < cnt[q]) {
      double cph, sph;
      ell_phi(q, cph, sph);
      double a = ra[t], b = rb[t];
      const int s0 = r0[t];
      for (int it = 0; it < kBisect; ++it) {
        const double m = 0.5 * (a + b);
        double k, g;
        kg_eps_ref(m, cph, sph, p.eps, p.ks, p.xt, p.mut4, k, g);
        if (sgn((1.0 - k) * (1.0 - k) - g * g) == s0) a = m; else b = m;
      }
      out[9 + 9 * q + r] = 0.5 * (a + b);
    }
    if (r == 0) out[8 + 9 * q] = cnt[q];
  }
}

// --------------------------------------------------------- block helpers --
// Exclusive scan of one int per thread over a 256-thread block (Hillis-Steele).
__device__ inline int block_excl_scan(int v, int* tmp, int& total) {
  const int t = threadIdx.x;
  tmp[t] = v;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int a = (t >= off) ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += a;
    __syncthreads();
  }
  total = tmp[255];
  const int ex = tmp[t] - v;
  __syncthreads();
  return ex;
}

// searchsorted(e, v, side="right") over the n sorted edges of the fine
// lattice, with a value on the last edge moved into the last bin
// (np.histogramdd, sgl_full.py:888-934): bin index 1..n-1 inside, 0 or n
// outside (NaN: outside).  The index is guessed from the edges' logarithmic
// spacing and then walked to the first edge above v: one or two loads
// instead of a ten-step bisection, the same result.
__device__ inline int walk_right(const double* __restrict__ e, int n, double v, int g) {
  g = min(max(g, 0), n);
  while (g > 0 && e[g - 1] > v) --g;
  while (g < n && e[g] <= v) ++g;
  if (g == n && v == e[n - 1]) --g;
  return g;
}
__device__ inline int bin_right_k(const double* __restrict__ ek, double v) {
  constexpr int n = kFineKRef + 1, npos = n / 2;        // 481 edges on each side
  if (!(v == v)) return n;
  const double av = fabs(v);
  int g = npos;
  if (av >= 1e-7) {
    const int m = (int)(log(av * 1e7) * ((npos - 1) / log(30.0 * 1e7)));
    g = (v > 0.0) ? npos + m + 1 : npos - 1 - m;
  }
  return walk_right(ek, n, v, g);
}
__device__ inline int bin_right_g(const double* __restrict__ eg, double v) {
  constexpr int n = kFineGRef + 1;
  if (!(v == v)) return n;
  const int g = (v >= 1e-9) ? (int)(log(v * 1e9) * ((n - 1) / log(30.0 * 1e9))) + 1 : 0;
  return walk_right(eg, n, v, g);
}

// The same search on the coarse lattice of the xi_lin term: 342 signed-log
// kappa edges, 111 log gamma edges (sgl_full.py:457, 1081-1130), and the key
// shell * kXiBins + bin of a point inside it.
__device__ inline int coarse_bin_k(const double* __restrict__ ekc, double v) {
  constexpr int n = kXiKc + 1, npos = n / 2;
  if (!(v == v)) return n;
  const double av = fabs(v);
  int g = npos;
  if (av >= 1e-7) {
    const int m = (int)(log(av * 1e7) * ((npos - 1) / log(30.0 * 1e7)));
    g = (v > 0.0) ? npos + m + 1 : npos - 1 - m;
  }
  return walk_right(ekc, n, v, g);
}
__device__ inline int coarse_bin_g(const double* __restrict__ egc, double v) {
  constexpr int n = kXiGc + 1;
  if (!(v == v)) return n;
  const int g = (v >= 1e-9) ? (int)(log(v * 1e9) * ((n - 1) / log(30.0 * 1e9))) + 1 : 0;
  return walk_right(egc, n, v, g);
}
__device__ inline int xi_ckey(const double* __restrict__ ekc, const double* __restrict__ egc,
                              int shell, double kap, double g) {
  const int a = coarse_bin_k(ekc, kap), b = coarse_bin_g(egc, g);
  return (a >= 1 && a <= kXiKc && b >= 1 && b <= kXiGc)
             ? shell * kXiBins + (a - 1) * kXiGc + (b - 1)
             : kXiKeyInvalid;
}

// ------------------------------------------------ cells (count or emit) --
// One workgroup per (row, orientation) task.  The row's edge list is merged
// in LDS: [0, geomspace(x_in, x_ell, n_in)], the three truncation edges kept
// below x_ell, the 44 offsets of every circular root and of this
// orientation's roots kept strictly between x_in and x_ell; then sorted
// (bitonic) and made unique, as np.unique does.  Count mode writes the task's
// cell count (the last orientation also carries the monopole tail and the
// filament cells, the reference's emission order); emit mode writes every
// cell's (kappa, gamma, weight) and fine-lattice key at the task's offset.
// Modes of the cell kernel: count the cells of every (row, orientation) task,
// emit them, or (subhalos) find per host row the largest radius of its
// circular grid and monopole tail within reach of a clump.
// With the xi_lin term (xe.ccounts / xe.pcid set) the count pass also counts
// the task's lens radii and the emit pass writes each point's radius id and
// coarse key: an orientation without roots of its own has the circular edge
// list, so its inner cells share the radii of the first such orientation
// (the reference groups a row's points by equal R, sgl_full.py:1397-1441).
constexpr int kModeCount = 0, kModeEmit = 1, kModeXref = 2;

template <int MODE>
__global__ __launch_bounds__(256, 2) void k_ref_cells(const double* __restrict__ cosmo,
                                                   const double* __restrict__ rows,
                                                   const double* __restrict__ frows,
                                                   const double* __restrict__ roots,
                                                   const double* __restrict__ ek,
                                                   const double* __restrict__ eg, int ell,
                                                   int fil, long long* __restrict__ counts,
                                                   const long long* __restrict__ offsets,
                                                   double* __restrict__ K, double* __restrict__ G,
                                                   double* __restrict__ W, int* __restrict__ key,
                                                   int* __restrict__ status,
                                                   double* __restrict__ subh,
                                                   const double* __restrict__ subc, XiEmit xe) {
  __shared__ double v[kMaxEdges];
  __shared__ double e[kMaxEdges];
  __shared__ int tmp[256];
  __shared__ int ndec_s;
  const int nph = (ell && MODE != kModeXref) ? kEllPhi : 1;
  const int task = blockIdx.x, t = threadIdx.x;
  const int row = task / nph, q = task - row * nph;
  const int i = row / kRefMasses;
  const Cosmo c = reinterpret_cast<const Cosmo*>(cosmo)[0];
  if (i >= (int)c.n_z) {
    if (MODE == kModeCount && t == 0) {
      counts[task] = 0;
      if (xe.ccounts) xe.ccounts[task] = 0;
    }
    return;
  }
  double* Hs = subh ? subh + (size_t)row * kHostStride : nullptr;
  const bool dec = Hs && Hs[kHrValid] != 0.0 && (MODE == kModeXref || Hs[kHrNq] > 0.0);
  if (MODE == kModeXref && !dec) return;
  const RefRow p = ref_row(rows + (size_t)row * kRowStride, ell);
  const double* rt = roots + (size_t)row * kRootStride;
  const int nc = (int)rt[0];
  const int ne = (ell && MODE != kModeXref) ? (int)rt[8 + 9 * q] : 0;
  // candidate slots: base n_in + 1, truncation 3, offsets 44 per root
  const int nb = p.n_in + 1, n_off = 2 * kEdgeOff;
  const int ncand = nb + 3 + n_off * (nc + ne);
  if (ncand > kMaxEdges) {
    if (t == 0) {
      atomicOr(status, kStatusRefEdges);
      if (MODE == kModeCount) {
        counts[task] = 0;
        if (xe.ccounts) xe.ccounts[task] = 0;
      }
    }
    return;
  }
  const double inf = __builtin_huge_val();
  for (int s = t; s < kMaxEdges; s += 256) {
    double x = inf;
    if (s < nb) {
      x = (s == 0) ? 0.0 : geom(p.x_in, p.x_ell, p.n_in, s - 1);
    } else if (s < nb + 3) {
      const int u = s - nb;
      const double xt = p.xt;
      const double y = (u == 0) ? xt * (1 - 1e-9) : (u == 1) ? xt : xt * (1 + 1e-9);
      x = (y <= p.x_ell) ? y : inf;
    } else if (s < ncand) {
      const int u = s - nb - 3, r = u / n_off, o = u - r * n_off;
      const double xc = (r < nc) ? rt[1 + r] : rt[9 + 9 * q + (r - nc)];
      const double off = edge_off(o % kEdgeOff);
      const double y = (o < kEdgeOff) ? xc * (1.0 + off) : xc * (1.0 - off);
      x = (y > p.x_in && y < p.x_ell) ? y : inf;
    }
    v[s] = x;
  }
  __syncthreads();
  int P = 1;
  while (P < ncand) P <<= 1;
  // bitonic sort of v[0..P)
  for (int k = 2; k <= P; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int s = t; s < P; s += 256) {
        const int l = s ^ jj;
        if (l > s) {
          const double a = v[s], b = v[l];
          const bool up = (s & k) == 0;
          if ((a > b) == up) { v[s] = b; v[l] = a; }
        }
      }
      __syncthreads();
    }
  // unique: four consecutive slots per thread
  int mine = 0;
  for (int u = 0; u < 4; ++u) {
    const int s = t * 4 + u;
    if (s < P && v[s] < inf && (s == 0 || v[s] != v[s - 1])) ++mine;
  }
  int total;
  int pos = block_excl_scan(mine, tmp, total);
  for (int u = 0; u < 4; ++u) {
    const int s = t * 4 + u;
    if (s < P && v[s] < inf && (s == 0 || v[s] != v[s - 1])) e[pos++] = v[s];
  }
  __syncthreads();
  const int n_e = total;
  const int n_in_cells = n_e - 1;
  const bool last = (q == nph - 1);
  const bool tail = last && (p.xm > p.x_ell * (1.0 + 1e-12));
  const int n_out = tail ? max(8, (int)ceil(12.0 * log10(p.xm / p.x_ell))) : 0;
  const int n_fil = (last && fil) ? kRefFil : 0;
  auto centre = [&](int ci) {
    const double xlo = e[ci], xhi = e[ci + 1];
    return (xlo > 0.0) ? sqrt(xlo * xhi) : 0.5 * xhi;
  };
  auto tail_centre = [&](int k) {
    return sqrt(geom(p.x_ell, p.xm, n_out + 1, k) * geom(p.x_ell, p.xm, n_out + 1, k + 1));
  };
  if (MODE == kModeXref) {
    // xref of host_cells: the circular centres and the tail centres within
    // x_sub; the coarse grid of the host's moment tables spans up to the
    // largest of them (subhalos_exact.py:1658, sgl_full.py:2096-2098)
    if (t == 0) {
      const double xs = Hs[kHrXsub];
      double xr = 0.0;
      for (int k = n_out - 1; k >= 0 && xr == 0.0; --k)
        if (tail_centre(k) <= xs) xr = tail_centre(k);
      for (int ci = n_in_cells - 1; ci >= 0 && xr == 0.0; --ci)
        if (centre(ci) <= xs) xr = centre(ci);
      Hs[kHrXref] = xr;
    }
    return;
  }
  // decorated cells: the inner cells and tail cells with centre <= x_sub
  // become nq points each (their centres increase, so they lead each part)
  int nq = 1, ndi = 0, ndo = 0;
  if (dec) {
    nq = (int)Hs[kHrNq];
    const double xs = Hs[kHrXsub];
    if (t == 0) ndec_s = 0;
    __syncthreads();
    int cnt = 0;
    for (int ci = t; ci < n_in_cells; ci += 256) cnt += (centre(ci) <= xs) ? 1 : 0;
    if (cnt) atomicAdd(&ndec_s, cnt);
    __syncthreads();
    ndi = ndec_s;
    for (int k = 0; k < n_out; ++k) ndo += (tail_centre(k) <= xs) ? 1 : 0;
  }
  const long long n_seg0 = (long long)ndi * nq;                 // inner decorated
  const long long n_seg1 = n_in_cells - ndi;                   // inner plain
  const long long n_seg2 = (long long)ndo * nq;                // tail decorated
  const long long n_seg3 = n_out - ndo;                        // tail plain
  const long long n_all = n_seg0 + n_seg1 + n_seg2 + n_seg3 + n_fil;
  // xi_lin radii: the first orientation without roots of its own carries the
  // circular inner cells' radii, an orientation with roots its own
  int q_can = 0;
  if (ell) {
    q_can = -1;
    for (int qq = 0; qq < nph; ++qq)
      if ((int)rt[8 + 9 * qq] == 0) { q_can = qq; break; }
  }
  const bool own_inner = (ne > 0) || (q == q_can);
  const long long n_own = own_inner ? n_in_cells : 0;
  if (MODE == kModeCount) {
    if (t == 0) {
      counts[task] = n_all;
      if (xe.ccounts) xe.ccounts[task] = n_own + n_out + (n_fil ? kFilU : 0);
      if (dec) {
        // the decorated cells of this task, for the point expansion
        int* ts = reinterpret_cast<int*>(subh + (size_t)kSubRows * kHostStride) + (size_t)task * 4;
        ts[0] = ndi;
        ts[1] = ndo;
        ts[2] = n_in_cells;
        ts[3] = nq;
      }
    }
    return;
  }
  const long long base = offsets[task];
  const double iw = 1.0 / (p.xm * p.xm);
  double cph = 1.0, sph = 0.0;
  if (ell) ell_phi(q, cph, sph);
  // work items: decorated inner cells, plain inner cells, decorated tail
  // cells, plain tail cells, filament cells.  A decorated cell leaves its
  // centre and weight in the first of its nq output slots; k_sub_emit
  // expands it into the points (decorated cells lead their parts, so the
  // output position of item r follows from the segment sizes).
  const long long m_all = ndi + n_seg1 + ndo + n_seg3 + n_fil;
  const bool xi = xe.pcid != nullptr;
  long long cinner = 0, ctail = 0;
  double a_r = 0.0, zp1 = 1.0, beta_h = 0.0;
  if (xi) {
    const long long cbase = xe.coffsets[task];
    cinner = own_inner ? cbase : xe.coffsets[(long long)row * nph + q_can];
    ctail = cbase + n_own;
    zp1 = 1.0 + xe.slices[(size_t)i * kSliceStride + kSlZ];
    a_r = p.rs * zp1;
    beta_h = rows[(size_t)row * kRowStride + kRowBeta];
  }
  for (long long it = t; it < m_all; it += 256) {
    double kap, gam, w;
    long long r = it, o;
    bool stash = false;
    int cid = 0;
    if (r < ndi + n_seg1) {
      const int ci = (int)r;
      const double xlo = e[ci], xhi = e[ci + 1];
      const double xc = (xlo > 0.0) ? sqrt(xlo * xhi) : 0.5 * xhi;
      w = p.Nbar * (xhi * xhi - xlo * xlo) * iw;
      if (ell) w = w / kEllPhi;
      if (xi) {
        cid = (int)(cinner + ci);
        if (own_inner) {
          xe.cells[2 * (size_t)cid] = xc * a_r;
          xe.cells[2 * (size_t)cid + 1] = beta_h;
        }
      }
      if (r < ndi) {
        o = base + r * nq;
        kap = xc;
        stash = true;
      } else {
        o = base + n_seg0 + (r - ndi);
        if (ell) kg_eps_ref(xc, cph, sph, p.eps, p.ks, p.xt, p.mut4, kap, gam);
        else trunc_kg(xc, p.ks, p.xt, kap, gam);
      }
    } else if ((r -= ndi + n_seg1) < ndo + n_seg3) {
      const int k = (int)r;
      const double e0 = geom(p.x_ell, p.xm, n_out + 1, k), e1 = geom(p.x_ell, p.xm, n_out + 1, k + 1);
      const double xo = sqrt(e0 * e1);
      w = p.Nbar * (e1 * e1 - e0 * e0) * iw;
      if (xi) {
        cid = (int)(ctail + k);
        xe.cells[2 * (size_t)cid] = xo * a_r;
        xe.cells[2 * (size_t)cid + 1] = beta_h;
      }
      const long long b2 = base + n_seg0 + n_seg1;
      if (r < ndo) {
        o = b2 + r * nq;
        kap = xo;
        stash = true;
      } else {
        o = b2 + n_seg2 + (r - ndo);
        kap = 0.0;
        gam = p.mut4 / (xo * xo);
      }
    } else {
      // filament cell (u bin l, orientation n), C order of the (120, 13) arrays
      const int f = (int)(r - ndo - n_seg3), l = f / kFilPhi, n = f - l * kFilPhi;
      o = base + n_seg0 + n_seg1 + n_seg2 + n_seg3 + f;
      const double* FR = frows + (size_t)row * kFilRowStride;
      double uc, du2, kf, wphi;
      fil_ubin(l, uc, du2);
      fil_orientation(n, kf, wphi);
      const double k0 = FR[0] * kf / FR[1];
      const double sq = sqrt(1.0 + uc * uc);
      kap = (uc <= 1.0) ? k0 * 2.0 * kPi / sq : 0.0;
      gam = k0 * 4.0 * kPi / (sq + 1.0) - kap;
      w = kPi * FR[2] * FR[1] * FR[1] * du2 * wphi;
      if (xi) {
        // the filament axis is not tangential: R < 0 marks m = 0 only
        cid = (int)(ctail + n_out + l);
        if (n == 0) {
          xe.cells[2 * (size_t)cid] = -(uc * (FR[1] * zp1));
          xe.cells[2 * (size_t)cid + 1] = FR[3];
        }
      }
    }
    if (stash) {
      K[o] = kap;
      W[o] = w;
      if (xi) xe.pcid[o] = cid;
      continue;
    }
    const double gc = fmax(gam, 1.0000001 * 1e-9);   // np.clip(G, 1.0000001 * 1e-9, None)
    const int a = bin_right_k(ek, kap);
    const int b = bin_right_g(eg, (gam == gam) ? gc : gam);
    const bool in = (a >= 1 && a <= kFineKRef && b >= 1 && b <= kFineGRef);
    K[o] = kap;
    G[o] = (gam == gam) ? gc : gam;
    W[o] = w;
    key[o] = in ? (a - 1) * kFineGRef + (b - 1) : kKeyInvalid;
    if (xi) {
      xe.pcid[o] = cid;
      xe.pck[o] = xi_ckey(xe.ekc, xe.egc, i, kap, (gam == gam) ? gc : gam);
    }
  }
}

// The points of the decorated cells (after k_ref_cells' emit pass): one
// workgroup per task; the stashed (x, w) of up to 256 cells go to LDS, then
// the cells' nq points each are written over their slots.
__global__ __launch_bounds__(256, 2) void k_sub_emit(const double* __restrict__ cosmo,
                                                  const double* __restrict__ ek,
                                                  const double* __restrict__ eg, int ell,
                                                  const long long* __restrict__ offsets,
                                                  const double* __restrict__ subh,
                                                  const double* __restrict__ subc,
                                                  double* __restrict__ K, double* __restrict__ G,
                                                  double* __restrict__ W, int* __restrict__ key,
                                                  XiEmit xe) {
  __shared__ double sx[256], sw[256];
  __shared__ int scid[256];
  const int nph = ell ? kEllPhi : 1;
  const int task = blockIdx.x, t = threadIdx.x;
  const int row = task / nph, q = task - row * nph;
  const int i = row / kRefMasses;
  const Cosmo c = reinterpret_cast<const Cosmo*>(cosmo)[0];
  if (i >= (int)c.n_z) return;
  const double* Hs = subh + (size_t)row * kHostStride;
  if (Hs[kHrValid] == 0.0 || !(Hs[kHrNq] > 0.0)) return;
  const int* ts = reinterpret_cast<const int*>(subh + (size_t)kSubRows * kHostStride) + (size_t)task * 4;
  const int ndi = ts[0], ndo = ts[1], n_in = ts[2], nq = ts[3];
  const double* Cm = subc + (size_t)row * kCoarseStride;
  double cph = 1.0, sph = 0.0;
  if (ell) ell_phi(q, cph, sph);
  const long long base = offsets[task];
  for (int seg = 0; seg < 2; ++seg) {
    const int ncell = seg ? ndo : ndi;
    const long long b0 = seg ? base + (long long)ndi * nq + (n_in - ndi) : base;
    const bool circ = seg ? true : !ell;
    for (int c0 = 0; c0 < ncell; c0 += 256) {
      const int nc = min(256, ncell - c0);
      __syncthreads();
      if (t < nc) {
        const long long o = b0 + (long long)(c0 + t) * nq;
        sx[t] = K[o];
        sw[t] = W[o];
        if (xe.pcid) scid[t] = xe.pcid[o];
      }
      __syncthreads();
      for (int pq = t; pq < nc * nq; pq += 256) {
        const int cl = pq / nq, qq = pq - cl * nq;
        double kap, gam, wq;
        sub_point(Hs, Cm, sx[cl], circ ? 1.0 : cph, circ ? 0.0 : sph, circ, qq, kap, gam, wq);
        const double gc = fmax(gam, 1.0000001 * 1e-9);
        const int a = bin_right_k(ek, kap);
        const int b = bin_right_g(eg, (gam == gam) ? gc : gam);
        const bool in = (a >= 1 && a <= kFineKRef && b >= 1 && b <= kFineGRef);
        const long long o = b0 + (long long)(c0 + cl) * nq + qq;
        K[o] = kap;
        G[o] = (gam == gam) ? gc : gam;
        W[o] = sw[cl] * wq;
        key[o] = in ? (a - 1) * kFineGRef + (b - 1) : kKeyInvalid;
        if (xe.pcid) {
          xe.pcid[o] = scid[cl];
          xe.pck[o] = xi_ckey(xe.ekc, xe.egc, i, kap, (gam == gam) ? gc : gam);
        }
      }
    }
  }
}

// --------------------------------------------------------- scans / sums --
// In-place exclusive scan of n int64 counts with the total at [n]: one
// workgroup, each thread scanning a contiguous chunk.
__global__ __launch_bounds__(256) void k_excl_scan_i64(long long* __restrict__ a, int n) {
  __shared__ long long tmp[256];
  const int t = threadIdx.x;
  const int per = (n + 255) / 256;
  const int lo = min(n, t * per), hi = min(n, lo + per);
  long long s = 0;
  for (int q = lo; q < hi; ++q) s += a[q];
  tmp[t] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const long long v = (t >= off) ? tmp[t - off] : 0;
    __syncthreads();
    tmp[t] += v;
    __syncthreads();
  }
  long long run = tmp[t] - s;
  for (int q = lo; q < hi; ++q) {
    const long long c = a[q];
    a[q] = run;
    run += c;
  }
  if (t == 255) a[n] = tmp[255];
}

// run-length slots past the run count hold no run: zero them before the scan
__global__ void k_zero_tail_runs(int* __restrict__ rlen, const int* __restrict__ nruns) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i <= kRefBins && i >= *nruns) rlen[i] = 0;
}

// a cell's (kappa, gamma, w), the payload the key sort carries
struct KGW {
  double k, g, w;
};
__global__ void k_pack_kgw(long long n, const double* __restrict__ K, const double* __restrict__ G,
                           const double* __restrict__ W, KGW* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = KGW{K[i], G[i], W[i]};
}

// Block sums of x[i] * y[i] (y may be null) over contiguous chunks, fixed order.
__global__ __launch_bounds__(256) void k_chunk_sums(const double* __restrict__ x,
                                                    const double* __restrict__ y, long long n,
                                                    long long chunk, double* __restrict__ part) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  const long long lo = (long long)blockIdx.x * chunk, hi = min(n, lo + chunk);
  double s = 0.0;
  for (long long q = lo + t; q < hi; q += 256) s += y ? x[q] * y[q] : x[q];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int n,
                                                   double* __restrict__ out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int q = t; q < n; q += 256) s += part[q];
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) *out = red[0];
}

// Runs of equal keys (the sorted cells of one fine bin, in emission order)
// are summed in parts of kRunPart cells, one wave per part (lane-strided
// partial sums, then a fixed shuffle tree), and the parts of a run are added
// in order: a fixed summation order, whatever the run lengths.
constexpr int kRunPart = 4096;

__global__ void k_run_nparts(const int* __restrict__ rlen, const int* __restrict__ nruns,
                             int* __restrict__ nparts) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > kRefBins) return;
  nparts[r] = (r < *nruns) ? (rlen[r] + kRunPart - 1) / kRunPart : 0;
}

__global__ __launch_bounds__(256) void k_run_parts(const int* __restrict__ rlen,
                                                   const int* __restrict__ nruns,
                                                   const long long* __restrict__ rstart,
                                                   const int* __restrict__ nparts,
                                                   const long long* __restrict__ poff,
                                                   const KGW* __restrict__ v,
                                                   double* __restrict__ psum) {
  const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int nr = *nruns;
  if (nr < 1) return;
  const long long total = poff[nr - 1] + nparts[nr - 1];
  if (wv >= total) return;
  int lo = 0, hi = nr - 1;                  // last run with poff <= wv
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (poff[mid] <= wv) lo = mid; else hi = mid - 1;
  }
  const int r = lo;
  const long long p = wv - poff[r];
  const long long a = rstart[r] + p * kRunPart;
  const long long b = min(rstart[r] + (long long)rlen[r], a + kRunPart);
  double sw = 0.0, sk = 0.0, sgm = 0.0;
  for (long long q = a + lane; q < b; q += 64) {
    const KGW c = v[q];
    sw += c.w;
    sk += c.w * c.k;
    sgm += c.w * c.g;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sw += __shfl_down(sw, off);
    sk += __shfl_down(sk, off);
    sgm += __shfl_down(sgm, off);
  }
  if (lane == 0) {
    double* o = psum + (size_t)wv * 3;
    o[0] = sw;
    o[1] = sk;
    o[2] = sgm;
  }
}

__global__ void k_run_combine(const int* __restrict__ ukeys, const int* __restrict__ nruns,
                              const int* __restrict__ nparts, const long long* __restrict__ poff,
                              const double* __restrict__ psum, double* __restrict__ rsum) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= *nruns) return;
  double sw = 0.0, sk = 0.0, sgm = 0.0;
  for (long long q = poff[r]; q < poff[r] + nparts[r]; ++q) {
    sw += psum[q * 3];
    sk += psum[q * 3 + 1];
    sgm += psum[q * 3 + 2];
  }
  double* o = rsum + (size_t)r * 4;
  o[0] = sw;
  o[1] = sk;
  o[2] = sgm;
  o[3] = (ukeys[r] != kKeyInvalid && sw != 0.0) ? 1.0 : 0.0;
}

// Exclusive scan of the runs' keep flags (one workgroup) and the compacted
// bin table (w, kappa centroid, gamma centroid), in bin order.
__global__ __launch_bounds__(256) void k_compact_bins(const double* __restrict__ rsum,
                                                      const int* __restrict__ nruns,
                                                      double* __restrict__ pw,
                                                      double* __restrict__ pk,
                                                      double* __restrict__ pg,
                                                      int* __restrict__ npop) {
  __shared__ int tmp[256];
  const int t = threadIdx.x, n = *nruns;
  const int per = (n + 255) / 256;
  const int lo = min(n, t * per), hi = min(n, lo + per);
  int c = 0;
  for (int q = lo; q < hi; ++q) c += rsum[(size_t)q * 4 + 3] != 0.0;
  int total;
  int pos = block_excl_scan(c, tmp, total);
  for (int q = lo; q < hi; ++q) {
    const double* r = rsum + (size_t)q * 4;
    if (r[3] == 0.0) continue;
    pw[pos] = r[0];
    pk[pos] = r[1] / r[0];
    pg[pos] = r[2] / r[0];
    ++pos;
  }
  if (t == 0) *npop = total;
}

// ------------------------------------------------------------- Lambda ----
// J0(x) - 1 without cancellation (sgl_full.py:2576-2582)
__device__ inline double j0m1(double x) {
  if (x < 1e-3) return -x * x / 4.0 * (1.0 - x * x / 16.0);
  return j0(x) - 1.0;
}

// e^{ix} - 1 - ix, compensated (sgl_full.py:2598-2611); also sin x, cos x
__device__ inline void cexpm1mx(double ph, double& cm1, double& smp, double& sn, double& cs) {
  sincos(ph, &sn, &cs);
  const double hs = sin(0.5 * ph);
  cm1 = -2.0 * hs * hs;
  if (fabs(ph) < 1e-2) {
    const double p2 = ph * ph;
    smp = -ph * p2 / 6.0 * (1.0 - p2 / 20.0 * (1.0 - p2 / 42.0));
  } else {
    smp = sn - ph;
  }
}

// Lambda on an axis for the k-reach probes (_reach, sgl_full.py:38-47): axis 0
// Lambda(k, 0) = sum w (e^{ik kap} - 1 - ik kap); axis 1 Lambda(0, k) = sum w (J0(k gam) - 1).
// One workgroup per probe, fixed-order sums.
__global__ __launch_bounds__(256) void k_ref_lam_axis(const double* __restrict__ pw,
                                                      const double* __restrict__ pk,
                                                      const double* __restrict__ pg,
                                                      const int* __restrict__ npop,
                                                      const double* __restrict__ kv, int axis,
                                                      double* __restrict__ out) {
  __shared__ double rr[256], ri[256];
  const int n = *npop, t = threadIdx.x;
  const double k = kv[blockIdx.x];
  double sr = 0.0, si = 0.0;
  for (int c = t; c < n; c += 256) {
    if (axis == 0) {
      double cm1, smp, sn, cs;
      cexpm1mx(k * pk[c], cm1, smp, sn, cs);
      sr += pw[c] * cm1;
      si += pw[c] * smp;
    } else {
      sr += pw[c] * j0m1(k * pg[c]);
    }
  }
  rr[t] = sr;
  ri[t] = si;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) { rr[t] += rr[t + w]; ri[t] += ri[t + w]; }
    __syncthreads();
  }
  if (t == 0) { out[2 * blockIdx.x] = rr[0]; out[2 * blockIdx.x + 1] = ri[0]; }
}

// A operand of the Lambda contraction, rows interleaved (2n: w cos, 2n+1: w sin)
// [2 n_k][lda], and the compensated first sum per k_k row as per-block partials.
__global__ __launch_bounds__(256) void k_ref_lam_E(const double* __restrict__ pw,
                                                   const double* __restrict__ pk,
                                                   const int* __restrict__ npop,
                                                   const double* __restrict__ kk, long lda,
                                                   double* __restrict__ E,
                                                   double* __restrict__ a1part) {
  __shared__ double rr[256], ri[256];
  const int n = *npop, t = threadIdx.x, r = blockIdx.y;
  const int cidx = blockIdx.x * 256 + t;
  const double k = kk[r];
  double sr = 0.0, si = 0.0;
  if (cidx < lda) {
    double cs = 0.0, sn = 0.0;
    if (cidx < n) {
      double cm1, smp;
      cexpm1mx(k * pk[cidx], cm1, smp, sn, cs);
      sr = pw[cidx] * cm1;
      si = pw[cidx] * smp;
      cs *= pw[cidx];
      sn *= pw[cidx];
    }
    E[(size_t)(2 * r) * lda + cidx] = cs;
    E[(size_t)(2 * r + 1) * lda + cidx] = sn;
  }
  rr[t] = sr;
  ri[t] = si;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) { rr[t] += rr[t + w]; ri[t] += ri[t + w]; }
    __syncthreads();
  }
  if (t == 0) {
    a1part[((size_t)r * gridDim.x + blockIdx.x) * 2] = rr[0];
    a1part[((size_t)r * gridDim.x + blockIdx.x) * 2 + 1] = ri[0];
  }
}

__global__ void k_ref_lam_a1(const double* __restrict__ a1part, int nblk, int nk,
                             double* __restrict__ a1) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nk) return;
  double sr = 0.0, si = 0.0;
  for (int q = 0; q < nblk; ++q) {
    sr += a1part[((size_t)r * nblk + q) * 2];
    si += a1part[((size_t)r * nblk + q) * 2 + 1];
  }
  a1[2 * r] = sr;
  a1[2 * r + 1] = si;
}

// B operand [N_bins][ldb]: J0(k_g gam_b) - 1 (zero past the bins).
__global__ void k_ref_lam_B(const double* __restrict__ pg, const int* __restrict__ npop,
                            const double* __restrict__ kg, int ng, long ldb, int rows_pad,
                            double* __restrict__ Bm) {
  const int c = blockIdx.x, m = blockIdx.y * blockDim.x + threadIdx.x;
  if (m >= ldb || c >= rows_pad) return;
  const int n = *npop;
  Bm[(size_t)c * ldb + m] = (c < n && m < ng) ? j0m1(kg[m] * pg[c]) : 0.0;
}

// ---------------------------------------------------- FP64 MFMA GEMM ----
// C[M][N] = A[M][K] B[K][N] (row-major, leading dimensions lda, ldb, ldc) on
// v_mfma_f64_16x16x4_f64.  64 x 64 tile per 256-thread workgroup (2 x 2 waves
// of 32 x 32, each 2 x 2 MFMA blocks), K in steps of 16 through LDS with a
// register prefetch of the next step.  Lane l of an MFMA holds A[l % 16][l / 16]
// and B[l / 16][l % 16]; its four results are rows 4 (l / 16) + r, column l % 16.
// Epilogues:
//   kEpiStore  C = acc
//   kEpiPhi    rows interleaved (2n: re, 2n+1: im): Lambda = acc + a1[n]
//              (+ psi[n][col] if given), C = exp(Lambda)
//   kEpiP3     C = cfac acc + a[i] + b[i] gam_col^2
constexpr int kEpiStore = 0, kEpiPhi = 1, kEpiP3 = 2;
typedef double v4d __attribute__((ext_vector_type(4)));
struct EpiArgs {
  const double* a1;    // kEpiPhi: [M/2][2]
  const double* psi;   // kEpiPhi: optional [M/2][N][2]
  const double* va;    // kEpiP3: a[i]
  const double* vb;    // kEpiP3: b[i]
  double cfac, gam_max;
  int n_gam;
  long sA, sB, sC;     // batch strides (blockIdx.z), zero for a single product
};

// TA: A is given transposed (stored K x M, leading dimension lda)
template <int EPI, bool TA = false>
__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K, const double* __restrict__ A,
                                                  long lda, const double* __restrict__ B, long ldb,
                                                  double* __restrict__ C, long ldc, EpiArgs ep) {
  constexpr int BM = 64, BN = 64, BK = 16, PAD = 2;
  __shared__ double As[BK][BM + PAD];
  __shared__ double Bs[BK][BN + PAD];
  A += (size_t)blockIdx.z * ep.sA;
  B += (size_t)blockIdx.z * ep.sB;
  C += (size_t)blockIdx.z * ep.sC;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int wm = wv >> 1, wn = wv & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  // global loads: A row m0 + t/4, k cols (t%4)*4..+3; B row k0 + t/16, cols (t%16)*4..+3
  const int am = t >> 2, ak = (t & 3) * 4;
  const int bk = t >> 4, bn = (t & 15) * 4;
  double ra[4], rb[4];
  auto load = [&](int k0) {
    const int gm = m0 + am;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int gk = k0 + ak + u;
      ra[u] = (gm < M && gk < K) ? (TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk])
                                 : 0.0;
    }
    const int gk = k0 + bk;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int gn = n0 + bn + u;
      rb[u] = (gk < K && gn < N) ? B[(size_t)gk * ldb + gn] : 0.0;
    }
  };
  v4d acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (v4d){0.0, 0.0, 0.0, 0.0};
  load(0);
  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int u = 0; u < 4; ++u) As[ak + u][am] = ra[u];
#pragma unroll
    for (int u = 0; u < 4; ++u) Bs[bk][bn + u] = rb[u];
    __syncthreads();
    if (k0 + BK < K) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      const int kr = kk + (lane >> 4), cl = lane & 15;
      double af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = As[kr][wm * 32 + a * 16 + cl];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[kr][wn * 32 + b * 16 + cl];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // result layout of v_mfma_f64_16x16x4_f64: lane l, register r holds row
  // (l / 16) + 4 r, column l % 16 of the 16 x 16 block
  const int cl = lane & 15, rq = lane >> 4;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = n0 + wn * 32 + b * 16 + cl;
      const int rbase = m0 + wm * 32 + a * 16 + rq;
      if (EPI == kEpiPhi) {
        // rows 2n (real) and 2n + 1 (imaginary) of k row n sit in lane groups
        // 2g and 2g + 1 (lanes l and l ^ 16) with the same register index
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double own = acc[a][b][r];
          const double other = __shfl_xor(own, 16);
          const int row = rbase + 4 * r;
          if (col >= N || row >= M) continue;
          const bool im = (row & 1) != 0;
          const int kr = row >> 1;
          double lr = (im ? other : own) + ep.a1[2 * kr];
          double li = (im ? own : other) + ep.a1[2 * kr + 1];
          if (ep.psi) {
            lr += ep.psi[((size_t)kr * N + col) * 2];
            li += ep.psi[((size_t)kr * N + col) * 2 + 1];
          }
          double s, cc;
          sincos(li, &s, &cc);
          C[(size_t)row * ldc + col] = exp(lr) * (im ? s : cc);
        }
      } else {
        if (col >= N) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rbase + 4 * r;
          if (row >= M) continue;
          double v = acc[a][b][r];
          if (EPI == kEpiP3) {
            const double g = (col == ep.n_gam - 1) ? ep.gam_max
                                                   : col * (ep.gam_max / (ep.n_gam - 1));
            v = ep.cfac * v + ep.va[row] + ep.vb[row] * g * g;
          }
          C[(size_t)row * ldc + col] = v;
        }
      }
    }
}

// -------------------------------------------------------- inversion ----
// np.linspace(lo, hi, n)[i]
__device__ inline double lin(double lo, double hi, int n, int i) {
  return (i == n - 1) ? hi : lo + i * ((hi - lo) / (n - 1));
}

// Hankel operand Mh[m][j] = k_g,m w_m J0(gam_j k_g,m), trapezoid weights w
// (half at the ends) on k_g = linspace(0, kg_max, n_kg), gam = linspace(0,
// gam_max, n_gam) (sgl.py:1143-1150).
__global__ void k_ref_hankel_M(double kg_max, int n_kg, double gam_max, int n_gam, long ldm,
                               double* __restrict__ Mh) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, m = blockIdx.y;
  if (j >= ldm) return;
  double v = 0.0;
  if (j < n_gam) {
    const double dk = lin(0.0, kg_max, n_kg, 1);
    const double kg = lin(0.0, kg_max, n_kg, m);
    const double w = (m == 0 || m == n_kg - 1) ? 0.5 * dk : dk;
    v = kg * w * j0(lin(0.0, gam_max, n_gam, j) * kg);
  }
  Mh[(size_t)m * ldm + j] = v;
}

// Fourier operand Af[i][2n] = w_n cos(kap_i k_n), Af[i][2n+1] = w_n sin(kap_i k_n)
// on k_k = linspace(0, kk_max, n_kk), kap = linspace(kap_min, kap_max, n_kap).
__global__ void k_ref_fourier_A(double kk_max, int n_kk, double kap_min, double kap_max,
                                int n_kap, long lda, double* __restrict__ Af) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (2 * n >= lda) return;
  double cs = 0.0, sn = 0.0;
  if (n < n_kk) {
    const double dk = lin(0.0, kk_max, n_kk, 1);
    const double w = (n == 0 || n == n_kk - 1) ? 0.5 * dk : dk;
    sincos(lin(kap_min, kap_max, n_kap, i) * lin(0.0, kk_max, n_kk, n), &sn, &cs);
    cs *= w;
    sn *= w;
  }
  Af[(size_t)i * lda + 2 * n] = cs;
  if (2 * n + 1 < lda) Af[(size_t)i * lda + 2 * n + 1] = sn;
}

// Euler-Maclaurin endpoint terms of the k_gamma trapezoid (sgl_full.py:277-298):
// a[i] = c Re sum_n w_n e^{-i kap_i k_n} va_n, b[i] likewise with vb_n,
// va = (h^2/12) Phi0 - (h^4/120) Phi2, vb = (h^4/480) Phi0, Phi2 = (Phi1 - Phi0)/h^2.
__global__ __launch_bounds__(256) void k_ref_em(const double* __restrict__ Af, long lda,
                                                const double* __restrict__ Phi, long ldp,
                                                int n_kk, double dkg, double cfac,
                                                double* __restrict__ va_out,
                                                double* __restrict__ vb_out) {
  __shared__ double ra[256], rbv[256];
  const int i = blockIdx.x, t = threadIdx.x;
  const double h2 = dkg * dkg, h4 = h2 * h2;
  double sa = 0.0, sb = 0.0;
  for (int n = t; n < n_kk; n += 256) {
    const double p0r = Phi[(size_t)(2 * n) * ldp], p0i = Phi[(size_t)(2 * n + 1) * ldp];
    const double p1r = Phi[(size_t)(2 * n) * ldp + 1], p1i = Phi[(size_t)(2 * n + 1) * ldp + 1];
    const double p2r = (p1r - p0r) / h2, p2i = (p1i - p0i) / h2;
    const double var = (h2 / 12.0) * p0r - (h4 / 120.0) * p2r;
    const double vai = (h2 / 12.0) * p0i - (h4 / 120.0) * p2i;
    const double vbr = (h4 / 480.0) * p0r, vbi = (h4 / 480.0) * p0i;
    const double c = Af[(size_t)i * lda + 2 * n], s = Af[(size_t)i * lda + 2 * n + 1];
    sa += c * var + s * vai;
    sb += c * vbr + s * vbi;
  }
  ra[t] = sa;
  rbv[t] = sb;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) { ra[t] += ra[t + w]; rbv[t] += rbv[t + w]; }
    __syncthreads();
  }
  if (t == 0) { va_out[i] = cfac * ra[0]; vb_out[i] = cfac * rbv[0]; }
}

struct ArcGrid {
  double kap_min, kap_max, gam_max, g_def;
  int n_kap, n_gam, n_int;
};

// interval i with g[i] <= x < g[i+1] on linspace(lo, hi, n), clamped to [0, n-2]
__device__ inline int lin_interval(double lo, double hi, int n, double x) {
  int i = (int)floor((x - lo) / ((hi - lo) / (n - 1)));
  i = max(0, min(i, n - 2));
  while (i > 0 && lin(lo, hi, n, i) > x) --i;
  while (i < n - 2 && lin(lo, hi, n, i + 1) <= x) ++i;
  return i;
}

// Linear RegularGridInterpolator on (kap, gam) with fill 0 outside (scipy's
// evaluate_linear_2d term order)
__device__ inline double bilin(const double* __restrict__ P3, long ldp, const ArcGrid& g, double k,
                               double gm) {
  if (!(k >= g.kap_min && k <= g.kap_max && gm >= 0.0 && gm <= g.gam_max)) return 0.0;
  const int i = lin_interval(g.kap_min, g.kap_max, g.n_kap, k);
  const int j = lin_interval(0.0, g.gam_max, g.n_gam, gm);
  const double k0 = lin(g.kap_min, g.kap_max, g.n_kap, i), k1 = lin(g.kap_min, g.kap_max, g.n_kap, i + 1);
  const double g0 = lin(0.0, g.gam_max, g.n_gam, j), g1 = lin(0.0, g.gam_max, g.n_gam, j + 1);
  const double y0 = (k - k0) / (k1 - k0), y1 = (gm - g0) / (g1 - g0);
  const double* r0 = P3 + (size_t)i * ldp + j;
  const double* r1 = r0 + ldp;
  return r0[0] * (1 - y0) * (1 - y1) + r0[1] * (1 - y0) * y1 + r1[0] * y0 * (1 - y1) +
         r1[1] * y0 * y1;
}

// P_i of invert (sgl_full.py:299-360) at one ln mu per workgroup: pi/mu^2 times
// the trapezoid of P3 along the constant-mu arc over n_int kappa points
// (gamma <= g_def), plus the tail-box segment parametrised by gamma at half the
// gamma step.  Divided by *norm when given.
__global__ __launch_bounds__(256) void k_ref_arcs(const double* __restrict__ P3, long ldp,
                                                  ArcGrid g, const double* __restrict__ xi,
                                                  const double* __restrict__ norm,
                                                  double* __restrict__ out) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  const double x = xi[blockIdx.x];
  const double mu = exp(x), rm = 1.0 / sqrt(mu);
  const double hi = 1.0 - rm;
  const double lo = fmax(g.kap_min, 1.0 - sqrt(1.0 / mu + g.g_def * g.g_def));
  double s = 0.0;
  const bool seg = hi > lo;
  if (seg) {
    // trapezoid over n_int points: thread t takes intervals [t * per, ...)
    const int n = g.n_int, per = (n - 1 + 255) / 256;
    const int a = min(n - 1, t * per), b = min(n - 1, a + per);
    if (a < b) {
      double kprev = lin(lo, hi, n, a);
      double vprev = bilin(P3, ldp, g, kprev, sqrt(fmax((hi - kprev) * (1.0 - kprev + rm), 0.0)));
      for (int q = a + 1; q <= b; ++q) {
        const double kq = lin(lo, hi, n, q);
        const double vq = bilin(P3, ldp, g, kq, sqrt(fmax((hi - kq) * (1.0 - kq + rm), 0.0)));
        s += (kq - kprev) * (vprev + vq) / 2.0;
        kprev = kq;
        vprev = vq;
      }
    }
  }
  red[t] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double main = red[0];
  __syncthreads();
  double s2 = 0.0;
  double gb = 0.0;
  int nt = 0;
  if (seg && g.gam_max > g.g_def && lo > g.kap_min) {
    gb = fmin(g.gam_max, sqrt(fmax((hi - g.kap_min) * (1.0 - g.kap_min + rm), 0.0)));
    if (gb > g.g_def) {
      const double dgam = lin(0.0, g.gam_max, g.n_gam, 1);
      nt = (int)ceil((gb - g.g_def) / (0.5 * dgam)) + 1;
      const int per = (nt - 1 + 255) / 256;
      const int a = min(nt - 1, t * per), b = min(nt - 1, a + per);
      auto f = [&](int q, double& gx) {
        gx = lin(g.g_def, gb, nt, q);
        const double r = sqrt(1.0 / mu + gx * gx);
        const double kx = fmin(fmax(1.0 - r, g.kap_min), g.kap_max);
        return bilin(P3, ldp, g, kx, gx) * gx / r;
      };
      if (a < b) {
        double gprev, vprev = f(a, gprev);
        for (int q = a + 1; q <= b; ++q) {
          double gq;
          const double vq = f(q, gq);
          s2 += (gq - gprev) * (vprev + vq) / 2.0;
          gprev = gq;
          vprev = vq;
        }
      }
    }
  }
  red[t] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) {
    const double pref = kPi / (mu * mu);
    double v = seg ? pref * main : 0.0;
    if (nt > 0) v += pref * red[0];
    out[blockIdx.x] = norm ? v / *norm : v;
  }
}

// Normalisation and moments of the wide grid (sgl_full.py:362-371):
// mom = [norm, mean, var, skew, sigma_DL/D_L], trapezoids on xw.
__global__ __launch_bounds__(256) void k_ref_moments(const double* __restrict__ xw,
                                                     const double* __restrict__ pw, int n,
                                                     double* __restrict__ mom) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  auto trap = [&](auto f) {
    double s = 0.0;
    for (int q = t; q + 1 < n; q += 256) s += (xw[q + 1] - xw[q]) * (f(q) + f(q + 1)) / 2.0;
    red[t] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) red[t] += red[t + w];
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  const double norm = trap([&](int q) { return pw[q]; });
  const double mean = trap([&](int q) { return xw[q] * (pw[q] / norm); });
  const double var = trap([&](int q) { const double d = xw[q] - mean; return d * d * (pw[q] / norm); });
  const double m3 = trap([&](int q) { const double d = xw[q] - mean; return d * d * d * (pw[q] / norm); });
  const double md = trap([&](int q) { return exp(-0.5 * xw[q]) * (pw[q] / norm); });
  const double v2 = trap([&](int q) {
    const double d = exp(-0.5 * xw[q]) - md;
    return d * d * (pw[q] / norm);
  });
  if (t == 0) {
    mom[0] = norm;
    mom[1] = mean;
    mom[2] = var;
    mom[3] = m3 / (var * sqrt(var));
    mom[4] = sqrt(v2) / md;
  }
}

// Per-block sums of w kappa^2 over the compacted bins (Lam.sigma, sgl_full.py:3439).
__global__ __launch_bounds__(256) void k_ref_wk2(const double* __restrict__ pw,
                                                 const double* __restrict__ pk,
                                                 const int* __restrict__ npop,
                                                 double* __restrict__ part) {
  __shared__ double red[256];
  const int t = threadIdx.x, c = blockIdx.x * 256 + t, n = *npop;
  red[t] = (c < n) ? pw[c] * pk[c] * pk[c] : 0.0;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = red[0];
}

// Edges of the fine lattice (sgl_full.py:456-468): kappa = concat(-pos[::-1], pos),
// pos = geomspace(1e-7, 30, 481); gamma = geomspace(1e-9, 30, 301).
__global__ void k_ref_fine_edges(double* __restrict__ ek, double* __restrict__ eg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kFineKRef + 1) {
    const int npos = (kFineKRef + 1) / 2;
    ek[i] = (i < npos) ? -geom(1e-7, 30.0, npos, npos - 1 - i) : geom(1e-7, 30.0, npos, i - npos);
  }
  if (i < kFineGRef + 1) eg[i] = geom(1e-9, 30.0, kFineGRef + 1, i);
}

// ------------------------------------------------------------ launchers --
struct RefTablesView {
  const double *cosmo, *rows, *frows, *ek, *eg;
};

int ref_roots(const RefTablesView& T, int n_z_max, int ell, double* roots, int* status,
              hipStream_t st) {
  hipLaunchKernelGGL(k_ref_roots, dim3(kRefMasses, n_z_max), dim3(256), 0, st, T.cosmo, T.rows,
                     n_z_max, ell, roots, status);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int ref_xref(const RefTablesView& T, int n_z_max, int ell, const double* roots, double* subh,
             int* status, hipStream_t st) {
  hipLaunchKernelGGL(k_ref_cells<kModeXref>, dim3(n_z_max * kRefMasses), dim3(256), 0, st, T.cosmo,
                     T.rows, T.frows, roots, T.ek, T.eg, ell, 0, (long long*)nullptr,
                     (const long long*)nullptr, (double*)nullptr, (double*)nullptr,
                     (double*)nullptr, (int*)nullptr, status, subh, (const double*)nullptr,
                     XiEmit{});
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

// counts: [ntask + 1] points, with xi (xe.ccounts = counts + ntask + 1) also
// [ntask + 1] radii; both scanned in place, totals last
int ref_cells(const RefTablesView& T, int n_z_max, int ell, int fil, const double* roots,
              long long* counts, int* status, double* subh, const double* subc, const XiEmit& xe,
              hipStream_t st) {
  const int nph = ell ? kEllPhi : 1;
  const int ntask = n_z_max * kRefMasses * nph;
  hipLaunchKernelGGL(k_ref_cells<kModeCount>, dim3(ntask), dim3(256), 0, st, T.cosmo, T.rows,
                     T.frows, roots, T.ek, T.eg, ell, fil, counts, (const long long*)nullptr,
                     (double*)nullptr, (double*)nullptr, (double*)nullptr, (int*)nullptr, status,
                     subh, subc, xe);
  hipLaunchKernelGGL(k_excl_scan_i64, dim3(1), dim3(256), 0, st, counts, ntask);
  if (xe.ccounts) hipLaunchKernelGGL(k_excl_scan_i64, dim3(1), dim3(256), 0, st, xe.ccounts, ntask);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int ref_emit(const RefTablesView& T, int n_z_max, int ell, int fil, const double* roots,
             const long long* offsets, double* K, double* G, double* W, int* key, int* status,
             double* subh, const double* subc, const XiEmit& xe, hipStream_t st) {
  const int nph = ell ? kEllPhi : 1;
  const int ntask = n_z_max * kRefMasses * nph;
  hipLaunchKernelGGL(k_ref_cells<kModeEmit>, dim3(ntask), dim3(256), 0, st, T.cosmo, T.rows,
                     T.frows, roots, T.ek, T.eg, ell, fil, (long long*)nullptr, offsets, K, G, W,
                     key, status, subh, subc, xe);
  if (subh)
    hipLaunchKernelGGL(k_sub_emit, dim3(ntask), dim3(256), 0, st, T.cosmo, T.ek, T.eg, ell, offsets,
                       subh, subc, K, G, W, key, xe);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

// Workspace of the binning (bytes): sorted keys and indices, cell indices,
// run-length outputs, run starts and sums, rocPRIM temporaries, partials.
struct BinWs {
  size_t skey, vin, vout, ukey, rlen, nruns, rstart, nparts, poff, psum, rsum, part, tmp, total,
      tmp_bytes;
  long long max_parts;
};
inline size_t al256(size_t n) { return (n + 255) & ~size_t(255); }
BinWs bin_ws(long long n) {
  BinWs L;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += al256(b); return r; };
  size_t sort_b = 0, rle_b = 0, scan_b = 0, scan2_b = 0;
  (void)rocprim::radix_sort_pairs((void*)nullptr, sort_b, (int*)nullptr, (int*)nullptr,
                                  (KGW*)nullptr, (KGW*)nullptr, (size_t)n, 0, 20);
  (void)rocprim::run_length_encode((void*)nullptr, rle_b, (int*)nullptr, (size_t)n, (int*)nullptr,
                                   (int*)nullptr, (int*)nullptr);
  (void)rocprim::exclusive_scan((void*)nullptr, scan_b, (int*)nullptr, (long long*)nullptr, 0LL,
                                (size_t)kRefBins + 1, rocprim::plus<long long>());
  scan2_b = scan_b;
  L.tmp_bytes = std::max(std::max(sort_b, rle_b), std::max(scan_b, scan2_b));
  const size_t nr = (size_t)kRefBins + 1;   // runs: every bin and the outside key
  L.max_parts = (long long)nr + n / kRunPart + 1;
  L.skey = take(n * sizeof(int));
  L.vin = take(n * sizeof(KGW));
  L.vout = take(n * sizeof(KGW));
  L.ukey = take(nr * sizeof(int));
  L.rlen = take(nr * sizeof(int));
  L.nruns = take(sizeof(int));
  L.rstart = take(nr * sizeof(long long));
  L.nparts = take(nr * sizeof(int));
  L.poff = take(nr * sizeof(long long));
  L.psum = take((size_t)L.max_parts * 3 * sizeof(double));
  L.rsum = take((size_t)kRefBins * 4 * sizeof(double) + 4 * sizeof(double));
  L.part = take(4096 * sizeof(double));
  L.tmp = take(L.tmp_bytes);
  L.total = o;
  return L;
}

// Bin the emitted cells: stable sort by key, runs, fixed-order run sums,
// compaction; pop = [w][k][g] (N_bins each, capacity kRefBins), npop, and
// scal = [sum w kappa^2 over the bins, sum W K over the cells].
int ref_bin(long long n, const int* key, const double* K, const double* G, const double* W,
            void* ws, double* pw, double* pk, double* pg, int* npop, double* scal,
            hipStream_t st) {
  const BinWs L = bin_ws(n);
  char* base = static_cast<char*>(ws);
  int* skey = (int*)(base + L.skey);
  KGW* vin = (KGW*)(base + L.vin);
  KGW* vout = (KGW*)(base + L.vout);
  int* ukey = (int*)(base + L.ukey);
  int* rlen = (int*)(base + L.rlen);
  int* nruns = (int*)(base + L.nruns);
  long long* rstart = (long long*)(base + L.rstart);
  int* nparts = (int*)(base + L.nparts);
  long long* poff = (long long*)(base + L.poff);
  double* psum = (double*)(base + L.psum);
  double* rsum = (double*)(base + L.rsum);
  double* part = (double*)(base + L.part);
  void* tmp = base + L.tmp;
  size_t tb = L.tmp_bytes;
  hipLaunchKernelGGL(k_pack_kgw, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, K, G, W,
                     vin);
  // stable: equal keys keep the cells' emission order
  if (rocprim::radix_sort_pairs(tmp, tb, key, skey, vin, vout, (size_t)n, 0, 20, st) != hipSuccess)
    return SGL_ERR_HIP;
  tb = L.tmp_bytes;
  if (rocprim::run_length_encode(tmp, tb, skey, (size_t)n, ukey, rlen, nruns, st) != hipSuccess)
    return SGL_ERR_HIP;
  tb = L.tmp_bytes;
  // run starts over every run slot (entries past the run count are never read)
  hipLaunchKernelGGL(k_zero_tail_runs, dim3((kRefBins + 1 + 255) / 256), dim3(256), 0, st, rlen,
                     nruns);
  if (rocprim::exclusive_scan(tmp, tb, rlen, rstart, 0LL, (size_t)kRefBins + 1,
                              rocprim::plus<long long>(), st) != hipSuccess)
    return SGL_ERR_HIP;
  tb = L.tmp_bytes;
  hipLaunchKernelGGL(k_run_nparts, dim3((kRefBins + 1 + 255) / 256), dim3(256), 0, st, rlen, nruns,
                     nparts);
  if (rocprim::exclusive_scan(tmp, tb, nparts, poff, 0LL, (size_t)kRefBins + 1,
                              rocprim::plus<long long>(), st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_run_parts, dim3((unsigned)((L.max_parts + 3) / 4)), dim3(256), 0, st, rlen,
                     nruns, rstart, nparts, poff, vout, psum);
  hipLaunchKernelGGL(k_run_combine, dim3((kRefBins + 1 + 255) / 256), dim3(256), 0, st, ukey,
                     nruns, nparts, poff, psum, rsum);
  hipLaunchKernelGGL(k_compact_bins, dim3(1), dim3(256), 0, st, rsum, nruns, pw, pk, pg, npop);
  // sum w kappa^2 = sum over bins of (w kappa) kappa; sum W K over every cell
  const long long chunk_b = 256;
  const int nb_blocks = (kRefBins + chunk_b - 1) / chunk_b;
  hipLaunchKernelGGL(k_ref_wk2, dim3(nb_blocks), dim3(256), 0, st, pw, pk, npop, part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, part, nb_blocks, scal);
  const long long chunk = std::max<long long>(256, (n + 4095) / 4096);
  const int nc = (int)((n + chunk - 1) / chunk);
  hipLaunchKernelGGL(k_chunk_sums, dim3(nc), dim3(256), 0, st, W, K, n, chunk, part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, st, part, nc, scal + 1);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int xi_gemm_batched(int M, int N, int K, const double* A, long lda, long sA, bool transA,
                    const double* B, long ldb, long sB, double* C, long ldc, long sC, int batch,
                    hipStream_t st) {
  if (M < 1 || N < 1 || K < 1 || batch < 1 || (M + 63) / 64 > 65535 || batch > 65535)
    return SGL_ERR_INVALID;
  EpiArgs ep = {};
  ep.sA = sA;
  ep.sB = sB;
  ep.sC = sC;
  const dim3 grid((N + 63) / 64, (M + 63) / 64, batch);
  if (transA)
    hipLaunchKernelGGL((k_gemm_f64<kEpiStore, true>), grid, dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, C, ldc, ep);
  else
    hipLaunchKernelGGL((k_gemm_f64<kEpiStore, false>), grid, dim3(256), 0, st, M, N, K, A, lda, B,
                       ldb, C, ldc, ep);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

template <int EPI>
static void gemm_f64(int M, int N, int K, const double* A, long lda, const double* B, long ldb,
                     double* C, long ldc, const EpiArgs& ep, hipStream_t st) {
  hipLaunchKernelGGL(k_gemm_f64<EPI>, dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, st, M, N, K,
                     A, lda, B, ldb, C, ldc, ep);
}

// Tables of one instance on the reference grid: the kernel-shape layout with
// the reference grid's dimensions, then the fine-lattice edges.
sgl_shape ref_shape(int flags) {
  sgl_shape s = {};
  s.n_z_max = kRefMaxShells;
  s.n_m = kRefMasses;
  s.n_b = 1;
  s.n_k = 2;
  s.n_out = 2;
  s.flags = (flags & (SGL_FLAG_ELL | SGL_FLAG_FILAMENTS | SGL_FLAG_EDGE_RVIR | SGL_FLAG_SUBHALOS)) |
            SGL_FLAG_REF_GRID;
  s.m_lo_h = 1.0; s.m_hi_h = 2.0; s.u_lo = 1.0; s.u_hi = 2.0;
  s.k_lo = 1.0; s.k_hi = 2.0; s.out_lo = 0.0; s.out_hi = 1.0;
  return s;
}
struct RefLayout {
  TablesLayout T;
  size_t ek, eg, total;
};
RefLayout ref_layout(int flags) {
  RefLayout L;
  L.T = tables_layout(ref_shape(flags), 1);
  L.ek = L.T.total;
  L.eg = L.ek + round32(kFineKRef + 1);
  L.total = L.eg + round32(kFineGRef + 1);
  return L;
}
RefTablesOffsets ref_tables_offsets(int flags) {
  const RefLayout L = ref_layout(flags);
  return RefTablesOffsets{L.T.cosmo, L.T.derived, L.T.slices, L.T.rows};
}
struct PhiWs {
  size_t E, Bm, a1p, a1, total;
  long lda;
  int nblk;
};
PhiWs phi_ws(int nk, int ng, int nbins) {
  PhiWs W;
  W.lda = ((long)std::max(nbins, 1) + 63) / 64 * 64;
  W.nblk = (int)((W.lda + 255) / 256);
  size_t o = 0;
  auto take = [&](size_t n) { size_t r = o; o += round32(n); return r; };
  W.E = take((size_t)2 * nk * W.lda);
  W.Bm = take((size_t)W.lda * ng);
  W.a1p = take((size_t)nk * W.nblk * 2);
  W.a1 = take((size_t)nk * 2);
  W.total = o;
  return W;
}
struct P3Ws {
  size_t Mh, F, Af, va, vb, total;
};
P3Ws p3_ws(int nk, int ng, int n_kap, int n_gam) {
  P3Ws W;
  size_t o = 0;
  auto take = [&](size_t n) { size_t r = o; o += round32(n); return r; };
  W.Mh = take((size_t)ng * n_gam);
  W.F = take((size_t)2 * nk * n_gam);
  W.Af = take((size_t)n_kap * 2 * nk);
  W.va = take((size_t)n_kap);
  W.vb = take((size_t)n_kap);
  W.total = o;
  return W;
}

__global__ void k_ref_init(const double* __restrict__ src, double* __restrict__ dst,
                           int* __restrict__ status) {
  const int t = threadIdx.x;
  if (t < SGL_COSMO_STRIDE) dst[t] = src[t];
  if (t == 0) {
    *status = 0;
    const int nz = (int)src[10];
    if (!(src[9] > 0.0) || nz < 1 || nz > kRefMaxShells) *status = kStatusBadShape;
    // engine shells: at most the 102 intervals of the 103-node engine grid;
    // gauss: whole shells of zint_n >= 1 nodes
    const double zi = src[SGL_REC_ZINT], zn = src[SGL_REC_ZINT_N];
    if (zi == 1.0 && nz > 102) *status = kStatusBadShape;
    if (zi == 2.0 && !(zn >= 1.0 && zn <= kRefMaxShells && nz % (int)zn == 0))
      *status = kStatusBadShape;
    if (zi != 0.0 && zi != 1.0 && zi != 2.0) *status = kStatusBadShape;
  }
}

// The Gaussian large-scale-structure term of clustering="limber":
// -var/2 (k_kappa^2 + k_gamma^2 / 2) added to ln Phi (sgl_full.py:3104-3105;
// the shear carries the convergence's variance split over two components).
// On the tensor grid into psi [nk][ng][2] (the clustering-term input of
// sgl_ref_phi), or added to the axis probes out [n][2].
__global__ __launch_bounds__(256) void k_ref_lss_psi(double var, const double* __restrict__ k_k,
                                                     int nk, const double* __restrict__ k_g,
                                                     int ng, double* __restrict__ psi) {
  const long long n = (long long)nk * ng;
  const double hv = 0.5 * var;
  for (long long q = blockIdx.x * 256ll + threadIdx.x; q < n; q += (long long)gridDim.x * 256) {
    const double a = k_k[q / ng], b = k_g[q % ng];
    psi[2 * q] = -(hv * (a * a + 0.5 * (b * b)));
    psi[2 * q + 1] = 0.0;
  }
}

__global__ __launch_bounds__(256) void k_ref_lss_axis(double var, const double* __restrict__ k,
                                                      int n, int axis, double* __restrict__ out) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double hv = 0.5 * var, kk = k[i];
  out[2 * i] -= (axis == 0) ? hv * (kk * kk + 0.0) : hv * (0.0 + 0.5 * (kk * kk));
}

}  // namespace sgl

using namespace sgl;

namespace {
hipStream_t ref_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
// subhalos only with the r200 edge (the reference pairs clump_edge with the
// halo edge; its r_vir clumps are not built here)
bool ref_flags_ok(int32_t f) {
  if (f & ~(SGL_FLAG_ELL | SGL_FLAG_FILAMENTS | SGL_FLAG_EDGE_RVIR | SGL_FLAG_SUBHALOS |
            SGL_FLAG_CLUSTERING))
    return false;
  return !((f & SGL_FLAG_SUBHALOS) && (f & SGL_FLAG_EDGE_RVIR));
}
sgl_status ref_rc(int rc) { return (sgl_status)rc; }
}  // namespace

extern "C" {

int32_t sgl_ref_shell_count(double z_s) {
  if (!(z_s > 0.0) || !std::isfinite(z_s)) return 0;
  const double l0 = std::log(0.01), l1 = std::log(12.341169644129371);
  int n = 0;
  for (int j = 0; j < 103; ++j) {
    const double z = std::exp((j == 102) ? l1 : l0 + j * ((l1 - l0) / 102));
    if (z < z_s) ++n;
  }
  return n + 1;
}

int32_t sgl_ref_tasks(int32_t flags) {
  if (!ref_flags_ok(flags)) return 0;
  return kRefMaxShells * kRefMasses * ((flags & SGL_FLAG_ELL) ? kEllPhi : 1);
}

size_t sgl_ref_tables_bytes(int32_t flags) {
  if (!ref_flags_ok(flags)) return 0;
  return ref_layout(flags).total * sizeof(double);
}

size_t sgl_ref_roots_bytes(void) {
  return (size_t)kRefMaxShells * kRefMasses * kRootStride * sizeof(double);
}

sgl_status sgl_ref_tables_build(const double* cosmo, int32_t flags, void* tables, int32_t* status,
                                void* stream) {
  if (!ref_flags_ok(flags) || !cosmo || !tables || !status) return SGL_ERR_INVALID;
  const RefLayout L = ref_layout(flags);
  const sgl_shape sh = ref_shape(flags);
  hipStream_t st = ref_stream(stream);
  double* T = static_cast<double*>(tables);
  hipLaunchKernelGGL(k_ref_init, dim3(1), dim3(64), 0, st, cosmo, T + L.T.cosmo, status);
  int rc = launch_fill_grids(sh, T + L.T.uedge, T + L.T.knodes, T + L.T.m_nodes, st);
  if (rc != SGL_OK) return ref_rc(rc);
  rc = ref_tables_build(T + L.T.cosmo, 1, sh, T + L.T.m_nodes, T + L.T.derived, T + L.T.slices,
                        T + L.T.mass, T + L.T.rows, T + L.T.frows, T + L.T.sigtab,
                        T + L.T.sigspl, T + L.T.gdense, st);
  if (rc != SGL_OK) return ref_rc(rc);
  hipLaunchKernelGGL(k_ref_fine_edges, dim3(4), dim3(256), 0, st, T + L.ek, T + L.eg);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

static RefTablesView ref_view(const void* tables, int flags) {
  const RefLayout L = ref_layout(flags);
  const double* T = static_cast<const double*>(tables);
  RefTablesView v;
  v.cosmo = T + L.T.cosmo;
  v.rows = T + L.T.rows;
  v.frows = (flags & SGL_FLAG_FILAMENTS) ? T + L.T.frows : nullptr;
  v.ek = T + L.ek;
  v.eg = T + L.eg;
  return v;
}

size_t sgl_ref_sub_bytes(void) { return sub_layout().total * sizeof(double); }

sgl_status sgl_ref_sub_build(const void* tables, int32_t flags, const double* roots, void* sub,
                             int32_t* status, void* stream) {
  if (!ref_flags_ok(flags) || !(flags & SGL_FLAG_SUBHALOS) || !tables || !roots || !sub || !status)
    return SGL_ERR_INVALID;
  const RefLayout L = ref_layout(flags);
  const double* T = static_cast<const double*>(tables);
  const SubLayout S = sub_layout();
  double* D = static_cast<double*>(sub);
  hipStream_t st = ref_stream(stream);
  int rc = sub_hosts_launch(T + L.T.cosmo, T + L.T.derived, T + L.T.slices, T + L.T.mass,
                            T + L.T.rows, T + L.T.sigspl, T + L.T.gdense, D + S.host, st);
  if (rc != SGL_OK) return ref_rc(rc);
  rc = ref_xref(ref_view(tables, flags), kRefMaxShells, (flags & SGL_FLAG_ELL) ? 1 : 0, roots,
                D + S.host, status, st);
  if (rc != SGL_OK) return ref_rc(rc);
  return ref_rc(sub_coarse_launch(T + L.T.cosmo, T + L.T.derived, T + L.T.slices, T + L.T.sigspl,
                                  (flags & SGL_FLAG_ELL) ? T + L.T.sigtab : nullptr, D + S.host,
                                  D + S.coarse, D + S.ring, st));
}

sgl_status sgl_ref_roots(const void* tables, int32_t flags, double* roots, int32_t* status,
                         void* stream) {
  if (!ref_flags_ok(flags) || !tables || !roots || !status) return SGL_ERR_INVALID;
  return ref_rc(ref_roots(ref_view(tables, flags), kRefMaxShells, (flags & SGL_FLAG_ELL) ? 1 : 0,
                          roots, status, ref_stream(stream)));
}

static bool sub_ok(int32_t flags, const void* sub) {
  return ((flags & SGL_FLAG_SUBHALOS) != 0) == (sub != nullptr);
}

sgl_status sgl_ref_cells_count(const void* tables, int32_t flags, const double* roots,
                               const void* sub, int64_t* counts, int32_t* status, void* stream) {
  if (!ref_flags_ok(flags) || !tables || !roots || !counts || !status || !sub_ok(flags, sub))
    return SGL_ERR_INVALID;
  const SubLayout S = sub_layout();
  double* D = const_cast<double*>(static_cast<const double*>(sub));
  XiEmit xe{};
  if (flags & SGL_FLAG_CLUSTERING)
    xe.ccounts = reinterpret_cast<long long*>(counts) + sgl_ref_tasks(flags) + 1;
  return ref_rc(ref_cells(ref_view(tables, flags), kRefMaxShells, (flags & SGL_FLAG_ELL) ? 1 : 0,
                          (flags & SGL_FLAG_FILAMENTS) ? 1 : 0, roots,
                          reinterpret_cast<long long*>(counts), status, D ? D + S.host : nullptr,
                          D ? D + S.coarse : nullptr, xe, ref_stream(stream)));
}

sgl_status sgl_ref_cells_emit(const void* tables, int32_t flags, const double* roots,
                              const void* sub, const int64_t* offsets, double* kap, double* gam,
                              double* w, int32_t* key, int32_t* status, void* stream) {
  if (!ref_flags_ok(flags) || !tables || !roots || !offsets || !kap || !gam || !w || !key ||
      !status || !sub_ok(flags, sub))
    return SGL_ERR_INVALID;
  const SubLayout S = sub_layout();
  double* D = const_cast<double*>(static_cast<const double*>(sub));
  return ref_rc(ref_emit(ref_view(tables, flags), kRefMaxShells, (flags & SGL_FLAG_ELL) ? 1 : 0,
                         (flags & SGL_FLAG_FILAMENTS) ? 1 : 0, roots,
                         reinterpret_cast<const long long*>(offsets), kap, gam, w, key, status,
                         D ? D + S.host : nullptr, D ? D + S.coarse : nullptr, XiEmit{},
                         ref_stream(stream)));
}

sgl_status sgl_ref_cells_emit_xi(const void* tables, int32_t flags, const double* roots,
                                 const void* sub, const int64_t* offsets, int64_t n_points,
                                 double* kap, double* gam, double* w, int32_t* key,
                                 const void* xi_state, int32_t* xi_pts, double* xi_cells,
                                 int32_t* status, void* stream) {
  if (!ref_flags_ok(flags) || !(flags & SGL_FLAG_CLUSTERING) || !tables || !roots || !offsets ||
      n_points < 1 || n_points > ((int64_t)1 << 31) - 1 || !kap || !gam || !w || !key ||
      !xi_state || !xi_pts || !xi_cells || !status || !sub_ok(flags, sub))
    return SGL_ERR_INVALID;
  const SubLayout S = sub_layout();
  double* D = const_cast<double*>(static_cast<const double*>(sub));
  const int ntask = sgl_ref_tasks(flags);
  const long long* off = reinterpret_cast<const long long*>(offsets);
  const RefLayout RL = ref_layout(flags);
  const XiState X = xi_state_layout();
  const double* xs = static_cast<const double*>(xi_state);
  XiEmit xe{};
  xe.coffsets = off + ntask + 1;
  xe.ekc = xs + X.ekc;
  xe.egc = xs + X.egc;
  xe.slices = static_cast<const double*>(tables) + RL.T.slices;
  xe.pcid = xi_pts;
  xe.pck = xi_pts + n_points;
  xe.cells = xi_cells;
  return ref_rc(ref_emit(ref_view(tables, flags), kRefMaxShells, (flags & SGL_FLAG_ELL) ? 1 : 0,
                         (flags & SGL_FLAG_FILAMENTS) ? 1 : 0, roots, off, kap, gam, w, key,
                         status, D ? D + S.host : nullptr, D ? D + S.coarse : nullptr, xe,
                         ref_stream(stream)));
}

size_t sgl_ref_bin_bytes(int64_t n_cells) {
  if (n_cells < 1 || n_cells > (int64_t)1 << 31) return 0;
  return bin_ws(n_cells).total;
}

sgl_status sgl_ref_bin(int64_t n_cells, const int32_t* key, const double* kap, const double* gam,
                       const double* w, void* ws, double* pop, int32_t* npop, double* scal,
                       void* stream) {
  if (n_cells < 1 || n_cells > (int64_t)1 << 31 || !key || !kap || !gam || !w || !ws || !pop ||
      !npop || !scal)
    return SGL_ERR_INVALID;
  return ref_rc(ref_bin(n_cells, key, kap, gam, w, ws, pop, pop + kRefBins, pop + 2 * (size_t)kRefBins,
                        npop, scal, ref_stream(stream)));
}

sgl_status sgl_ref_lam_axis(const double* pop, const int32_t* npop, const double* k, int32_t n,
                            int32_t axis, double* out, void* stream) {
  if (!pop || !npop || !k || !out || n < 1 || n > 65535 || (axis != 0 && axis != 1))
    return SGL_ERR_INVALID;
  hipLaunchKernelGGL(k_ref_lam_axis, dim3(n), dim3(256), 0, ref_stream(stream), pop, pop + kRefBins,
                     pop + 2 * (size_t)kRefBins, npop, k, axis, out);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_phi_bytes(int32_t nk, int32_t ng, int32_t nbins) {
  if (nk < 1 || ng < 1 || nbins < 1 || nbins > kRefBins) return 0;
  return phi_ws(nk, ng, nbins).total * sizeof(double);
}

sgl_status sgl_ref_phi(const double* pop, const int32_t* npop, int32_t nbins, const double* kk,
                       int32_t nk, const double* kg, int32_t ng, const double* psi, void* ws,
                       double* phi, void* stream) {
  if (!pop || !npop || !kk || !kg || !ws || !phi || nk < 1 || nk > 65535 || ng < 1 ||
      nbins < 1 || nbins > kRefBins)
    return SGL_ERR_INVALID;
  hipStream_t st = ref_stream(stream);
  const PhiWs W = phi_ws(nk, ng, nbins);
  double* D = static_cast<double*>(ws);
  const double *pw = pop, *pk = pop + kRefBins, *pg = pop + 2 * (size_t)kRefBins;
  hipLaunchKernelGGL(k_ref_lam_E, dim3(W.nblk, nk), dim3(256), 0, st, pw, pk, npop, kk, W.lda,
                     D + W.E, D + W.a1p);
  hipLaunchKernelGGL(k_ref_lam_a1, dim3((nk + 255) / 256), dim3(256), 0, st, D + W.a1p, W.nblk, nk,
                     D + W.a1);
  hipLaunchKernelGGL(k_ref_lam_B, dim3((unsigned)W.lda, (ng + 255) / 256), dim3(256), 0, st, pg,
                     npop, kg, ng, (long)ng, (int)W.lda, D + W.Bm);
  EpiArgs ep = {};
  ep.a1 = D + W.a1;
  ep.psi = psi;
  gemm_f64<kEpiPhi>(2 * nk, ng, (int)W.lda, D + W.E, W.lda, D + W.Bm, ng, phi, ng, ep, st);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_p3_bytes(int32_t nk, int32_t ng, int32_t n_kap, int32_t n_gam) {
  if (nk < 2 || ng < 2 || n_kap < 2 || n_gam < 2) return 0;
  return p3_ws(nk, ng, n_kap, n_gam).total * sizeof(double);
}

sgl_status sgl_ref_p3(const double* phi, int32_t nk, int32_t ng, double kk_max, double kg_max,
                      double kap_min, double kap_max, int32_t n_kap, double gam_max, int32_t n_gam,
                      void* ws, double* p3, void* stream) {
  if (!phi || !ws || !p3 || nk < 2 || ng < 2 || n_kap < 2 || n_gam < 2 || ng > 65535 ||
      n_kap > 65535 || !(kk_max > 0) || !(kg_max > 0) || !(kap_max > kap_min) || !(gam_max > 0))
    return SGL_ERR_INVALID;
  hipStream_t st = ref_stream(stream);
  const P3Ws W = p3_ws(nk, ng, n_kap, n_gam);
  double* D = static_cast<double*>(ws);
  hipLaunchKernelGGL(k_ref_hankel_M, dim3((n_gam + 255) / 256, ng), dim3(256), 0, st, kg_max, ng,
                     gam_max, n_gam, (long)n_gam, D + W.Mh);
  EpiArgs ep = {};
  gemm_f64<kEpiStore>(2 * nk, n_gam, ng, phi, ng, D + W.Mh, n_gam, D + W.F, n_gam, ep, st);
  hipLaunchKernelGGL(k_ref_fourier_A, dim3((nk + 255) / 256, n_kap), dim3(256), 0, st, kk_max, nk,
                     kap_min, kap_max, n_kap, (long)2 * nk, D + W.Af);
  const double cfac = 1.0 / (2.0 * kPi * kPi);
  const double dkg = (kg_max - 0.0) / (ng - 1) * 1.0;
  hipLaunchKernelGGL(k_ref_em, dim3(n_kap), dim3(256), 0, st, D + W.Af, (long)2 * nk, phi, (long)ng,
                     nk, dkg, cfac, D + W.va, D + W.vb);
  ep.va = D + W.va;
  ep.vb = D + W.vb;
  ep.cfac = cfac;
  ep.gam_max = gam_max;
  ep.n_gam = n_gam;
  gemm_f64<kEpiP3>(n_kap, n_gam, 2 * nk, D + W.Af, 2 * nk, D + W.F, n_gam, p3, n_gam, ep, st);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_arcs(const double* p3, double kap_min, double kap_max, int32_t n_kap,
                        double gam_max, int32_t n_gam, double g_def, const double* xi, int32_t n_xi,
                        const double* norm, double* out, void* stream) {
  if (!p3 || !xi || !out || n_kap < 2 || n_gam < 2 || n_xi < 1 || !(kap_max > kap_min) ||
      !(gam_max > 0) || !(g_def >= 0))
    return SGL_ERR_INVALID;
  ArcGrid g;
  g.kap_min = kap_min;
  g.kap_max = kap_max;
  g.gam_max = gam_max;
  g.g_def = g_def;
  g.n_kap = n_kap;
  g.n_gam = n_gam;
  g.n_int = 2400;
  hipLaunchKernelGGL(k_ref_arcs, dim3(n_xi), dim3(256), 0, ref_stream(stream), p3, (long)n_gam, g, xi,
                     norm, out);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_gemm_f64(int32_t M, int32_t N, int32_t K, const double* A, int64_t lda,
                            const double* B, int64_t ldb, double* C, int64_t ldc, void* stream) {
  if (!A || !B || !C || M < 1 || N < 1 || K < 1 || (M + 63) / 64 > 65535 || lda < K || ldb < N ||
      ldc < N)
    return SGL_ERR_INVALID;
  EpiArgs ep = {};
  gemm_f64<kEpiStore>(M, N, K, A, (long)lda, B, (long)ldb, C, (long)ldc, ep, ref_stream(stream));
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_moments(const double* xw, const double* pw, int32_t n, double* mom,
                           void* stream) {
  if (!xw || !pw || !mom || n < 2) return SGL_ERR_INVALID;
  hipLaunchKernelGGL(k_ref_moments, dim3(1), dim3(256), 0, ref_stream(stream), xw, pw, n, mom);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_lss_bytes(void) { return ref_lss_doubles() * sizeof(double); }

sgl_status sgl_ref_lss_var(const void* tables, int32_t flags, void* ws, double* var, void* stream) {
  if (!ref_flags_ok(flags) || !tables || !ws || !var) return SGL_ERR_INVALID;
  const RefLayout L = ref_layout(flags);
  const double* T = static_cast<const double*>(tables);
  return ref_rc(launch_ref_lss(T + L.T.cosmo, T + L.T.derived, T + L.T.gdense,
                               static_cast<double*>(ws), var, ref_stream(stream)));
}

sgl_status sgl_ref_lss_psi(double var, const double* k_k, int32_t nk, const double* k_g,
                           int32_t ng, double* psi, void* stream) {
  if (!k_k || !k_g || !psi || nk < 1 || ng < 1 || !(var >= 0.0) || !std::isfinite(var))
    return SGL_ERR_INVALID;
  const long long n = (long long)nk * ng;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_ref_lss_psi, dim3(blocks), dim3(256), 0, ref_stream(stream), var, k_k, nk,
                     k_g, ng, psi);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_lss_axis(double var, const double* k, int32_t n, int32_t axis, double* out,
                            void* stream) {
  if (!k || !out || n < 1 || (axis != 0 && axis != 1) || !(var >= 0.0) || !std::isfinite(var))
    return SGL_ERR_INVALID;
  hipLaunchKernelGGL(k_ref_lss_axis, dim3((n + 255) / 256), dim3(256), 0, ref_stream(stream), var,
                     k, n, axis, out);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

}  // extern "C"
