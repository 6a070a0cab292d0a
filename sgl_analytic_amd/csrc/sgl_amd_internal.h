// Internal layout and helpers shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/sgl_amd.h"
#include "sgl_device.h"

namespace sgl {

constexpr int kMaxSlices = 96;
// Reference grid (SGL_FLAG_REF_GRID): up to 208 lens nodes ("slices"): the
// 104 midpoint shells (the 103 engine nodes below z_s = 12.34 plus the last
// shell) at two Gauss-Legendre nodes each, and the engine's 99 mass nodes.
constexpr int kRefMaxShells = 208;
constexpr int kRefMasses = 99;
constexpr int kGrowthMaxNodes = 208;   // slice nodes the growth kernel stages in LDS
constexpr int kCovMaxModes = 2048;
constexpr int kStatusCovModes = 1;      // covariance mode count clamped
constexpr int kStatusBadShape = 2;
constexpr int kStatusVpTransition = 4;  // positive closure: transition balancing not converged
constexpr int kStatusVpRate = 8;        // positive closure: a rate variance beyond the node cap

// slice record (8 doubles): z, chi, dchi, dchi (1+z)^2, D(z) (the growth of
// sigma(M, z)), Sigma_cr, E^2, Dg(z) (the growth in the bias; D(z) unless the
// legacy growth is selected)
constexpr int kSliceStride = 8;
constexpr int kSlZ = 0, kSlChi = 1, kSlDchi = 2, kSlWz = 3, kSlD = 4, kSlScr = 5, kSlE2 = 6,
              kSlDg = 7;
// mass record (4 doubles): M [Msun], sigma(M, 0), dln sigma / dln M
constexpr int kMassStride = 4;
// row record (8 doubles): k_s, r_s, x_t, r_vir, w_row, beta, eps (halo ellipticity,
// 0 without SGL_FLAG_ELL), 0
constexpr int kRowStride = 8;
// Filament rows (k0base, r_s, w_row, beta) and the filament cell lattice:
// kFilU u bins log-uniform from 1e-5 to 1 (first bin the disc u < 1e-5) x
// kFilPhi orientation nodes.
constexpr int kFilRowStride = 4;
constexpr int kFilU = 120;
constexpr int kFilPhi = 13;
constexpr int kRowKs = 0, kRowRs = 1, kRowXt = 2, kRowRvir = 3, kRowW = 4, kRowBeta = 5;

// Offsets (in doubles) of the sub-arrays of the tables buffer.
struct TablesLayout {
  size_t cosmo, derived, slices, mass, rows, frows, cov, sigtab, sigspl, gdense, m_nodes, uedge,
      knodes, status, total;
};

// Halo ellipticity (SGL_FLAG_ELL): the sigma(M) table of the M*(z) root,
// M = logspace(0, 17, kMstarN) Msun (sgl_full.py:551-567).
constexpr int kMstarN = 1200;

// Subhalos (SGL_FLAG_SUBHALOS): the reference's sigma(M, 0) spline, ln sigma on
// M = logspace(4, 17, kSigSplN) Msun with scipy's not-a-knot CubicSpline in
// ln M (sgl.py:536-538): [kSigSplN][kMassStride] window-sum records, then
// the knots x = ln M, values y = ln sigma and slopes s (kSigSplN each); and
// the growth ODE's state on every RK4 step node x_n = ln a of the growth
// kernel, (D, dD/dx) / D(a = 1) [kGrowthSteps + 1][2], for D at any z.
// ln M*(z) (sgl_full.py:551-567): sigma(M*, 0) Dg(z) = delta_c0 by linear
// interpolation of ln M against sigma, clamped to the table's ends
// (np.interp).  The reference first cuts its table to the largest strictly
// decreasing suffix, against its spline's extrapolation below 1e4 Msun: the
// reference-grid table (k_ref_mstar_table) holds that spline with the entries
// before the suffix flat at its first value; the kernel-shape table holds the
// window sums themselves, strictly decreasing over the whole table.
__device__ inline double ln_mstar(const double* __restrict__ tab, double target) {
  constexpr double kLnStep = 17.0 * 2.302585092994046 / (kMstarN - 1);
  if (target >= tab[0]) {
    // a table whose head is flat (k_ref_mstar_table) clamps to the end of that plateau
    int j = 0;
    while (j < kMstarN - 1 && tab[j + 1] == tab[0]) ++j;
    return kLnStep * j;
  }
  if (target <= tab[kMstarN - 1]) return kLnStep * (kMstarN - 1);
  int lo = 0, hi = kMstarN - 1;           // tab[lo] > target >= tab[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tab[mid] > target) lo = mid; else hi = mid;
  }
  const double f = (tab[lo] - target) / (tab[lo] - tab[hi]);
  return kLnStep * (lo + f);
}

constexpr int kSigSplN = 500;
constexpr int kSigSplDoubles = kSigSplN * 4 + 3 * kSigSplN;
constexpr int kGrowthSteps = 4096;

// The shape-only grids come first, so their offsets do not depend on B: a
// tables buffer built for any batch size serves every call that reads only
// the grids (phi packing, Filon weights).
inline TablesLayout tables_layout(const sgl_shape& s, int B) {
  TablesLayout L;
  size_t o = 0;
  auto take = [&](size_t n) { size_t r = o; o += (n + 31) & ~size_t(31); return r; };
  L.m_nodes = take((size_t)s.n_m);
  L.uedge = take((size_t)s.n_b + 1);
  L.knodes = take((size_t)s.n_k);
  L.status = take(1);
  L.cosmo = take((size_t)B * 16);
  L.derived = take((size_t)B * 16);
  L.slices = take((size_t)B * s.n_z_max * kSliceStride);
  L.mass = take((size_t)B * s.n_m * kMassStride);
  L.rows = take((size_t)B * s.n_z_max * s.n_m * kRowStride);
  L.frows = take((s.flags & SGL_FLAG_FILAMENTS) ? (size_t)B * s.n_z_max * s.n_m * kFilRowStride : 0);
  L.cov = take((size_t)B * s.n_z_max * s.n_z_max);
  L.sigtab = take((s.flags & SGL_FLAG_ELL) ? (size_t)B * kMstarN : 0);
  const bool sub = (s.flags & SGL_FLAG_SUBHALOS) != 0;
  // the reference's sigma spline: subhalos, and M* of the reference grid's ellipticity
  const bool spl = sub || ((s.flags & SGL_FLAG_REF_GRID) && (s.flags & SGL_FLAG_ELL));
  L.sigspl = take(spl ? (size_t)B * kSigSplDoubles : 0);
  // the dense growth table: subhalos, and every reference-grid build (the
  // Limber term's D(z))
  L.gdense = take((sub || (s.flags & SGL_FLAG_REF_GRID)) ? (size_t)B * (kGrowthSteps + 1) * 2 : 0);
  L.total = o;
  return L;
}

// Positive ("vp") closure workspace (doubles), per instance b and slice s:
//   eta[2B]                      calibrated variance scale eta[b], then the
//                                calibration target A^T C A at eta[B + b]
//   wk[B][NZ][G]                 sum over a group's cells of w xi
//   K[B][NZ][kVpN][kVpN]         transition slice s -> s + 1 (row = from-state)
//   tab[B][NZ][G][3]             rate of group g at the latent nodes u_j >= node
//                                bucket: lambda(u) = u minv - tm, (minv, tm, bucket)
//   ytot[B][NZ][n_k][2]          sum over the slice's groups of Y_g(k)
//   zpart[B][NZ][kVpN][n_k][2]   u_j P1_j - P0_j (the rate-weighted group sums)
// G = n_m groups (the halo mass rows), 2 n_m with filaments (then the
// filament rows).
constexpr int kVpN = 48;
__host__ __device__ inline int vp_groups(const sgl_shape& s) {
  return s.n_m * ((s.flags & SGL_FLAG_FILAMENTS) ? 2 : 1);
}
struct VpLayout {
  size_t eta, wk, K, tab, ytot, zpart, total;
};
inline VpLayout vp_layout(const sgl_shape& s, int B) {
  VpLayout L;
  size_t o = 0;
  auto take = [&](size_t n) { size_t r = o; o += (n + 31) & ~size_t(31); return r; };
  const size_t nz = (size_t)B * s.n_z_max, G = (size_t)vp_groups(s);
  L.eta = take(2 * (size_t)B);
  L.wk = take(nz * G);
  L.K = take(nz * kVpN * kVpN);
  L.tab = take(nz * G * 3);
  L.ytot = take(nz * s.n_k * 2);
  L.zpart = take(nz * kVpN * s.n_k * 2);
  L.total = o;
  return L;
}

// Fine jump lattice of the batched path: signed log bins.
//   positive side: kFinePos bins log-uniform on [kXiLo, kXiHi], bin 0 also
//   takes 0 < xi < kXiLo; negative side: kFineNeg bins log-uniform on |xi|
//   in [kXiNegLo, kXiNegHi].
constexpr int kFinePos = 1024;
constexpr int kFineNeg = 128;
constexpr int kFineBins = kFinePos + kFineNeg;
constexpr double kXiLo = 1e-12;
constexpr double kXiHi = 30.0;
constexpr double kXiNegLo = 1e-8;
constexpr double kXiNegHi = 30.0;
// Per-slice coarse lattice of the two-point amplitudes: 16 fine bins each.
constexpr int kCoarse = 16;
constexpr int kSliceBins = kFineBins / kCoarse;
constexpr int kMoments = 3;        // fine bins carry w, w xi, w xi^2 (centroid and in-bin variance)
constexpr int kRec = 4;            // compacted bin record: xi/(4 pi), xi, w, v
constexpr int kMaxMasses = 256;
constexpr int kMaxB = 1024;
// Small-phase split of the fine-record sum: records with k xi <= kTayEps
// enter through prefix sums of w xi^p, p = 2 .. 9 (kTay moments).
constexpr int kTay = 8;
constexpr double kTayEps = 0.5;
// Coarse (per-slice) records with a Gaussian in-bin spread: prefix sums of
// the raw Gaussian moments w mu_n, n = 1 .. 9, over the positive bins.
constexpr int kTayC = 9;
constexpr int kCoarsePos = kFinePos / kCoarse;

// Two-point term of the batched path on a sub-grid of the k nodes: every
// kPsi2Stride-th node (node jc is k node jc * S of the log grid; the last one
// may lie up to S - 1 node spacings beyond k_hi, on the same log grid), then
// cubic Lagrange in ln k back onto every node.  Grids too short for a
// four-point rule keep every node (S = 1).
constexpr int kPsi2Stride = SGL_PSI2_STRIDE;
__host__ __device__ inline int psi2_stride(const sgl_shape& s) {
  return (s.n_k - 1 >= 3 * kPsi2Stride) ? kPsi2Stride : 1;
}
__host__ __device__ inline int psi2_nodes(const sgl_shape& s) {
  const int S = psi2_stride(s);
  return (s.n_k - 1 + S - 1) / S + 1;
}
// k at node n of the log grid (the formula that fills the knodes array; n
// may exceed n_k - 1).
__device__ inline double knode_at(const sgl_shape& s, int n) {
  const double l0 = log10(s.k_lo), l1 = log10(s.k_hi);
  return pow(10.0, l0 + (l1 - l0) * n / (s.n_k - 1));
}

// Histogram buffer (doubles): per-slice partial fine moments, compacted fine
// records + count per instance, compacted coarse records + count per slice,
// per instance the bin -> first-record map (kFineBins + 1 entries) and the
// prefix moments of the positive records ((kFinePos + 1) x kTay), the same
// two per slice for the positive coarse records, the 80-byte RowF record of
// every (slice, mass) row, and the 32-byte record of every column.
struct HistLayout {
  size_t partial, cpart, frec, fcnt, srec, scnt, fpos, fpfx, cpos, cpfx, rowf, cols, items, total;
};
__host__ __device__ inline size_t round32(size_t n) { return (n + 31) & ~size_t(31); }
__host__ __device__ inline HistLayout hist_layout(const sgl_shape& s, int B) {
  HistLayout L;
  L.partial = 0;
  L.cpart = L.partial + round32((size_t)B * s.n_z_max * kFineBins * kMoments);
  L.frec = L.cpart + round32((size_t)B * s.n_z_max * kSliceBins * 3);
  L.fcnt = L.frec + round32((size_t)B * kFineBins * kRec);
  L.srec = L.fcnt + round32((size_t)B);
  L.scnt = L.srec + round32((size_t)B * s.n_z_max * kSliceBins * kRec);
  L.fpos = L.scnt + round32((size_t)B * s.n_z_max);
  L.fpfx = L.fpos + round32((size_t)B * (kFineBins + 1));
  L.cpos = L.fpfx + round32((size_t)B * (kFinePos + 1) * kTay);
  L.cpfx = L.cpos + round32((size_t)B * s.n_z_max * (kCoarsePos + 1));
  L.rowf = L.cpfx + round32((size_t)B * s.n_z_max * (kCoarsePos + 1) * kTayC);
  L.cols = L.rowf + round32((size_t)B * s.n_z_max * s.n_m * (sizeof(RowF) / sizeof(double)));
  // work queue of the persistent binning schedule (ints): [0] next item,
  // [1] item count, [2 ..] one (instance, slice) item per slice
  L.items = L.cols + round32((size_t)s.n_b * 4);
  L.total = L.items + round32(((size_t)B * s.n_z_max + 2 + 1) / 2);
  return L;
}

// Launch-shape knobs set by the autotuner (sgl_set_tuning); host-side, read
// by the launchers at launch time.
extern int g_sgl_tune[SGL_TUNE_COUNT];

int tables_build(const double* cosmo, int B, const sgl_shape& sh, const double* m_nodes_h,
                 double* derived, double* slices, double* mass, double* rows, double* frows,
                 double* cov, double* sigtab, int* status, hipStream_t st);
int ref_tables_build(const double* cosmo, int B, const sgl_shape& sh, const double* m_nodes,
                     double* derived, double* slices, double* mass, double* rows, double* frows,
                     double* sigtab, double* sigspl, double* gdense, hipStream_t st);
size_t ref_lss_doubles();
int launch_ref_lss(const double* cosmo, const double* derived, const double* gdense, double* ws,
                   double* var, hipStream_t st);
int launch_fill_grids(const sgl_shape& sh, double* ue, double* kn, double* m_nodes_h,
                      hipStream_t st);
int launch_lnphi_exact(const double* rows, const double* frows, const double* cosmo, const double* uedge,
                       const double* knodes, const sgl_shape& sh, int B, double* lam, double* amp,
                       double* work, hipStream_t st);
int launch_bin(const double* rows, const double* frows, const double* cosmo, const double* uedge, const sgl_shape& sh,
               int B, double* hist, hipStream_t st);
int launch_bin_pass(const double* rows, const double* frows, const double* cosmo, const double* uedge, const sgl_shape& sh,
               int B, double* hist, hipStream_t st);
int launch_bin_reduce(const double* rows, const double* cosmo, const double* uedge, const sgl_shape& sh,
               int B, double* hist, hipStream_t st);
int launch_lnphi_binned(const double* hist, const double* cosmo, const double* knodes,
                        const sgl_shape& sh, int B, double* lam, float* amp, hipStream_t st);
int launch_quadform(const double* cov, const double* cosmo, const void* amp, int B,
                    const sgl_shape& sh, int prec, double* psi2, hipStream_t st);
int launch_vp_cells(const double* rows, const double* frows, const double* cosmo,
                    const double* uedge, const double* knodes, const sgl_shape& sh, int B,
                    const VpLayout& V, double* vp, double* work, int pass, hipStream_t st);
int launch_vp_setup(const double* rows, const double* frows, const double* cov,
                    const double* cosmo, const sgl_shape& sh, int B, const VpLayout& V, double* vp,
                    int* status, hipStream_t st);
int launch_vp_chain(const double* cosmo, const sgl_shape& sh, int B, const VpLayout& V,
                    const double* vp, double* lnr, hipStream_t st);
int launch_sum_slices(const double* work, const sgl_shape& sh, int B, double* lam, hipStream_t st);
int launch_trunc_nfw(const double* x, int n, double ks, double xt, int prec, double* kap,
                     double* gam, hipStream_t st);
int launch_psi2_interp(const double* psi2c, const sgl_shape& sh, int B, double* psi2,
                       hipStream_t st);
int launch_phi_pack(const double* lam, const double* psi2, const double* knodes,
                    const sgl_shape& sh, int B, int prec, void* out, hipStream_t st);
int launch_filon(const double* knodes, const sgl_shape& sh, int prec, void* out, hipStream_t st);
int launch_invert_fp64(const double* phi, const double* knodes, const sgl_shape& sh, int B,
                       const double* t, int n_t, double* p, hipStream_t st);
int launch_gemm(const void* E, const void* Phi, const sgl_shape& sh, int B, int prec, float* P,
                hipStream_t st);
int launch_diag(const void* p, int is_double, const sgl_shape& sh, int B, double* diag,
                hipStream_t st);
inline int k_padded_host(const sgl_shape& sh) { return ((2 * (sh.n_k + 1) + 63) / 64) * 64; }

}  // namespace sgl
