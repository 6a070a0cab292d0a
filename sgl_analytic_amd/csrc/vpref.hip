// Positive ("vp") clustering closure on the reference-shape path: the
// reference's xi_closure="vp" (pipeline.py:154-207, pos_closure.install,
// pos_closure.py:708-733), which replaces the Gaussian pair term Psi_2 by
//   ln R = ln E_U exp( sum_s sum_c W_sc (lam_sc(U_s) - 1) F_c ),
//   F_c = e^{i k_kappa kappa_c} J0(k_gamma gamma_c) - 1,
// minus shear_var k_gamma^2 / 4, U_s a stationary chain of 48 Gauss-Hermite
// states along the line of sight.
//
//   cells       the per-shell fine-lattice bins (pos_pieces,
//               sgl_full.py:1693-1699, 1800-1802): the emitted cells sorted by
//               (shell, fine bin) (stable), summed per run in fixed parts;
//               each takes the pooled centroid of its bin (the one Lambda
//               uses) and, by its own centroid, a coarse parent whose
//               loading phi it inherits (pos_closure.population,
//               pos_closure.py:388-421);
//   groups      cells of one shell with one coarse parent (equal loading:
//               pos_closure.py:465-468 groups by the loading's value; cells
//               without a loading form the shell's zero group, whose rates
//               are 1);
//   rho, f      top-hat line-of-sight correlation of consecutive shells on
//               30001 k_par nodes, weighted by the leading k_perp profiles
//               (los_rho, pos_closure.py:372-385, 422-437), f_s = (1 - rho_bar)
//               / (1 + rho_bar);
//   eta         variance(eta) of the chain's kappa clustering (pos_closure.py:
//               560-591) evaluated here, the root found by the host;
//   rates       vp rates (pos_closure.py:273-308) in interval form
//               (pos_closure_fused.py:18-49): D_j = c + a (u_j - u_anchor) + b
//               on the active states, c elsewhere;
//   ln R        per shell and point tile: the group sums Y_g streamed over
//               the groups ordered by (linear, left anchor ascending, right
//               anchor descending), the state values built by the reference's
//               prefix rule (pos_closure_fused.py:52-104), then the log-scaled
//               forward recursion per point (pos_closure.py:183-240);
//   qxi         Im ln R unwrapped along k_kappa at the first k_gamma, then
//               along k_gamma (numpy.unwrap), minus the shear term
//               (pos_closure.py:690-705);
//   spline      the not-a-knot bicubic interpolant of qxi on the coarse grid
//               k = k_max s^2 (nk x n nodes, sgl_full.py:226-238, 2953-3022)
//               on the full grid, as one-dimensional not-a-knot cubics along
//               k_kappa then k_gamma (the same interpolant);
//   window      the positive-definite window on Phi when the reach did not
//               converge (sgl_full.py:83-94, 328-333).
#include <algorithm>
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "ref_xi.h"
#include "sgl_device.h"

namespace sgl {
namespace {

// The positive closure runs on the midpoint shells only (one Limber node
// per shell): at most 104, the 103 engine nodes below z_s plus the last shell.
constexpr int kVpShells = 104;
static_assert(kVpShells <= kRefMaxShells, "vp shells within the table capacity");
constexpr int kNs = 48;                                  // latent states
constexpr int kFine = 961 * 300;                         // fine (kappa, gamma) bins
constexpr int kCkeyBits = 25;
constexpr int kCkeyInvalid = (1 << kCkeyBits) - 1;       // outside the lattice: sorts last
static_assert(kVpShells * kFine < kCkeyInvalid, "composite keys exceed the sort bits");
constexpr int kGroupSlots = kXiBins + 1;                 // coarse parents + the zero group
constexpr int kGkeyBits = 22;
static_assert(kVpShells * kGroupSlots < (1 << kGkeyBits), "group keys exceed the sort bits");
constexpr int kOkeyBits = 14;                            // (shell, type, anchor) visiting order
static_assert(kVpShells * 3 * kNs < (1 << kOkeyBits), "order keys exceed the sort bits");
constexpr int kLosN = 30001;                             // los_rho k_par nodes on [0, 3]
constexpr double kLosKmax = 3.0;
constexpr int kPart = 4096;                              // cells per fixed summation part
constexpr double kPiV = 3.141592653589793;

struct VKGW {
  double k, g, w;
};

inline size_t al256(size_t n) { return (n + 255) & ~size_t(255); }
inline unsigned nblk(long long n, int b = 256) { return (unsigned)((n + b - 1) / b); }
inline hipStream_t vst(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline bool hip_ok() { return hipGetLastError() == hipSuccess; }

// ------------------------------------------------------ run workspace ----
// The emitted cells keyed by (shell, fine bin), stably sorted, their runs and
// per-run sums (w, w kappa, w gamma) in fixed parts.
struct RunWs {
  size_t ckey, skey, vin, vout, ukey, rlen, nruns, rstart, nparts, poff, psum, rsum, tmp, total;
  size_t tmp_bytes;
  long long cap, max_parts;
};
RunWs run_ws(long long n) {
  RunWs L;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += al256(b); return r; };
  L.cap = std::min<long long>(n, (long long)kVpShells * kFine) + 1;
  L.max_parts = L.cap + n / kPart + 1;
  size_t sort_b = 0, rle_b = 0, scan_b = 0;
  (void)rocprim::radix_sort_pairs((void*)nullptr, sort_b, (int*)nullptr, (int*)nullptr,
                                  (VKGW*)nullptr, (VKGW*)nullptr, (size_t)n, 0, kCkeyBits);
  (void)rocprim::run_length_encode((void*)nullptr, rle_b, (int*)nullptr, (size_t)n, (int*)nullptr,
                                   (int*)nullptr, (int*)nullptr);
  (void)rocprim::exclusive_scan((void*)nullptr, scan_b, (int*)nullptr, (long long*)nullptr, 0LL,
                                (size_t)L.cap, rocprim::plus<long long>());
  L.tmp_bytes = std::max(sort_b, std::max(rle_b, scan_b));
  L.ckey = take((size_t)n * sizeof(int));
  L.skey = take((size_t)n * sizeof(int));
  L.vin = take((size_t)n * sizeof(VKGW));
  L.vout = take((size_t)n * sizeof(VKGW));
  L.ukey = take((size_t)L.cap * sizeof(int));
  L.rlen = take((size_t)L.cap * sizeof(int));
  L.nruns = take(64 * sizeof(int));
  L.rstart = take((size_t)L.cap * sizeof(long long));
  L.nparts = take((size_t)L.cap * sizeof(int));
  L.poff = take((size_t)L.cap * sizeof(long long));
  L.psum = take((size_t)L.max_parts * 3 * sizeof(double));
  L.rsum = take((size_t)L.cap * 3 * sizeof(double));
  L.tmp = take(L.tmp_bytes);
  L.total = o;
  return L;
}

// ------------------------------------------------------ closure state ----
// Per PDF, sized by the cell count n (groups <= n + shells).
struct VpState {
  size_t meta, cu, cw, ckap, cphi, gkey_in, gkey_out, idx_in, idx_out, gkey, glen, gnum,
      gstart, gshell, gphi, gwk, gminv, gtm, gbk, gtype, ganch, ga, gb, okey_in, okey_out,
      gord_in, gord_out, vlen, vstart, vcu, vcw, vlast, vgrp, bmap, bocc, bocc64, shg, f, rho, K,
      ms, P0, los, tmp, total;
  size_t tmp_bytes;
  long long n, gcap;
};
// meta (int64): [0] cells, [1] groups, [2] shells, [3] groups visited (nonzero
// rates), [4] cells visited, [5] status (1: bins disagree with the pooled
// lattice, 2: a shell without a kappa mode, 4: a rate beyond the node cap)
VpState vp_state(long long n) {
  VpState L;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += al256(b); return r; };
  const long long nn = std::max<long long>(n, 1);
  L.n = nn;
  L.gcap = nn + kVpShells + 1;
  size_t sort_b = 0, sort2_b = 0, rle_b = 0, scan_b = 0;
  (void)rocprim::radix_sort_pairs((void*)nullptr, sort_b, (int*)nullptr, (int*)nullptr,
                                  (int*)nullptr, (int*)nullptr, (size_t)nn, 0, kGkeyBits);
  (void)rocprim::radix_sort_pairs((void*)nullptr, sort2_b, (int*)nullptr, (int*)nullptr,
                                  (int*)nullptr, (int*)nullptr, (size_t)L.gcap, 0, kOkeyBits);
  (void)rocprim::run_length_encode((void*)nullptr, rle_b, (int*)nullptr, (size_t)nn,
                                   (int*)nullptr, (int*)nullptr, (int*)nullptr);
  size_t scan2_b = 0, scan3_b = 0;
  (void)rocprim::exclusive_scan((void*)nullptr, scan_b, (int*)nullptr, (long long*)nullptr, 0LL,
                                (size_t)L.gcap, rocprim::plus<long long>());
  (void)rocprim::exclusive_scan((void*)nullptr, scan2_b, (long long*)nullptr,
                                (long long*)nullptr, 0LL, (size_t)kFine,
                                rocprim::plus<long long>());
  (void)rocprim::exclusive_scan((void*)nullptr, scan3_b, (long long*)nullptr,
                                (long long*)nullptr, 0LL, (size_t)L.gcap,
                                rocprim::plus<long long>());
  L.tmp_bytes = std::max(std::max(std::max(sort_b, sort2_b), std::max(rle_b, scan_b)),
                         std::max(scan2_b, scan3_b));
  L.meta = take(16 * sizeof(long long));
  L.cu = take(nn * sizeof(int));
  L.cw = take(nn * sizeof(double));
  L.ckap = take(nn * sizeof(double));
  L.cphi = take(nn * sizeof(double));
  L.gkey_in = take(nn * sizeof(int));
  L.gkey_out = take(nn * sizeof(int));
  L.idx_in = take(nn * sizeof(int));
  L.idx_out = take(nn * sizeof(int));
  L.gkey = take(L.gcap * sizeof(int));
  L.glen = take(L.gcap * sizeof(int));
  L.gnum = take(64 * sizeof(int));
  L.gstart = take(L.gcap * sizeof(long long));
  L.gshell = take(L.gcap * sizeof(int));
  L.gphi = take(L.gcap * sizeof(double));
  L.gwk = take(L.gcap * sizeof(double));
  L.gminv = take(L.gcap * sizeof(double));
  L.gtm = take(L.gcap * sizeof(double));
  L.gbk = take(L.gcap * sizeof(int));
  L.gtype = take(L.gcap * sizeof(int));
  L.ganch = take(L.gcap * sizeof(int));
  L.ga = take(L.gcap * sizeof(double));
  L.gb = take(L.gcap * sizeof(double));
  L.okey_in = take(L.gcap * sizeof(int));
  L.okey_out = take(L.gcap * sizeof(int));
  L.gord_in = take(L.gcap * sizeof(int));
  L.gord_out = take(L.gcap * sizeof(int));
  L.vlen = take(L.gcap * sizeof(long long));
  L.vstart = take(L.gcap * sizeof(long long));
  L.vcu = take(nn * sizeof(int));
  L.vcw = take(nn * sizeof(double));
  L.vlast = take(nn * sizeof(int));
  L.vgrp = take(nn * sizeof(int));
  L.bmap = take((size_t)kFine * sizeof(long long));
  L.bocc = take((size_t)kFine * sizeof(int));
  L.bocc64 = take((size_t)kFine * sizeof(long long));
  L.shg = take((size_t)(kVpShells + 1) * 2 * sizeof(long long));
  L.f = take(kVpShells * sizeof(double));
  L.rho = take(kVpShells * sizeof(double));
  L.K = take((size_t)kVpShells * kNs * kNs * sizeof(double));
  L.ms = take((size_t)kVpShells * kNs * sizeof(double));
  L.P0 = take((size_t)kXiNk * kLosN * sizeof(double));
  L.los = take((size_t)kVpShells * kXiNk * sizeof(double));
  L.tmp = take(L.tmp_bytes);
  L.total = o;
  return L;
}

template <typename T>
__device__ __host__ inline T* at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(base) + off);
}

__device__ inline int first_greater(const double* e, int n, double v) {
  int lo = 0, hi = n;                    // first index with e[i] > v (searchsorted right)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (e[mid] > v) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// ---------------------------------------------------------- cell runs ----
__global__ void k_vp_ckey(long long n, const long long* __restrict__ offsets, int tasks_per_shell,
                          int n_z, const int* __restrict__ key, const double* __restrict__ K,
                          const double* __restrict__ G, const double* __restrict__ W,
                          int* __restrict__ ckey, VKGW* __restrict__ v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = 0, hi = n_z - 1;              // last shell whose first cell is <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (offsets[(size_t)mid * tasks_per_shell] <= i) lo = mid; else hi = mid - 1;
  }
  const int k = key[i];
  ckey[i] = (k >= 0 && k < kFine) ? lo * kFine + k : kCkeyInvalid;
  v[i] = VKGW{K[i], G[i], W[i]};
}

__global__ void k_vp_zero_tail(int* __restrict__ rlen, const int* __restrict__ nruns,
                               long long cap) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < cap && r >= *nruns) rlen[r] = 0;
}

__global__ void k_vp_nparts(const int* __restrict__ rlen, const int* __restrict__ nruns,
                            long long cap, int* __restrict__ nparts) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= cap) return;
  nparts[r] = (r < *nruns) ? (rlen[r] + kPart - 1) / kPart : 0;
}

// one wave per part: lane-strided sums, then a fixed shuffle tree
__global__ __launch_bounds__(256) void k_vp_parts(const int* __restrict__ rlen,
                                                  const int* __restrict__ nruns,
                                                  const long long* __restrict__ rstart,
                                                  const int* __restrict__ nparts,
                                                  const long long* __restrict__ poff,
                                                  const VKGW* __restrict__ v,
                                                  double* __restrict__ psum) {
  const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  const int nr = *nruns;
  if (nr < 1) return;
  const long long total = poff[nr - 1] + nparts[nr - 1];
  if (wv >= total) return;
  int lo = 0, hi = nr - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (poff[mid] <= wv) lo = mid; else hi = mid - 1;
  }
  const long long p = wv - poff[lo];
  const long long a = rstart[lo] + p * kPart;
  const long long b = min(rstart[lo] + (long long)rlen[lo], a + kPart);
  double sw = 0.0, sk = 0.0, sg = 0.0;
  for (long long q = a + lane; q < b; q += 64) {
    const VKGW c = v[q];
    sw += c.w;
    sk += c.w * c.k;
    sg += c.w * c.g;
  }
  for (int off = 32; off > 0; off >>= 1) {
    sw += __shfl_down(sw, off);
    sk += __shfl_down(sk, off);
    sg += __shfl_down(sg, off);
  }
  if (lane == 0) {
    psum[wv * 3] = sw;
    psum[wv * 3 + 1] = sk;
    psum[wv * 3 + 2] = sg;
  }
}

__global__ void k_vp_combine(const int* __restrict__ ukey, const int* __restrict__ nruns,
                             const int* __restrict__ nparts, const long long* __restrict__ poff,
                             const double* __restrict__ psum, double* __restrict__ rsum,
                             long long* __restrict__ n_valid) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int nr = *nruns;
  if (r == 0) *n_valid = (nr > 0 && ukey[nr - 1] == kCkeyInvalid) ? nr - 1 : nr;
  if (r >= nr) return;
  double sw = 0.0, sk = 0.0, sg = 0.0;
  for (long long q = poff[r]; q < poff[r] + nparts[r]; ++q) {
    sw += psum[q * 3];
    sk += psum[q * 3 + 1];
    sg += psum[q * 3 + 2];
  }
  rsum[(size_t)r * 3] = sw;
  rsum[(size_t)r * 3 + 1] = sk;
  rsum[(size_t)r * 3 + 2] = sg;
}

// ------------------------------------------------------------- cells ----
// occupied bins over every shell (any run with weight), for the pooled index
__global__ void k_vp_occ(long long n, const int* __restrict__ ukey, const double* __restrict__ rsum,
                         int* __restrict__ occ) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (rsum[(size_t)c * 3] != 0.0) occ[ukey[c] % kFine] = 1;
}

__global__ void k_vp_int_to_ll(const int* __restrict__ a, long long* __restrict__ b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = a[i];
}

// the pooled lattice has one entry per occupied bin: the counts must agree
__global__ void k_vp_check_occ(const long long* __restrict__ bmap, const int* __restrict__ occ,
                               const int* __restrict__ npop, long long* __restrict__ meta) {
  if (threadIdx.x == 0 && blockIdx.x == 0)
    if (bmap[kFine - 1] + occ[kFine - 1] != (long long)*npop) meta[5] |= 1;
}

// per cell: pooled index and centroid, its own coarse parent and the
// parent's loading, its group key
__global__ void k_vp_cells(long long n, const int* __restrict__ ukey,
                           const double* __restrict__ rsum, const long long* __restrict__ bmap,
                           const double* __restrict__ pk, const double* __restrict__ xs,
                           const int* __restrict__ xukey, const double* __restrict__ phic,
                           int* __restrict__ cu, double* __restrict__ cw,
                           double* __restrict__ ckap, double* __restrict__ cphi,
                           int* __restrict__ gkey, int* __restrict__ idx) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const XiState L = xi_state_layout();
  const int* meta = reinterpret_cast<const int*>(xs + L.meta);
  const int key = ukey[c], s = key / kFine, bin = key - s * kFine;
  const double w = rsum[(size_t)c * 3], wk = rsum[(size_t)c * 3 + 1], wg = rsum[(size_t)c * 3 + 2];
  const int u = (int)bmap[bin];
  cu[c] = u;
  cw[c] = w;
  ckap[c] = pk[u];
  const double kc = wk / w, gc = wg / w;
  const int a = first_greater(xs + L.ekc, kXiKc + 1, kc) - 1;
  const int b = first_greater(xs + L.egc, kXiGc + 1, gc) - 1;
  double phi = 0.0;
  int slot = kXiBins;
  if (a >= 0 && a < kXiKc && b >= 0 && b < kXiGc) {
    const int cb = a * kXiGc + b, want = s * kXiBins + cb;
    int lo = meta[kXmRfirst + s], hi = meta[kXmRfirst + s + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (xukey[mid] < want) lo = mid + 1; else hi = mid;
    }
    if (lo < meta[kXmRfirst + s + 1] && xukey[lo] == want) phi = phic[lo];
    if (phi != 0.0) slot = cb;
  }
  cphi[c] = phi;
  // a bin whose sums are all zero is not a cell of the shell (the pieces keep
  // h0, h0k or h0g != 0): its key sorts after every shell
  const bool absent = w == 0.0 && wk == 0.0 && wg == 0.0;
  gkey[c] = absent ? kVpShells * kGroupSlots : s * kGroupSlots + slot;
  idx[c] = (int)c;
}

// per group: shell, loading, wk = sum W kappa (pooled centroid) in cell order
__global__ void k_vp_groups(const int* __restrict__ gkey, const int* __restrict__ glen,
                            const int* __restrict__ gnum, const long long* __restrict__ gstart,
                            const int* __restrict__ idx, const double* __restrict__ cw,
                            const double* __restrict__ ckap, const double* __restrict__ cphi,
                            int* __restrict__ gshell, double* __restrict__ gphi,
                            double* __restrict__ gwk) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= *gnum) return;
  const long long a = gstart[g], b = a + glen[g];
  double wk = 0.0;
  for (long long q = a; q < b; ++q) {
    const int c = idx[q];
    wk += cw[c] * ckap[c];
  }
  gshell[g] = gkey[g] / kGroupSlots;
  gphi[g] = (gkey[g] % kGroupSlots == kXiBins || gkey[g] >= kVpShells * kGroupSlots)
                ? 0.0
                : cphi[idx[a]];
  gwk[g] = wk;
}

// groups of every shell: shg[s] = first group of shell s, shg[n_z] = the
// groups and (meta) cells of the shells (the absent bins' group excluded)
__global__ void k_vp_shell_groups(const int* __restrict__ gshell, const int* __restrict__ gnum,
                                  const long long* __restrict__ gstart, int n_z,
                                  long long* __restrict__ shg, long long* __restrict__ meta) {
  const int s = threadIdx.x;
  const int ng = *gnum;
  if (s > n_z) return;
  int lo = 0, hi = ng;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (gshell[mid] < s) lo = mid + 1; else hi = mid;
  }
  shg[s] = lo;
  if (s == n_z) {
    meta[0] = gstart[lo];
    meta[1] = lo;
    meta[2] = n_z;
  }
}

// ---------------------------------------------------------- rho, f, K ----
// P0(sqrt(kp_j^2 + kpar_i^2)) on the los_rho grid (growth-free, sgl_full.py:2428-2432)
__global__ void k_vp_p0(const double* __restrict__ cosmo, const double* __restrict__ derived,
                        const double* __restrict__ xs, double* __restrict__ P0) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)kXiNk * kLosN) return;
  const XiState L = xi_state_layout();
  const int j = (int)(e / kLosN), i = (int)(e - (long long)j * kLosN);
  const Cosmo c = reinterpret_cast<const Cosmo*>(cosmo)[0];
  const Derived d = reinterpret_cast<const Derived*>(derived)[0];
  const double kpar = (i == kLosN - 1) ? kLosKmax : i * (kLosKmax / (kLosN - 1));
  const double kp = xs[L.kp + j];
  const double k = sqrt(kp * kp + kpar * kpar), lk = log(k);
  const double T = eh98_T(eh98_const(c), k, lk);
  P0[e] = d.Anorm * d.Anorm * exp(c.ns * lk) * T * T;
}

__device__ inline double np_sinc(double x) {
  const double y = kPiV * ((x == 0.0) ? 1.0e-20 : x);
  return sin(y) / y;
}

// los_rho of the shell pair (s, s + 1) at k_perp node j: one workgroup,
// trapezoids over k_par with a fixed-order block sum
__global__ __launch_bounds__(256) void k_vp_los(const double* __restrict__ slices,
                                                const double* __restrict__ P0,
                                                double* __restrict__ los) {
  __shared__ double red[3][256];
  const int s = blockIdx.x, j = blockIdx.y, t = threadIdx.x;
  const double* S1 = slices + (size_t)s * kSliceStride;
  const double* S2 = slices + (size_t)(s + 1) * kSliceStride;
  const double d1 = S1[kSlDchi], d2 = S2[kSlDchi], dc = S2[kSlChi] - S1[kSlChi];
  const double* P = P0 + (size_t)j * kLosN;
  double a12 = 0.0, a11 = 0.0, a22 = 0.0;
  auto kp = [](int i) { return (i == kLosN - 1) ? kLosKmax : i * (kLosKmax / (kLosN - 1)); };
  for (int i = t; i < kLosN - 1; i += 256) {
    const double x0 = kp(i), x1 = kp(i + 1), h = 0.5 * (x1 - x0);
    const double s10 = np_sinc(x0 * d1 / (2 * kPiV)), s11 = np_sinc(x1 * d1 / (2 * kPiV));
    const double s20 = np_sinc(x0 * d2 / (2 * kPiV)), s21 = np_sinc(x1 * d2 / (2 * kPiV));
    const double p0 = P[i], p1 = P[i + 1];
    a12 += h * (p0 * (s10 * s20 * cos(x0 * dc)) + p1 * (s11 * s21 * cos(x1 * dc)));
    a11 += h * (p0 * (s10 * s10) + p1 * (s11 * s11));
    a22 += h * (p0 * (s20 * s20) + p1 * (s21 * s21));
  }
  red[0][t] = a12;
  red[1][t] = a11;
  red[2][t] = a22;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] += red[0][t + w];
      red[1][t] += red[1][t + w];
      red[2][t] += red[2][t + w];
    }
    __syncthreads();
  }
  if (t == 0) los[(size_t)s * kXiNk + j] = red[0][0] / sqrt(red[1][0] * red[2][0]);
}

// rho of consecutive shells, weighted by |v0|; f = (1 - rho_bar) / (1 + rho_bar);
// a shell whose kappa mode is empty is flagged (status 2)
__global__ void k_vp_rho(const double* __restrict__ los, const double* __restrict__ v0, int n_z,
                         double* __restrict__ rho, double* __restrict__ f,
                         long long* __restrict__ meta) {
  __shared__ double r[kVpShells];
  const int t = threadIdx.x;
  for (int s = t; s < n_z - 1; s += blockDim.x) {
    const double* a = v0 + (size_t)s * kXiNk;
    const double* b = v0 + (size_t)(s + 1) * kXiNk;
    double num = 0.0, na = 0.0, nb = 0.0;
    for (int j = 0; j < kXiNk; ++j) {
      num += a[j] * b[j] * los[(size_t)s * kXiNk + j];
      na += a[j] * a[j];
      nb += b[j] * b[j];
    }
    r[s] = num / sqrt(na * nb);
    rho[s] = r[s];
  }
  for (int s = t; s < n_z; s += blockDim.x) {
    double nrm = 0.0;
    for (int j = 0; j < kXiNk; ++j) nrm += v0[(size_t)s * kXiNk + j] * v0[(size_t)s * kXiNk + j];
    if (!(nrm > 0.0)) atomicOr(reinterpret_cast<unsigned long long*>(meta + 5), 2ull);
  }
  __syncthreads();
  for (int s = t; s < n_z; s += blockDim.x) {
    double sum = 0.0;
    int cnt = 0;
    if (s >= 1) { sum += r[s - 1]; ++cnt; }
    if (s + 1 < n_z) { sum += r[s]; ++cnt; }
    const double rb = cnt ? sum / cnt : 0.0;
    f[s] = (1.0 - rb) / (1.0 + rb);
  }
}

__device__ __forceinline__ double wave_max_v(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// stationary reversible transition of correlation rho (pos_closure.py:251-267)
__global__ __launch_bounds__(64) void k_vp_transition(const double* __restrict__ rho,
                                                      double* __restrict__ K,
                                                      long long* __restrict__ meta) {
  __shared__ double J[kNs][kNs + 1];
  __shared__ double v[kNs];
  const int s = blockIdx.x, i = threadIdx.x;
  const double r = rho[s];
  const double den = 1.0 / (2.0 * (1.0 - r * r));
  const bool own = i < kNs;
  const double ui = own ? kGhU[i] : 0.0, wi = own ? kGhW[i] : 1.0;
  if (own) {
    for (int j = 0; j < kNs; ++j) {
      const double uj = kGhU[j];
      J[i][j] = wi * kGhW[j] * exp((2.0 * r * ui * uj - r * r * (ui * ui + uj * uj)) * den);
    }
    v[i] = 1.0;
  }
  double err = 1.0;
  for (int it = 0; it < 10000; ++it) {
    __syncthreads();
    double marg = 1.0;
    if (own) {
      double jv = 0.0;
      for (int j = 0; j < kNs; ++j) jv = fma(J[i][j], v[j], jv);
      marg = v[i] * jv;
    }
    err = wave_max_v(own ? fabs(marg / wi - 1.0) : 0.0);
    if (err < 3e-14) break;
    __syncthreads();
    if (own) v[i] *= sqrt(wi / marg);
  }
  if (i == 0 && err > 1e-10) atomicOr(reinterpret_cast<unsigned long long*>(meta + 5), 8ull);
  __syncthreads();
  if (own) {
    double* Ks = K + ((size_t)s * kNs + i) * kNs;
    const double vi = v[i] / wi;
    for (int j = 0; j < kNs; ++j) Ks[j] = vi * J[i][j] * v[j];
  }
}

// ---------------------------------------------------------- rates ----
// suffix sums of the nodes (A_j = sum_{i>=j} w u, B_j = sum w, C_j = sum w u^2,
// from the top down) and the relative variance rv_j at t = u_j
struct GhTab {
  double u[kNs], w[kNs], A[kNs], B[kNs], C[kNs], rv[kNs - 1];
};
__device__ inline void gh_tab(GhTab& S) {
  const int t = threadIdx.x;
  if (t < kNs) {
    S.u[t] = kGhU[t];
    S.w[t] = kGhW[t];
  }
  __syncthreads();
  if (t == 0) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = kNs - 1; j >= 0; --j) {
      a += S.w[j] * S.u[j];
      b += S.w[j];
      c += S.w[j] * S.u[j] * S.u[j];
      S.A[j] = a; S.B[j] = b; S.C[j] = c;
    }
  }
  __syncthreads();
  if (t < kNs - 1) {
    const double tb = S.u[t];
    const double E1 = S.A[t + 1] - tb * S.B[t + 1];
    const double E2 = S.C[t + 1] - 2.0 * tb * S.A[t + 1] + tb * tb * S.B[t + 1];
    S.rv[t] = E2 / (E1 * E1) - 1.0;
  }
  __syncthreads();
}
// rate of relative variance v for a positive loading: lam(u_j) = u_j minv - tm
// for j >= bk (bk = 0: the linear rate, minv = sqrt(v), tm = -1); false beyond
// the cap (pos_closure.py:273-308)
__device__ inline bool gh_rate(const GhTab& S, double v, double& minv, double& tm, int& bk) {
  if (!(v < S.rv[kNs - 2])) {
    minv = 0.0; tm = -1.0; bk = 0;
    return false;
  }
  int lo = 0, hi = kNs - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (S.rv[mid] <= v) lo = mid + 1; else hi = mid;
  }
  bk = lo;
  if (lo == 0) {
    minv = sqrt(v);
    tm = -1.0;
    return true;
  }
  const double a = S.A[lo], b = S.B[lo], c = S.C[lo];
  const double q2 = b - (1.0 + v) * b * b;
  const double q1 = 2.0 * a * ((1.0 + v) * b - 1.0);
  const double q0 = c - (1.0 + v) * a * a;
  const double lu = S.u[lo - 1], hu = S.u[lo];
  double t;
  if (q2 != 0.0) {
    const double disc = sqrt(fmax(q1 * q1 - 4.0 * q2 * q0, 0.0));
    const double r1 = (-q1 - disc) / (2.0 * q2), r2 = (-q1 + disc) / (2.0 * q2);
    t = (r1 >= lu && r1 < hu) ? r1 : r2;
  } else {
    t = -q0 / q1;
  }
  const double m = a - t * b;
  minv = 1.0 / m;
  tm = t / m;
  return true;
}

__global__ __launch_bounds__(256) void k_vp_grates(const int* __restrict__ gnum,
                                                   const int* __restrict__ gshell,
                                                   const double* __restrict__ gphi,
                                                   const double* __restrict__ f, double eta,
                                                   double* __restrict__ gminv,
                                                   double* __restrict__ gtm, int* __restrict__ gbk,
                                                   long long* __restrict__ meta) {
  __shared__ GhTab S;
  gh_tab(S);
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= *gnum) return;
  const double ph = gphi[g];
  double minv, tm;
  int bk;
  if (!gh_rate(S, eta * f[gshell[g]] * ph * ph, minv, tm, bk))
    atomicOr(reinterpret_cast<unsigned long long*>(meta + 5), 4ull);
  gminv[g] = minv;
  gtm[g] = tm;
  gbk[g] = (ph < 0.0) ? -1 - bk : bk;    // a negative loading mirrors the states
}

// m_s(u_j) = sum_g lam_g(u_j) wk_g, one wave per shell, lane j (groups in order)
__global__ __launch_bounds__(64) void k_vp_ms(const long long* __restrict__ shg,
                                              const double* __restrict__ gwk,
                                              const double* __restrict__ gminv,
                                              const double* __restrict__ gtm,
                                              const int* __restrict__ gbk,
                                              double* __restrict__ ms) {
  const int s = blockIdx.x, j = threadIdx.x;
  if (j >= kNs) return;
  const long long g0 = shg[s], g1 = shg[s + 1];
  double acc = 0.0;
  for (long long g = g0; g < g1; ++g) {
    int bk = gbk[g];
    int jj = j;
    if (bk < 0) {
      bk = -1 - bk;
      jj = kNs - 1 - j;
    }
    const double lam = (jj >= bk) ? kGhU[jj] * gminv[g] - gtm[g] : 0.0;
    acc += lam * gwk[g];
  }
  ms[(size_t)s * kNs + j] = acc;
}

// chain variance of sum_s m_s(U_s) (pos_closure.py:352-369), one wave
__global__ __launch_bounds__(64) void k_vp_chainvar(const double* __restrict__ ms,
                                                    const double* __restrict__ K, int n_z,
                                                    double* __restrict__ out) {
  __shared__ double h[2][kNs];
  const int t = threadIdx.x;
  const bool own = t < kNs;
  const double w = own ? kGhW[t] : 0.0;
  double acc = 0.0;
  for (int s = 0; s < n_z; ++s) {
    const int cur = s & 1;
    const double m = own ? ms[(size_t)s * kNs + t] : 0.0;
    double mean = w * m;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mean += __shfl_xor(mean, o);
    const double mu = m - mean;
    acc += w * mu * mu;
    double hp = 0.0;
    if (s > 0 && own) {
      const double* Ks = K + (size_t)(s - 1) * kNs * kNs;
      for (int i = 0; i < kNs; ++i) hp = fma(Ks[(size_t)i * kNs + t], h[cur ^ 1][i], hp);
      acc += 2.0 * mu * hp;
    }
    if (own) h[cur][t] = hp + w * mu;
    __syncthreads();
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (t == 0) *out = acc;
}

// interval form of the rates at the calibrated eta and the visiting order key
// (shell, type, anchor): type 0 linear (D = a u), 1 clipped rising (active
// from the anchor to the top, D = -1 + a (u - u_anchor) + b), 2 clipped
// falling (active from the bottom to the anchor); a group whose rates are all
// 1 (zero loading) is left out of the visit (key past every shell)
__global__ void k_vp_intervals(const int* __restrict__ gnum, const int* __restrict__ gshell,
                               const double* __restrict__ gminv, const double* __restrict__ gtm,
                               const int* __restrict__ gbk, const double* __restrict__ gphi,
                               int* __restrict__ gtype, int* __restrict__ ganch,
                               double* __restrict__ ga, double* __restrict__ gb,
                               int* __restrict__ okey, int* __restrict__ gord) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= *gnum) return;
  int bk = gbk[g];
  const bool neg = bk < 0;
  if (neg) bk = -1 - bk;
  const double minv = gminv[g], tm = gtm[g];
  int type, anch;
  double a, b;
  if (bk == 0) {
    type = 0;
    anch = 0;
    a = neg ? -minv : minv;
    b = 0.0;
  } else if (!neg) {
    type = 1;
    anch = bk;
    a = minv;
    b = kGhU[bk] * minv - tm;
  } else {
    type = 2;
    anch = kNs - 1 - bk;
    a = -minv;
    b = kGhU[bk] * minv - tm;          // lam at u_anch = -u_bk
  }
  gtype[g] = type;
  ganch[g] = anch;
  ga[g] = a;
  gb[g] = b;
  const bool skip = gphi[g] == 0.0;
  const int within = (type == 0) ? 0 : (type == 1) ? kNs + anch : 2 * kNs + (kNs - 1 - anch);
  okey[g] = skip ? (1 << kOkeyBits) - 1 : gshell[g] * 3 * kNs + within;
  gord[g] = g;
}

// the visited groups' cells, contiguous in visiting order
__global__ void k_vp_vlen(const int* __restrict__ gnum, const int* __restrict__ okey,
                          const int* __restrict__ gord, const int* __restrict__ glen,
                          long long gcap, long long* __restrict__ vlen) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= gcap) return;
  vlen[i] = (i < *gnum && okey[i] != (1 << kOkeyBits) - 1) ? glen[gord[i]] : 0;
}

__global__ void k_vp_visit(const int* __restrict__ gnum, const int* __restrict__ okey,
                           const int* __restrict__ gord, const long long* __restrict__ gstart,
                           const int* __restrict__ glen, const long long* __restrict__ vstart,
                           const int* __restrict__ idx, const int* __restrict__ cu,
                           const double* __restrict__ cw, int* __restrict__ vcu,
                           double* __restrict__ vcw, int* __restrict__ vlast,
                           int* __restrict__ vgrp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= *gnum || okey[i] == (1 << kOkeyBits) - 1) return;
  const int g = gord[i];
  const long long a = gstart[g], n = glen[g], o = vstart[i];
  for (long long q = 0; q < n; ++q) {
    const int c = idx[a + q];
    vcu[o + q] = cu[c];
    vcw[o + q] = cw[c];
    vlast[o + q] = (q == n - 1) ? 1 : 0;
    vgrp[o + q] = g;
  }
}

// per shell: first visited cell and group, and the visited totals
__global__ void k_vp_visit_shells(const int* __restrict__ gnum, const int* __restrict__ okey,
                                  const long long* __restrict__ vstart, int n_z,
                                  long long* __restrict__ shv, long long* __restrict__ meta) {
  const int s = threadIdx.x;
  if (s > n_z) return;
  const int ng = *gnum;
  const int want = s * 3 * kNs;
  int lo = 0, hi = ng;                  // first visiting slot with key >= want
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (okey[mid] < want) lo = mid + 1; else hi = mid;
  }
  shv[s] = vstart[lo];
  if (s == n_z) {
    meta[3] = lo;
    meta[4] = vstart[lo];
  }
}

// ------------------------------------------------------------- ln R ----
// J0(q gamma_u) for every pooled bin u and column q
__global__ void k_vp_jtab(int nU, const double* __restrict__ pg, const double* __restrict__ qv,
                          int nq, double* __restrict__ Jt) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nU * nq) return;
  const int u = (int)(e / nq), q = (int)(e - (long long)u * nq);
  Jt[e] = j0(pg[u] * qv[q]);
}

constexpr int kTq = 64;                 // columns per tile (lanes)
constexpr int kRowsPer = 4;             // rows per thread
constexpr int kTk = 4 * kRowsPer;       // rows per tile (4 waves)
constexpr int kChunk = 64;              // cells staged per step

// z[s][j][p] (left and right sweeps) and c[s][p] of one shell on a tile of
// points p = (k row, q column); the groups of the shell visited in the order
// linear, rising by anchor, falling by anchor descending
__global__ __launch_bounds__(256) void k_vp_zstate(
    const long long* __restrict__ shv, const int* __restrict__ vcu,
    const double* __restrict__ vcw, const int* __restrict__ vlast, const int* __restrict__ vgrp,
    const int* __restrict__ gtype, const int* __restrict__ ganch, const double* __restrict__ ga,
    const double* __restrict__ gb, const double* __restrict__ pk, const double* __restrict__ Jt,
    const double* __restrict__ kv, int k0, int nk, int nq, int ntq, double* __restrict__ z,
    double* __restrict__ cz) {
  __shared__ double Jc[kChunk][kTq];
  __shared__ double Er[kChunk][kTk], Ei[kChunk][kTk];
  __shared__ double Wc[kChunk];
  __shared__ int Lc[kChunk], Gc[kChunk];
  const int s = blockIdx.y;
  const int tile = blockIdx.x, tk = tile / ntq, tq = tile - tk * ntq;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int q = tq * kTq + lane;
  const int rbase = tk * kTk + wave * kRowsPer;      // local row (within the chunk of rows)
  const long long P = (long long)nk * nq;
  const long long c0 = shv[s], c1 = shv[s + 1];
  double Yr[kRowsPer], Yi[kRowsPer], lar[kRowsPer], lai[kRowsPer], cr[kRowsPer], ci[kRowsPer];
  double axr[kRowsPer], axi[kRowsPer], bxr[kRowsPer], bxi[kRowsPer];
#pragma unroll
  for (int r = 0; r < kRowsPer; ++r) {
    Yr[r] = Yi[r] = lar[r] = lai[r] = cr[r] = ci[r] = 0.0;
    axr[r] = axi[r] = bxr[r] = bxi[r] = 0.0;
  }
  int phase = 0;          // 0: linear groups, 1: rising sweep, 2: falling sweep
  int jcur = 0;           // rising: next state to emit; falling: next state (descending)
  auto emit_rising = [&](int j) {
#pragma unroll
    for (int r = 0; r < kRowsPer; ++r) {
      const int row = rbase + r;
      if (row < nk && q < nq) {
        const long long p = (long long)row * nq + q;
        double* zz = z + (((size_t)s * kNs + j) * P + p) * 2;
        zz[0] = kGhU[j] * lar[r] + bxr[r];
        zz[1] = kGhU[j] * lai[r] + bxi[r];
      }
    }
  };
  auto emit_falling = [&](int j) {
#pragma unroll
    for (int r = 0; r < kRowsPer; ++r) {
      const int row = rbase + r;
      if (row < nk && q < nq) {
        const long long p = (long long)row * nq + q;
        double* zz = z + (((size_t)s * kNs + j) * P + p) * 2;
        zz[0] += bxr[r];
        zz[1] += bxi[r];
      }
    }
  };
  // advance the rising sweep to state j (emitting the states before it)
  auto rise_to = [&](int j) {
    while (jcur < j) {
      emit_rising(jcur);
      ++jcur;
      const double du = kGhU[jcur] - kGhU[jcur - 1];
#pragma unroll
      for (int r = 0; r < kRowsPer; ++r) {
        bxr[r] += du * axr[r];
        bxi[r] += du * axi[r];
      }
    }
  };
  auto fall_to = [&](int j) {
    while (jcur > j) {
      emit_falling(jcur);
      --jcur;
      const double du = kGhU[jcur] - kGhU[jcur + 1];
#pragma unroll
      for (int r = 0; r < kRowsPer; ++r) {
        bxr[r] += du * axr[r];
        bxi[r] += du * axi[r];
      }
    }
  };
  auto start_falling = [&]() {
    // the rising sweep ends: every remaining state emitted (from the linear
    // part alone when no rising group came)
    if (phase < 1) { phase = 1; jcur = 0; }
    while (jcur < kNs) {
      emit_rising(jcur);
      ++jcur;
      if (jcur < kNs) {
        const double du = kGhU[jcur] - kGhU[jcur - 1];
#pragma unroll
        for (int r = 0; r < kRowsPer; ++r) {
          bxr[r] += du * axr[r];
          bxi[r] += du * axi[r];
        }
      }
    }
    phase = 2;
    jcur = kNs - 1;
#pragma unroll
    for (int r = 0; r < kRowsPer; ++r) axr[r] = axi[r] = bxr[r] = bxi[r] = 0.0;
  };
  for (long long cb = c0; cb < c1; cb += kChunk) {
    const int nc = (int)min((long long)kChunk, c1 - cb);
    __syncthreads();
    for (int e = t; e < nc * kTq; e += 256) {
      const int c = e / kTq, l = e - c * kTq;
      const int qq = tq * kTq + l;
      Jc[c][l] = (qq < nq) ? Jt[(size_t)vcu[cb + c] * nq + qq] : 0.0;
    }
    for (int e = t; e < nc * kTk; e += 256) {
      const int c = e / kTk, r = e - c * kTk;
      const int row = tk * kTk + r;
      double er = 0.0, ei = 0.0;
      if (row < nk) {
        const double y = kv[k0 + row] * pk[vcu[cb + c]];
        const double h = sin(0.5 * y);
        er = -2.0 * h * h;
        ei = sin(y);
      }
      const double w = vcw[cb + c];
      Er[c][r] = w * er;
      Ei[c][r] = w * ei;
    }
    if (t < nc) {
      Wc[t] = vcw[cb + t];
      Lc[t] = vlast[cb + t];
      Gc[t] = vgrp[cb + t];
    }
    __syncthreads();
    for (int c = 0; c < nc; ++c) {
      const double J = Jc[c][lane];
      const double wjm1 = Wc[c] * (J - 1.0);
#pragma unroll
      for (int r = 0; r < kRowsPer; ++r) {
        const int rr = wave * kRowsPer + r;
        Yr[r] += Er[c][rr] * J + wjm1;
        Yi[r] += Ei[c][rr] * J;
      }
      if (Lc[c]) {
        const int g = __builtin_amdgcn_readfirstlane(Gc[c]);
        const int type = gtype[g], anch = ganch[g];
        const double a = ga[g], b = gb[g];
        if (type == 0) {
#pragma unroll
          for (int r = 0; r < kRowsPer; ++r) {
            lar[r] += a * Yr[r];
            lai[r] += a * Yi[r];
          }
        } else {
          if (type == 1) {
            if (phase == 0) { phase = 1; jcur = 0; }
            rise_to(anch);
          } else {
            if (phase < 2) start_falling();
            fall_to(anch);
          }
#pragma unroll
          for (int r = 0; r < kRowsPer; ++r) {
            cr[r] -= Yr[r];
            ci[r] -= Yi[r];
            bxr[r] += b * Yr[r];
            bxi[r] += b * Yi[r];
            axr[r] += a * Yr[r];
            axi[r] += a * Yi[r];
          }
        }
#pragma unroll
        for (int r = 0; r < kRowsPer; ++r) Yr[r] = Yi[r] = 0.0;
      }
    }
  }
  if (phase < 2) start_falling();
  // the falling sweep ends: the remaining states down to 0
  while (true) {
    emit_falling(jcur);
    if (jcur == 0) break;
    --jcur;
    const double du = kGhU[jcur] - kGhU[jcur + 1];
#pragma unroll
    for (int r = 0; r < kRowsPer; ++r) {
      bxr[r] += du * axr[r];
      bxi[r] += du * axi[r];
    }
  }
#pragma unroll
  for (int r = 0; r < kRowsPer; ++r) {
    const int row = rbase + r;
    if (row < nk && q < nq) {
      const long long p = (long long)row * nq + q;
      cz[((size_t)s * P + p) * 2] = cr[r];
      cz[((size_t)s * P + p) * 2 + 1] = ci[r];
    }
  }
}

__device__ __forceinline__ double read_lane_v(double x, int l) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// forward recursion per point (pos_closure.py:183-240), one wave per point
constexpr int kRecWaves = 8;
__global__ __launch_bounds__(64 * kRecWaves) void k_vp_recur(const double* __restrict__ z,
                                                             const double* __restrict__ cz,
                                                             const double* __restrict__ K, int n_z,
                                                             long long P,
                                                             double* __restrict__ out) {
  __shared__ double Ks[kNs * kNs];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long p = (long long)blockIdx.x * kRecWaves + wave;
  const bool own = lane < kNs, live = p < P;
  double ar = own ? kGhW[lane] : 0.0, ai = 0.0, scale = 0.0;
  for (int s = 0; s < n_z; ++s) {
    if (s > 0) {
      __syncthreads();
      for (int e = threadIdx.x; e < kNs * kNs; e += 64 * kRecWaves)
        Ks[e] = K[(size_t)(s - 1) * kNs * kNs + e];
      __syncthreads();
      double nr = 0.0, ni = 0.0;
      for (int i = 0; i < kNs; ++i) {
        const double kij = own ? Ks[i * kNs + lane] : 0.0;
        nr = fma(kij, read_lane_v(ar, i), nr);
        ni = fma(kij, read_lane_v(ai, i), ni);
      }
      ar = nr;
      ai = ni;
    }
    double zr = -1e308, zi = 0.0;
    if (own && live) {
      const double* zz = z + (((size_t)s * kNs + lane) * P + p) * 2;
      const double* cc = cz + ((size_t)s * P + p) * 2;
      zr = zz[0] + cc[0];
      zi = zz[1] + cc[1];
    }
    const double mx = live ? wave_max_v(zr) : 0.0;
    if (own && live) {
      const double e = exp(zr - mx);
      double sn, cs;
      sincos(zi, &sn, &cs);
      const double er = e * cs, ei = e * sn;
      const double xr = ar * er - ai * ei;
      ai = ar * ei + ai * er;
      ar = xr;
    }
    const double nrm = live ? wave_max_v(own ? sqrt(ar * ar + ai * ai) : 0.0) : 1.0;
    ar /= nrm;
    ai /= nrm;
    scale += mx + log(nrm);
  }
  double sr = ar, si = ai;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sr += __shfl_xor(sr, o);
    si += __shfl_xor(si, o);
  }
  if (live && lane == 0) {
    out[2 * p] = scale + 0.5 * log(sr * sr + si * si);
    out[2 * p + 1] = atan2(si, sr);
  }
}

// --------------------------------------------------- unwrap and shear ----
// numpy.unwrap's correction of one step d (period 2 pi)
__device__ inline double unwrap_step(double d) {
  const double P = 2.0 * kPiV;
  double m = fmod(d + kPiV, P);
  if (m != 0.0 && m < 0.0) m += P;
  double dd = m - kPiV;
  if (dd == -kPiV && d > 0.0) dd = kPiV;
  return (fabs(d) < kPiV) ? 0.0 : dd - d;
}
__global__ void k_vp_unwrap_col(double* __restrict__ v, int nk, int nq) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double corr = 0.0, prev = v[1];
  for (int i = 1; i < nk; ++i) {
    const double cur = v[((size_t)i * nq) * 2 + 1];
    corr += unwrap_step(cur - prev);
    prev = cur;
    v[((size_t)i * nq) * 2 + 1] = cur + corr;
  }
}
__global__ void k_vp_unwrap_rows(double* __restrict__ v, const double* __restrict__ qv, int nk,
                                 int nq, double shear_var) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nk) return;
  double* row = v + (size_t)i * nq * 2;
  double corr = 0.0, prev = row[1];
  row[0] -= 0.25 * shear_var * qv[0] * qv[0];
  for (int l = 1; l < nq; ++l) {
    const double cur = row[2 * l + 1];
    corr += unwrap_step(cur - prev);
    prev = cur;
    row[2 * l + 1] = cur + corr;
    row[2 * l] -= 0.25 * shear_var * qv[l] * qv[l];
  }
}

// ------------------------------------------------------------ spline ----
// not-a-knot cubic slopes on n uniform nodes x_i = i / (n - 1) (the
// tridiagonal system of scipy's CubicSpline, Thomas algorithm); y and the
// slopes strided by ys
__device__ inline void nak_slopes(const double* y, long long ys, int n, double* sl,
                                  double* cp) {
  const double h = 1.0 / (n - 1);
  auto dl = [&](int i) { return (y[(size_t)(i + 1) * ys] - y[(size_t)i * ys]) / h; };
  if (n == 2) {
    sl[0] = sl[ys] = dl(0);
    return;
  }
  if (n == 3) {                           // the parabola through three points
    const double d0 = dl(0), d1 = dl(1);
    sl[0] = (3.0 * d0 - d1) * 0.5;
    sl[ys] = (d0 + d1) * 0.5;
    sl[2 * ys] = (3.0 * d1 - d0) * 0.5;
    return;
  }
  // row 0: h s0 + 2h s1 = ((h + 4h) h d0 + h^2 d1) / (2h)
  {
    const double diag = h, up = 2.0 * h;
    const double r = ((h + 4.0 * h) * h * dl(0) + h * h * dl(1)) / (2.0 * h);
    cp[0] = up / diag;
    sl[0] = r / diag;
  }
  for (int i = 1; i < n - 1; ++i) {
    const double lo = h, di = 4.0 * h, up = h;
    const double r = 3.0 * (h * dl(i - 1) + h * dl(i));
    const double den = di - lo * cp[i - 1];
    cp[i] = up / den;
    sl[(size_t)i * ys] = (r - lo * sl[(size_t)(i - 1) * ys]) / den;
  }
  {
    const double lo = 2.0 * h, di = h;
    const double r = (h * h * dl(n - 3) + (4.0 * h + h) * h * dl(n - 2)) / (2.0 * h);
    const double den = di - lo * cp[n - 2];
    sl[(size_t)(n - 1) * ys] = (r - lo * sl[(size_t)(n - 2) * ys]) / den;
  }
  for (int i = n - 2; i >= 0; --i) sl[(size_t)i * ys] -= cp[i] * sl[(size_t)(i + 1) * ys];
}
// the cubic Hermite piece at x in [0, 1] on n uniform nodes
__device__ inline double nak_eval(const double* y, const double* sl, long long ys, int n, double x) {
  const double h = 1.0 / (n - 1);
  int i = (int)floor(x / h);
  if (i < 0) i = 0;
  if (i > n - 2) i = n - 2;
  const double xi = i * h, t = x - xi;
  const double y0 = y[(size_t)i * ys], y1 = y[(size_t)(i + 1) * ys];
  const double s0 = sl[(size_t)i * ys], s1 = sl[(size_t)(i + 1) * ys];
  const double d = (y1 - y0) / h;
  const double c3 = (s0 + s1 - 2.0 * d) / (h * h);
  const double c2 = (3.0 * d - 2.0 * s0 - s1) / h;
  return ((c3 * t + c2) * t + s0) * t + y0;
}

__device__ inline double lin_at(double hi, int n, int i) {
  return (i == n - 1) ? hi : i * (hi / (n - 1));
}

// slopes along k_kappa of every coarse column (thread per (part, column))
__global__ void k_vp_spl0(const double* __restrict__ Q, int nkc, int nc, double* __restrict__ S0,
                          double* __restrict__ cp) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * nc) return;
  const int part = e / nc, l = e - part * nc;
  nak_slopes(Q + 2 * l + part, 2LL * nc, nkc, S0 + 2 * l + part, cp + (size_t)e * nkc);
}
// the k_kappa interpolant on every full-grid row, for each coarse column
__global__ void k_vp_eval0(const double* __restrict__ Q, const double* __restrict__ S0, int nkc,
                           int nc, int nk, double* __restrict__ R) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nk * nc * 2) return;
  const int i = (int)(e / (2 * nc)), rem = (int)(e - (long long)i * 2 * nc);
  const int l = rem >> 1, part = rem & 1;
  const double x = sqrt(fmin(fmax(lin_at(1.0, nk, i), 0.0), 1.0));
  R[e] = nak_eval(Q + 2 * l + part, S0 + 2 * l + part, 2LL * nc, nkc, x);
}
// then along k_gamma on every full row (thread per (row, part)), evaluated at
// every full column: psi[i][j][2]
__global__ void k_vp_spl1(const double* __restrict__ R, int nc, int nk, int ng,
                          double* __restrict__ S1, double* __restrict__ cp,
                          double* __restrict__ psi) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 2 * nk) return;
  const int i = e >> 1, part = e & 1;
  const double* y = R + (size_t)i * nc * 2 + part;
  double* sl = S1 + (size_t)i * nc * 2 + part;
  nak_slopes(y, 2, nc, sl, cp + (size_t)e * nc);
  for (int j = 0; j < ng; ++j) {
    const double x = sqrt(fmin(fmax(lin_at(1.0, ng, j), 0.0), 1.0));
    psi[((size_t)i * ng + j) * 2 + part] = nak_eval(y, sl, 2, nc, x);
  }
}

// --------------------------------------------------------------- window ----
__global__ void k_vp_window(double* __restrict__ phi, int nk, int ng) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)nk * ng) return;
  const int i = (int)(e / ng), j = (int)(e - (long long)i * ng);
  const double x = fabs(lin_at(1.0, nk, i)) * 2.0;
  const double wk = ((x < 1.0) ? 2.0 / 3.0 - x * x + x * x * x / 2.0
                               : (x < 2.0) ? (2.0 - x) * (2.0 - x) * (2.0 - x) / 6.0 : 0.0) /
                    (2.0 / 3.0);
  const double r = fmin(fmax(fabs(lin_at(1.0, ng, j)), 0.0), 1.0);
  const double om = 1.0 - r;
  const double wg = om * om * om * om * (4.0 * r + 1.0);
  const double f = wk * wg;
  phi[(size_t)(2 * i) * ng + j] *= f;
  phi[(size_t)(2 * i + 1) * ng + j] *= f;
}

struct RefTabs {
  const double *cosmo, *derived, *slices;
};
RefTabs ref_tabs(const void* tables, int32_t flags) {
  const RefTablesOffsets o = ref_tables_offsets(flags);
  const double* T = static_cast<const double*>(tables);
  return RefTabs{T + o.cosmo, T + o.derived, T + o.slices};
}

}  // namespace
}  // namespace sgl

using namespace sgl;

extern "C" {

size_t sgl_ref_vp_runs_bytes(int64_t n_points) {
  if (n_points < 1 || n_points > ((int64_t)1 << 31) - 1) return 0;
  return run_ws(n_points).total;
}

sgl_status sgl_ref_vp_runs(const void* tables, int32_t flags, int64_t n_points,
                           const int64_t* offsets, const int32_t* key, const double* kap,
                           const double* gam, const double* w, void* ws, int64_t* n_cells,
                           void* stream) {
  if (!tables || !offsets || !key || !kap || !gam || !w || !ws || !n_cells || n_points < 1 ||
      n_points > ((int64_t)1 << 31) - 1)
    return SGL_ERR_INVALID;
  const RunWs L = run_ws(n_points);
  char* b = static_cast<char*>(ws);
  hipStream_t st = vst(stream);
  int* ckey = (int*)(b + L.ckey);
  int* skey = (int*)(b + L.skey);
  VKGW* vin = (VKGW*)(b + L.vin);
  VKGW* vout = (VKGW*)(b + L.vout);
  int* ukey = (int*)(b + L.ukey);
  int* rlen = (int*)(b + L.rlen);
  int* nruns = (int*)(b + L.nruns);
  long long* rstart = (long long*)(b + L.rstart);
  int* nparts = (int*)(b + L.nparts);
  long long* poff = (long long*)(b + L.poff);
  double* psum = (double*)(b + L.psum);
  double* rsum = (double*)(b + L.rsum);
  void* tmp = b + L.tmp;
  const RefTablesOffsets o = ref_tables_offsets(flags);
  const double* cosmo = static_cast<const double*>(tables) + o.cosmo;
  int n_z = 0;
  {
    double nz = 0.0;
    if (hipMemcpyAsync(&nz, cosmo + 10, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return SGL_ERR_HIP;
    n_z = (int)nz;
  }
  if (n_z < 1 || n_z > kVpShells) return SGL_ERR_INVALID;
  const int tps = sgl_ref_tasks(flags) / kRefMaxShells;
  const long long n = n_points;
  hipLaunchKernelGGL(k_vp_ckey, dim3(nblk(n)), dim3(256), 0, st, n,
                     reinterpret_cast<const long long*>(offsets), tps, n_z, key, kap, gam, w, ckey,
                     vin);
  size_t tb = L.tmp_bytes;
  if (rocprim::radix_sort_pairs(tmp, tb, ckey, skey, vin, vout, (size_t)n, 0, kCkeyBits, st) !=
      hipSuccess)
    return SGL_ERR_HIP;
  tb = L.tmp_bytes;
  if (rocprim::run_length_encode(tmp, tb, skey, (size_t)n, ukey, rlen, nruns, st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_vp_zero_tail, dim3(nblk(L.cap)), dim3(256), 0, st, rlen, nruns, L.cap);
  tb = L.tmp_bytes;
  if (rocprim::exclusive_scan(tmp, tb, rlen, rstart, 0LL, (size_t)L.cap, rocprim::plus<long long>(),
                              st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_vp_nparts, dim3(nblk(L.cap)), dim3(256), 0, st, rlen, nruns, L.cap, nparts);
  tb = L.tmp_bytes;
  if (rocprim::exclusive_scan(tmp, tb, nparts, poff, 0LL, (size_t)L.cap, rocprim::plus<long long>(),
                              st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_vp_parts, dim3(nblk(L.max_parts * 64)), dim3(256), 0, st, rlen, nruns,
                     rstart, nparts, poff, vout, psum);
  hipLaunchKernelGGL(k_vp_combine, dim3(nblk(L.cap)), dim3(256), 0, st, ukey, nruns, nparts, poff,
                     psum, rsum, reinterpret_cast<long long*>(n_cells));
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_vp_state_bytes(int64_t n_cells) {
  if (n_cells < 1 || n_cells > ((int64_t)1 << 31) - 1) return 0;
  return vp_state(n_cells).total;
}

sgl_status sgl_ref_vp_setup(const void* tables, int32_t flags, int64_t n_points,
                            const void* runs_ws, int64_t n_cells, const double* pop,
                            const int32_t* npop, const void* xi_state, const void* xi_rows_ws,
                            int64_t n_xi_runs, int64_t n_xi_parts, void* state, int64_t* meta_out,
                            void* stream) {
  if (!tables || !runs_ws || !pop || !npop || !xi_state || !xi_rows_ws || !state || !meta_out ||
      n_cells < 1 || n_points < 1 || n_xi_runs < 1 || n_xi_parts < 1 ||
      !(flags & SGL_FLAG_CLUSTERING))
    return SGL_ERR_INVALID;
  const RunWs R = run_ws(n_points);
  const VpState L = vp_state(n_cells);
  const RefTabs T = ref_tabs(tables, flags);
  const XiRowsView X = xi_rows_view(xi_rows_ws, n_xi_runs, n_xi_parts);
  const char* rb = static_cast<const char*>(runs_ws);
  const int* ukey = (const int*)(rb + R.ukey);
  const double* rsum = (const double*)(rb + R.rsum);
  void* S = state;
  hipStream_t st = vst(stream);
  const double* xs = static_cast<const double*>(xi_state);
  long long* meta = at<long long>(S, L.meta);
  const double* pk = pop + kFine;
  const long long n = n_cells;
  int n_z = 0;
  {
    double nz = 0.0;
    if (hipMemcpyAsync(&nz, T.cosmo + 10, sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return SGL_ERR_HIP;
    n_z = (int)nz;
  }
  if (n_z < 1 || n_z > kVpShells) return SGL_ERR_INVALID;
  if (hipMemsetAsync(meta, 0, 16 * sizeof(long long), st) != hipSuccess) return SGL_ERR_HIP;
  int* occ = at<int>(S, L.bocc);
  long long* bmap = at<long long>(S, L.bmap);
  long long* bocc64 = at<long long>(S, L.bocc64);
  if (hipMemsetAsync(occ, 0, (size_t)kFine * sizeof(int), st) != hipSuccess) return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_vp_occ, dim3(nblk(n)), dim3(256), 0, st, n, ukey, rsum, occ);
  hipLaunchKernelGGL(k_vp_int_to_ll, dim3(nblk(kFine)), dim3(256), 0, st, occ, bocc64, kFine);
  void* tmp = at<char>(S, L.tmp);
  size_t tb = L.tmp_bytes;
  if (rocprim::exclusive_scan(tmp, tb, bocc64, bmap, 0LL, (size_t)kFine,
                              rocprim::plus<long long>(), st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_vp_check_occ, dim3(1), dim3(64), 0, st, bmap, occ, npop, meta);
  int* cu = at<int>(S, L.cu);
  double* cw = at<double>(S, L.cw);
  double* ckap = at<double>(S, L.ckap);
  double* cphi = at<double>(S, L.cphi);
  int* gk_in = at<int>(S, L.gkey_in);
  int* gk_out = at<int>(S, L.gkey_out);
  int* id_in = at<int>(S, L.idx_in);
  int* id_out = at<int>(S, L.idx_out);
  hipLaunchKernelGGL(k_vp_cells, dim3(nblk(n)), dim3(256), 0, st, n, ukey, rsum, bmap, pk, xs,
                     X.ukey, X.phic, cu, cw, ckap, cphi, gk_in, id_in);
  tb = L.tmp_bytes;
  if (rocprim::radix_sort_pairs(tmp, tb, gk_in, gk_out, id_in, id_out, (size_t)n, 0, kGkeyBits,
                                st) != hipSuccess)
    return SGL_ERR_HIP;
  int* gkey = at<int>(S, L.gkey);
  int* glen = at<int>(S, L.glen);
  int* gnum = at<int>(S, L.gnum);
  long long* gstart = at<long long>(S, L.gstart);
  tb = L.tmp_bytes;
  if (rocprim::run_length_encode(tmp, tb, gk_out, (size_t)n, gkey, glen, gnum, st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_vp_zero_tail, dim3(nblk(L.gcap)), dim3(256), 0, st, glen, gnum, L.gcap);
  tb = L.tmp_bytes;
  if (rocprim::exclusive_scan(tmp, tb, glen, gstart, 0LL, (size_t)L.gcap,
                              rocprim::plus<long long>(), st) != hipSuccess)
    return SGL_ERR_HIP;
  int* gshell = at<int>(S, L.gshell);
  double* gphi = at<double>(S, L.gphi);
  double* gwk = at<double>(S, L.gwk);
  hipLaunchKernelGGL(k_vp_groups, dim3(nblk(L.gcap)), dim3(256), 0, st, gkey, glen, gnum, gstart,
                     id_out, cw, ckap, cphi, gshell, gphi, gwk);
  long long* shg = at<long long>(S, L.shg);
  hipLaunchKernelGGL(k_vp_shell_groups, dim3(1), dim3(128), 0, st, gshell, gnum, gstart, n_z, shg,
                     meta);
  // rho, f and the transitions
  double* P0 = at<double>(S, L.P0);
  double* los = at<double>(S, L.los);
  double* rho = at<double>(S, L.rho);
  double* f = at<double>(S, L.f);
  double* K = at<double>(S, L.K);
  hipLaunchKernelGGL(k_vp_p0, dim3(nblk((long long)kXiNk * kLosN)), dim3(256), 0, st, T.cosmo,
                     T.derived, xs, P0);
  if (n_z > 1)
    hipLaunchKernelGGL(k_vp_los, dim3(n_z - 1, kXiNk), dim3(256), 0, st, T.slices, P0, los);
  hipLaunchKernelGGL(k_vp_rho, dim3(1), dim3(128), 0, st, los, X.v0, n_z, rho, f, meta);
  if (n_z > 1)
    hipLaunchKernelGGL(k_vp_transition, dim3(n_z - 1), dim3(64), 0, st, rho, K, meta);
  if (hipMemcpyAsync(meta_out, meta, 16 * sizeof(long long), hipMemcpyDeviceToDevice, st) !=
      hipSuccess)
    return SGL_ERR_HIP;
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

}  // extern "C"

namespace sgl {
namespace {
__global__ void k_vp_fill_order(const int* __restrict__ gnum, long long gcap,
                                int* __restrict__ okey, int* __restrict__ gord) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < gcap && g >= *gnum) {
    okey[g] = (1 << kOkeyBits) - 1;
    gord[g] = (int)g;
  }
}
constexpr long long kZBytes = 2LL << 30;    // z states of one row chunk
struct LnrWs {
  size_t jt, z, cz, total;
  int rows;
};
LnrWs lnr_ws(int n_z, long long nU, int nk, int nq) {
  LnrWs W;
  const long long per_row = (long long)n_z * kNs * nq * 16;
  long long rows = std::max<long long>(kTk, kZBytes / std::max<long long>(per_row, 1) / kTk * kTk);
  rows = std::min<long long>(rows, ((long long)nk + kTk - 1) / kTk * kTk);
  W.rows = (int)rows;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += al256(b); return r; };
  W.jt = take((size_t)nU * nq * sizeof(double));
  W.z = take((size_t)n_z * kNs * rows * nq * 2 * sizeof(double));
  W.cz = take((size_t)n_z * rows * nq * 2 * sizeof(double));
  W.total = o;
  return W;
}
struct SplWs {
  size_t s0, cp0, r, s1, cp1, total;
};
SplWs spl_ws(int nkc, int nc, int nk) {
  SplWs W;
  size_t o = 0;
  auto take = [&](size_t b) { size_t r = o; o += al256(b); return r; };
  W.s0 = take((size_t)nkc * nc * 2 * sizeof(double));
  W.cp0 = take((size_t)2 * nc * nkc * sizeof(double));
  W.r = take((size_t)nk * nc * 2 * sizeof(double));
  W.s1 = take((size_t)nk * nc * 2 * sizeof(double));
  W.cp1 = take((size_t)2 * nk * nc * sizeof(double));
  W.total = o;
  return W;
}
}  // namespace
}  // namespace sgl

extern "C" {

sgl_status sgl_ref_vp_variance(void* state, int64_t n_cells, int32_t n_z, double eta, double* out,
                               void* stream) {
  if (!state || !out || n_cells < 1 || n_z < 1 || n_z > kVpShells || !(eta > 0.0))
    return SGL_ERR_INVALID;
  const VpState L = vp_state(n_cells);
  void* S = state;
  hipStream_t st = vst(stream);
  hipLaunchKernelGGL(k_vp_grates, dim3(nblk(L.gcap)), dim3(256), 0, st, at<int>(S, L.gnum),
                     at<int>(S, L.gshell), at<double>(S, L.gphi), at<double>(S, L.f), eta,
                     at<double>(S, L.gminv), at<double>(S, L.gtm), at<int>(S, L.gbk),
                     at<long long>(S, L.meta));
  hipLaunchKernelGGL(k_vp_ms, dim3(n_z), dim3(64), 0, st, at<long long>(S, L.shg),
                     at<double>(S, L.gwk), at<double>(S, L.gminv), at<double>(S, L.gtm),
                     at<int>(S, L.gbk), at<double>(S, L.ms));
  hipLaunchKernelGGL(k_vp_chainvar, dim3(1), dim3(64), 0, st, at<double>(S, L.ms),
                     at<double>(S, L.K), (int)n_z, out);
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_vp_rates(void* state, int64_t n_cells, int32_t n_z, double eta,
                            int64_t* meta_out, void* stream) {
  if (!state || !meta_out || n_cells < 1 || n_z < 1 || n_z > kVpShells || !(eta > 0.0))
    return SGL_ERR_INVALID;
  const VpState L = vp_state(n_cells);
  void* S = state;
  hipStream_t st = vst(stream);
  int* gnum = at<int>(S, L.gnum);
  hipLaunchKernelGGL(k_vp_grates, dim3(nblk(L.gcap)), dim3(256), 0, st, gnum,
                     at<int>(S, L.gshell), at<double>(S, L.gphi), at<double>(S, L.f), eta,
                     at<double>(S, L.gminv), at<double>(S, L.gtm), at<int>(S, L.gbk),
                     at<long long>(S, L.meta));
  int* ok_in = at<int>(S, L.okey_in);
  int* ok_out = at<int>(S, L.okey_out);
  int* go_in = at<int>(S, L.gord_in);
  int* go_out = at<int>(S, L.gord_out);
  hipLaunchKernelGGL(k_vp_intervals, dim3(nblk(L.gcap)), dim3(256), 0, st, gnum,
                     at<int>(S, L.gshell), at<double>(S, L.gminv), at<double>(S, L.gtm),
                     at<int>(S, L.gbk), at<double>(S, L.gphi), at<int>(S, L.gtype),
                     at<int>(S, L.ganch), at<double>(S, L.ga), at<double>(S, L.gb), ok_in, go_in);
  hipLaunchKernelGGL(k_vp_fill_order, dim3(nblk(L.gcap)), dim3(256), 0, st, gnum, L.gcap, ok_in,
                     go_in);
  void* tmp = at<char>(S, L.tmp);
  size_t tb = L.tmp_bytes;
  // stable: equal keys keep the groups' (coarse bin) order
  if (rocprim::radix_sort_pairs(tmp, tb, ok_in, ok_out, go_in, go_out, (size_t)L.gcap, 0,
                                kOkeyBits, st) != hipSuccess)
    return SGL_ERR_HIP;
  long long* vlen = at<long long>(S, L.vlen);
  long long* vstart = at<long long>(S, L.vstart);
  hipLaunchKernelGGL(k_vp_vlen, dim3(nblk(L.gcap)), dim3(256), 0, st, gnum, ok_out, go_out,
                     at<int>(S, L.glen), L.gcap, vlen);
  tb = L.tmp_bytes;
  if (rocprim::exclusive_scan(tmp, tb, vlen, vstart, 0LL, (size_t)L.gcap,
                              rocprim::plus<long long>(), st) != hipSuccess)
    return SGL_ERR_HIP;
  hipLaunchKernelGGL(k_vp_visit, dim3(nblk(L.gcap)), dim3(256), 0, st, gnum, ok_out, go_out,
                     at<long long>(S, L.gstart), at<int>(S, L.glen), vstart,
                     at<int>(S, L.idx_out), at<int>(S, L.cu), at<double>(S, L.cw),
                     at<int>(S, L.vcu), at<double>(S, L.vcw), at<int>(S, L.vlast),
                     at<int>(S, L.vgrp));
  long long* shv = at<long long>(S, L.shg) + (kVpShells + 1);
  hipLaunchKernelGGL(k_vp_visit_shells, dim3(1), dim3(128), 0, st, gnum, ok_out, vstart, n_z, shv,
                     at<long long>(S, L.meta));
  if (hipMemcpyAsync(meta_out, at<long long>(S, L.meta), 16 * sizeof(long long),
                     hipMemcpyDeviceToDevice, st) != hipSuccess)
    return SGL_ERR_HIP;
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_vp_lnr_bytes(int32_t n_z, int32_t n_bins, int32_t nk, int32_t nq) {
  if (n_z < 1 || n_z > kVpShells || n_bins < 1 || nk < 1 || nq < 1) return 0;
  return lnr_ws(n_z, n_bins, nk, nq).total;
}

sgl_status sgl_ref_vp_lnr(const void* state, int64_t n_cells, int32_t n_z, const double* pop,
                          int32_t n_bins, const double* kv, int32_t nk, const double* qv,
                          int32_t nq, void* ws, double* out, void* stream) {
  if (!state || !pop || !kv || !qv || !ws || !out || n_cells < 1 || n_z < 1 ||
      n_z > kVpShells || n_bins < 1 || nk < 1 || nq < 1)
    return SGL_ERR_INVALID;
  const VpState L = vp_state(n_cells);
  const LnrWs W = lnr_ws(n_z, n_bins, nk, nq);
  void* S = const_cast<void*>(state);
  char* b = static_cast<char*>(ws);
  hipStream_t st = vst(stream);
  double* Jt = (double*)(b + W.jt);
  double* z = (double*)(b + W.z);
  double* cz = (double*)(b + W.cz);
  const double* pk = pop + kFine;
  const double* pg = pop + 2 * kFine;
  hipLaunchKernelGGL(k_vp_jtab, dim3(nblk((long long)n_bins * nq)), dim3(256), 0, st, n_bins, pg,
                     qv, nq, Jt);
  const long long* shv = at<long long>(S, L.shg) + (kVpShells + 1);
  const int ntq = (nq + kTq - 1) / kTq;
  for (int k0 = 0; k0 < nk; k0 += W.rows) {
    const int rows = std::min(W.rows, nk - k0);
    const int ntk = (rows + kTk - 1) / kTk;
    hipLaunchKernelGGL(k_vp_zstate, dim3(ntk * ntq, n_z), dim3(256), 0, st, shv,
                       at<int>(S, L.vcu), at<double>(S, L.vcw), at<int>(S, L.vlast),
                       at<int>(S, L.vgrp), at<int>(S, L.gtype), at<int>(S, L.ganch),
                       at<double>(S, L.ga), at<double>(S, L.gb), pk, Jt, kv, k0, rows, nq, ntq, z,
                       cz);
    const long long P = (long long)rows * nq;
    hipLaunchKernelGGL(k_vp_recur, dim3(nblk(P, kRecWaves)), dim3(64 * kRecWaves), 0, st, z, cz,
                       at<double>(S, L.K), (int)n_z, P, out + (size_t)k0 * nq * 2);
  }
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_vp_qxi(double* lnr, int32_t nk, int32_t nq, const double* qv,
                          double shear_var, void* stream) {
  if (!lnr || !qv || nk < 1 || nq < 1) return SGL_ERR_INVALID;
  hipStream_t st = vst(stream);
  hipLaunchKernelGGL(k_vp_unwrap_col, dim3(1), dim3(64), 0, st, lnr, nk, nq);
  hipLaunchKernelGGL(k_vp_unwrap_rows, dim3(nblk(nk)), dim3(256), 0, st, lnr, qv, nk, nq,
                     shear_var);
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

size_t sgl_ref_vp_spline_bytes(int32_t nkc, int32_t nc, int32_t nk) {
  if (nkc < 2 || nc < 2 || nk < 1) return 0;
  return spl_ws(nkc, nc, nk).total;
}

sgl_status sgl_ref_vp_spline(const double* coarse, int32_t nkc, int32_t nc, int32_t nk,
                             int32_t ng, void* ws, double* psi, void* stream) {
  if (!coarse || !ws || !psi || nkc < 2 || nc < 2 || nk < 1 || ng < 1) return SGL_ERR_INVALID;
  const SplWs W = spl_ws(nkc, nc, nk);
  char* b = static_cast<char*>(ws);
  hipStream_t st = vst(stream);
  double* S0 = (double*)(b + W.s0);
  double* cp0 = (double*)(b + W.cp0);
  double* R = (double*)(b + W.r);
  double* S1 = (double*)(b + W.s1);
  double* cp1 = (double*)(b + W.cp1);
  hipLaunchKernelGGL(k_vp_spl0, dim3(nblk(2 * nc)), dim3(256), 0, st, coarse, nkc, nc, S0, cp0);
  hipLaunchKernelGGL(k_vp_eval0, dim3(nblk((long long)nk * nc * 2)), dim3(256), 0, st, coarse, S0,
                     nkc, nc, nk, R);
  hipLaunchKernelGGL(k_vp_spl1, dim3(nblk(2 * nk)), dim3(256), 0, st, R, nc, nk, ng, S1, cp1, psi);
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_ref_window(double* phi, int32_t nk, int32_t ng, void* stream) {
  if (!phi || nk < 2 || ng < 2) return SGL_ERR_INVALID;
  hipLaunchKernelGGL(k_vp_window, dim3(nblk((long long)nk * ng)), dim3(256), 0, vst(stream), phi,
                     nk, ng);
  return hip_ok() ? SGL_OK : SGL_ERR_HIP;
}

}  // extern "C"
