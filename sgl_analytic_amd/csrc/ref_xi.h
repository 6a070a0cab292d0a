// The xi_lin (Limber) two-point term of the reference-shape path
// (SGL_FLAG_BIAS): constants, the per-PDF state layout and the emission hook
// shared by the cell kernels (ref2d.hip) and the xi kernels (xilin.hip).
#pragma once
#include "sgl_amd_internal.h"

namespace sgl {

// k_perp quadrature: Gauss-Legendre in ln k on [1e-4, 1e4] Mpc^-1 with
// (2, 2, 12, 8, 6, 4, 4, 4) nodes per decade (sgl_full.py:2480-2511)
constexpr int kXiNk = 42;
constexpr int kXiVals = 3 * kXiNk;        // J0, J2, J4 (m = 0, 1, 2) per node
// coarse (kappa, gamma) lattice: signed-log kappa edges on 1e-7..30 (170 per
// side), log gamma edges on 1e-9..30 (110) (sgl_full.py:457-468)
constexpr int kXiKc = 341, kXiGc = 110;
constexpr int kXiBins = kXiKc * kXiGc;
constexpr int kXiKeyInvalid = kRefMaxShells * kXiBins;   // outside: sorts last
constexpr int kXiKeyBits = 23;
static_assert(kXiKeyInvalid < (1 << kXiKeyBits), "coarse keys exceed the sort bits");
constexpr int kXiGroups = 3 * kRefMaxShells;             // (shell, m)
constexpr int kXiCoarse = 101;                           // Q's coarse grid (Lam xi_coarse)
constexpr double kXiSvdTol = 1e-6, kXiPrune = 1e-7;      // sgl_full.py:2436, Lam xi_prune
constexpr int kXiPart = 2048;                            // points per walk part

// Per-PDF state (doubles).  meta (int32): [0] n_runs (valid keys), [1]
// n_parts, [2] n_z, [3] nP, [4] nQ, then pmap[341], qmap[110] (lattice bin
// -> compact index, -1 unoccupied), plist[341], qlist[110], and the first run
// of each shell rfirst[kRefMaxShells + 1].
constexpr int kXmRuns = 0, kXmParts = 1, kXmNz = 2, kXmNP = 3, kXmNQ = 4, kXmPmap = 8;
constexpr int kXmQmap = kXmPmap + kXiKc;
constexpr int kXmPlist = kXmQmap + kXiGc;
constexpr int kXmQlist = kXmPlist + kXiKc;
constexpr int kXmRfirst = kXmQlist + kXiGc;
constexpr int kXmInts = kXmRfirst + kRefMaxShells + 1;
struct XiState {
  size_t meta, kp, c, ekc, egc, kmc, gmc, lx, sw, scal, lam, V, res, srow, W0, U, minv, qc, cs,
      total;
};
// scal: [0] clustering variance, [1] largest singular value over groups,
// [2] k_kappa max and [3] k_gamma max of the coarse grid, [4] the shear
// variance of the m = 1 groups (the positive closure's shear target)
__host__ __device__ inline XiState xi_state_layout() {
  XiState L;
  size_t o = 0;
  auto take = [&](size_t n) { size_t r = o; o += round32(n); return r; };
  L.meta = take((kXmInts + 1) / 2);
  L.kp = take(kXiNk);
  L.c = take(kXiNk);
  L.ekc = take(kXiKc + 1);
  L.egc = take(kXiGc + 1);
  L.kmc = take(kXiKc);
  L.gmc = take(kXiGc);
  L.lx = take((size_t)kRefMaxShells * 2 * kXiNk);
  L.sw = take((size_t)kRefMaxShells * kXiNk);
  L.scal = take(8);
  L.lam = take((size_t)kXiGroups * kXiNk);
  L.V = take((size_t)kXiGroups * kXiNk * kXiNk);
  L.res = take((size_t)kXiGroups * kXiNk);
  L.srow = take((size_t)kXiGroups * kXiNk);
  L.W0 = take((size_t)kRefMaxShells * kXiKc * kXiNk);
  L.U = take((size_t)kXiGroups * kXiGc * kXiNk);
  L.minv = take((size_t)kXiCoarse * kXiCoarse);
  L.qc = take((size_t)2 * kXiCoarse * kXiCoarse);
  L.cs = take((size_t)2 * kXiCoarse * kXiCoarse);
  L.total = o;
  return L;
}

// Offsets (doubles) of the reference tables that the xi kernels read, and the
// FP64 MFMA contraction batched over blockIdx.z, A optionally transposed
// (A^T stored K x M with leading dimension lda) (ref2d.hip).
struct RefTablesOffsets {
  size_t cosmo, derived, slices, rows;
};
RefTablesOffsets ref_tables_offsets(int flags);
int xi_gemm_batched(int M, int N, int K, const double* A, long lda, long sA, bool transA,
                    const double* B, long ldb, long sB, double* C, long ldc, long sC, int batch,
                    hipStream_t st);

// Emission hook of the cell kernels.  Count pass: cells per task (the
// distinct lens radii: a task's inner cells when it owns them, its tail and
// filament radii).  Emit pass: per point the cell id and the coarse key
// shell * kXiBins + bin, per cell (R [Mpc], beta), R < 0 for points without
// tangential shear (filaments).
struct XiEmit {
  long long* ccounts;
  const long long* coffsets;
  const double* ekc;
  const double* egc;
  const double* slices;
  int* pcid;
  int* pck;
  double* cells;
};

// What the positive closure reads of the xi rows workspace (sgl_ref_xi_rows):
// the run keys, each run's loading phic and lens rate, the shells' leading
// k_perp profiles |v0|.
struct XiRowsView {
  const int* ukey;
  const double *phic, *wrun, *v0;
};
XiRowsView xi_rows_view(const void* rows_ws, long long n_runs, long long n_parts);

}  // namespace sgl
