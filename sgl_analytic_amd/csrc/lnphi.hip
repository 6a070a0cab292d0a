// Compound-Poisson exponent of the lens population.
//
//   Lambda(k) = sum_c w_c (e^{i k xi_c} - 1 - i k xi_c)         (compensated)
//   a_i(k)    = sum_{c in slice i} beta_c w_c (e^{i k xi_c} - 1)
//
// Cells are never stored: every workgroup regenerates the cells of its slice
// from the (slice, mass) rows -- truncated NFW (kappa, gamma), jump
// xi = -ln detA, tilted weight w detA -- and streams them against its k nodes
// or its histogram.
//
// Exact path (k_lnphi_exact): every cell, FP64 sincos, the reference's
// cancellation-free forms cos - 1 = -2 sin^2(ph/2), sin - ph by series below
// |ph| = 1e-2 (sgl_full.py:2599-2613).
// Binned path:
//   k_rowf / k_cols   per-row and per-column records of the slice lattice;
//   k_bin_cells       every cell with the FP32 lens profile (a form without
//                     cancellation at x = 1) into a per-slice LDS histogram of fine
//                     signed-log jump bins (w and its centred first and second
//                     xi moments, beta w; 64-bit fixed point, so the sums do
//                     not depend on the order of the atomics), then
//                     the slice's fine partials and its coarse two-point
//                     moments (beta w, beta w xi, beta w xi^2 per kCoarse
//                     fine bins);
//   k_reduce_fine     slice-order sum of the partials (no atomics across
//                     workgroups: reproducible), compaction into Gaussian
//                     records (centroid, in-bin variance), prefix moments;
//   k_lnphi_binned    per k: records with k xi <= 1/2 through Taylor series
//                     on prefix moments, the rest as w e^{i k xibar - k^2 v/2}
//                     with the phase in FP64 reduced to revolutions, sin/cos
//                     on the FP32 hardware path, compensated small-phase terms
//                     and accumulation in FP64;
//   k_amp_binned      the slice amplitudes from the coarse records, the same
//                     way, at every kPsi2Stride-th k node.
#include <algorithm>
#include <mutex>

#include "sgl_amd_internal.h"

namespace sgl {

__global__ void k_fill_uedges(const sgl_shape sh, double* __restrict__ ue, double* __restrict__ kn,
                              double* __restrict__ m_nodes_h) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= sh.n_b) {
    double l0 = log10(sh.u_lo), l1 = log10(sh.u_hi);
    ue[t] = pow(10.0, l0 + (l1 - l0) * t / sh.n_b);
  }
  if (t < sh.n_k) {
    double l0 = log10(sh.k_lo), l1 = log10(sh.k_hi);
    kn[t] = pow(10.0, l0 + (l1 - l0) * t / (sh.n_k - 1));
  }
  if (m_nodes_h != nullptr && t < sh.n_m && (sh.flags & SGL_FLAG_REF_GRID)) {
    // engine mass node t + 1 of exp(linspace(ln 1e7, ln 1e17, 100)) Msun; the
    // last node is exp(ln 1e17), as linspace ends exactly at its stop value
    const double e0 = log(1e7), e1 = log(1e17);
    m_nodes_h[t] = exp((t + 1 == 99) ? e1 : e0 + (t + 1) * ((e1 - e0) / 99.0));
  } else if (m_nodes_h != nullptr && t < sh.n_m) {
    double e0 = log(sh.m_lo_h), e1 = log(sh.m_hi_h);
    double a = e0 + (e1 - e0) * t / sh.n_m, b = e0 + (e1 - e0) * (t + 1) / sh.n_m;
    m_nodes_h[t] = exp(0.5 * (a + b));
  }
}

// Per-slice LDS staging: row constants and the impact-parameter geometry
// (geometric bin centre and u1^2 - u0^2 of every column).
struct SliceStage {
  RowConst rc[kMaxMasses];
  double ug[kMaxB];
  double du2[kMaxB];
};

__device__ inline void stage_slice(SliceStage& S, const double* __restrict__ rows,
                                   const double* __restrict__ uedge, const sgl_shape& sh) {
  for (int j = threadIdx.x; j < sh.n_m; j += blockDim.x) S.rc[j] = row_const(rows + (size_t)j * kRowStride);
  for (int l = threadIdx.x; l < sh.n_b; l += blockDim.x) {
    double u0 = uedge[l], u1 = uedge[l + 1];
    S.ug[l] = sqrt(u0 * u1);
    S.du2[l] = u1 * u1 - u0 * u0;
  }
}

// Cell c = (row j, column l, orientation q) of the staged slice, FP64
// throughout; n_phi = kEllPhi with halo ellipticity (each node 1/n_phi of the
// cell's weight, the pseudo-elliptical profile), else 1.
__device__ __forceinline__ void gen_cell(const SliceStage& S, int n_b, int n_phi, int c,
                                         double& xi, double& ws, double& bw) {
  const int per = n_b * n_phi;
  const int j = c / per, r = c - j * per, l = r / n_phi, q = r - l * n_phi;
  const RowConst& rc = S.rc[j];
  if (rc.valid == 0.0) { xi = 0.0; ws = 0.0; bw = 0.0; return; }
  double kap, gam;
  const double x = S.ug[l] * rc.s;
  if (n_phi == 1) trunc_kg_row(x, rc, kap, gam);
  else kg_eps_row(x, ell_cos2(q), rc, kap, gam);
  cell_jump(kap, gam, rc.warea * S.du2[l] / n_phi, xi, ws);
  bw = rc.beta * ws;
}

// Filament rows of a slice and the filament cell lattice (shape-only part),
// staged in LDS: per u bin the profile factors fk = 2 pi / s and
// fg = 4 pi / (s + 1) - fk (s = sqrt(1 + u^2); kappa = k0 fk, gamma = k0 fg,
// u <= 1 on the whole lattice) and u1^2 - u0^2; per orientation node the
// amplitude factor and weight.
struct FilStage {
  double fr[kMaxMasses][kFilRowStride];
  double fk[kFilU], fg[kFilU], du2[kFilU];
  double kf[kFilPhi], wp[kFilPhi];
};

__device__ inline void stage_fil(FilStage& F, const double* __restrict__ frows, const sgl_shape& sh) {
  for (int q = threadIdx.x; q < sh.n_m * kFilRowStride; q += blockDim.x)
    F.fr[q / kFilRowStride][q % kFilRowStride] = frows[q];
  for (int l = threadIdx.x; l < kFilU; l += blockDim.x) {
    double uc, du2;
    fil_ubin(l, uc, du2);
    const double s = sqrt(1.0 + uc * uc);
    F.fk[l] = 2.0 * kPi / s;
    F.fg[l] = 4.0 * kPi / (s + 1.0) - F.fk[l];
    F.du2[l] = du2;
  }
  for (int n = threadIdx.x; n < kFilPhi; n += blockDim.x) fil_orientation(n, F.kf[n], F.wp[n]);
}

// Filament cell c = (row j, u bin l, orientation n) of the staged slice, FP64.
__device__ __forceinline__ void gen_fil_cell(const FilStage& F, int c, double& xi, double& ws,
                                             double& bw) {
  const int per = kFilU * kFilPhi;
  const int j = c / per, r = c - j * per, l = r / kFilPhi, n = r - l * kFilPhi;
  const double* R = F.fr[j];
  if (R[2] == 0.0) { xi = 0.0; ws = 0.0; bw = 0.0; return; }
  const double k0 = R[0] * F.kf[n] / R[1];
  cell_jump(k0 * F.fk[l], k0 * F.fg[l], R[2] * kPi * R[1] * R[1] * F.du2[l] * F.wp[n], xi, ws);
  bw = R[3] * ws;
}

// ------------------------------------------------------------ exact path --
constexpr int kCellTile = 256;
constexpr int kKpt = 2;  // k nodes per thread

// Body of the exact exponent kernel; the filament variant stages the
// filament rows (FilStage) as well, the halo-only variant leaves that LDS out
// so that three blocks fit per CU.
template <bool kFil>
__device__ __forceinline__ void lnphi_exact_body(const double* __restrict__ rows,
                                                 const double* __restrict__ frows,
                                                 const double* __restrict__ cosmo,
                                                 const double* __restrict__ uedge,
                                                 const double* __restrict__ knodes,
                                                 const sgl_shape& sh, double* __restrict__ part,
                                                 double* __restrict__ amp, SliceStage& S,
                                                 FilStage* FS, double* sxi, double* sws,
                                                 double* sbw) {
  const int t = threadIdx.x;
  const int i = blockIdx.y, b = blockIdx.z;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  const int n_k = sh.n_k;
  const int kbase = blockIdx.x * (256 * kKpt);
  double kv[kKpt], lr[kKpt], li[kKpt], ar[kKpt], ai[kKpt];
#pragma unroll
  for (int q = 0; q < kKpt; ++q) {
    int n = kbase + t + 256 * q;
    kv[q] = (n < n_k) ? knodes[n] : 0.0;
    lr[q] = li[q] = ar[q] = ai[q] = 0.0;
  }
  if (i < n_z) {
    stage_slice(S, rows + ((size_t)b * sh.n_z_max + i) * sh.n_m * kRowStride, uedge, sh);
    constexpr bool fil = kFil;
    if constexpr (kFil) stage_fil(*FS, frows + ((size_t)b * sh.n_z_max + i) * sh.n_m * kFilRowStride, sh);
    const int n_phi = (sh.flags & SGL_FLAG_ELL) ? kEllPhi : 1;
    const int n_halo = sh.n_m * sh.n_b * n_phi;
    const int n_cells = n_halo + (fil ? sh.n_m * kFilU * kFilPhi : 0);
    for (int c0 = 0; c0 < n_cells; c0 += kCellTile) {
      __syncthreads();
      {
        int c = c0 + t;
        double xi = 0.0, ws = 0.0, bw = 0.0;
        if (c < n_halo) gen_cell(S, sh.n_b, n_phi, c, xi, ws, bw);
        else if constexpr (kFil) {
          if (c < n_cells) gen_fil_cell(*FS, c - n_halo, xi, ws, bw);
        }
        sxi[t] = xi; sws[t] = ws; sbw[t] = bw;
      }
      __syncthreads();
      const int nc = min(kCellTile, n_cells - c0);
      for (int c = 0; c < nc; ++c) {
        const double w = sws[c];
        if (w == 0.0) continue;
        const double xi = sxi[c], bw = sbw[c];
#pragma unroll
        for (int q = 0; q < kKpt; ++q) {
          double ph = kv[q] * xi;
          double hs, hc;
          sincos(0.5 * ph, &hs, &hc);
          double sn = 2.0 * hs * hc;
          double cm1 = -2.0 * hs * hs;
          double smp;
          if (fabs(ph) < 1e-2) {
            double p2 = ph * ph;
            smp = -ph * p2 / 6.0 * (1.0 - p2 / 20.0 * (1.0 - p2 / 42.0));
          } else {
            smp = sn - ph;
          }
          lr[q] = fma(w, cm1, lr[q]);
          li[q] = fma(w, smp, li[q]);
          ar[q] = fma(bw, cm1, ar[q]);
          ai[q] = fma(bw, sn, ai[q]);
        }
      }
    }
  }
  const size_t base = ((size_t)b * sh.n_z_max + i) * n_k;
#pragma unroll
  for (int q = 0; q < kKpt; ++q) {
    int n = kbase + t + 256 * q;
    if (n < n_k) {
      part[2 * (base + n)] = lr[q];
      part[2 * (base + n) + 1] = li[q];
      amp[2 * (base + n)] = ar[q];
      amp[2 * (base + n) + 1] = ai[q];
    }
  }
}

// ---------------------------------------------- positive closure cells --
// Cells of group g of the staged slice: halo mass row g (n_b x n_phi cells)
// for g < n_m, else filament row g - n_m (kFilU x kFilPhi cells).
constexpr int kVpWkSplit = 8;

struct GroupCells {
  int first, count;
  bool fil;
};
__device__ __forceinline__ GroupCells group_cells(const sgl_shape& sh, int n_phi, int g) {
  if (g < sh.n_m) return {g * sh.n_b * n_phi, sh.n_b * n_phi, false};
  return {(g - sh.n_m) * kFilU * kFilPhi, kFilU * kFilPhi, true};
}

// wk[b][s][g] = sum over the cells of group g of w xi (the k^2 coefficient
// of the group's jump sum): one wave per group, lanes over the cells; the
// groups of a slice are spread over kVpWkSplit workgroups.
__global__ __launch_bounds__(256) void k_vp_wk(const double* __restrict__ rows,
                                               const double* __restrict__ frows,
                                               const double* __restrict__ cosmo,
                                               const double* __restrict__ uedge, sgl_shape sh,
                                               double* __restrict__ wk) {
  __shared__ SliceStage S;
  __shared__ FilStage FS;
  const int i = blockIdx.x, b = blockIdx.y;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  if (i >= n_z) return;
  const bool fil = (sh.flags & SGL_FLAG_FILAMENTS) != 0;
  stage_slice(S, rows + ((size_t)b * sh.n_z_max + i) * sh.n_m * kRowStride, uedge, sh);
  if (fil) stage_fil(FS, frows + ((size_t)b * sh.n_z_max + i) * sh.n_m * kFilRowStride, sh);
  __syncthreads();
  const int n_phi = (sh.flags & SGL_FLAG_ELL) ? kEllPhi : 1;
  const int G = vp_groups(sh);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int g = blockIdx.z * 4 + wave; g < G; g += 4 * gridDim.z) {
    const GroupCells gc = group_cells(sh, n_phi, g);
    double acc = 0.0;
    for (int l = lane; l < gc.count; l += 64) {
      double xi, ws, bw;
      if (gc.fil) gen_fil_cell(FS, gc.first + l, xi, ws, bw);
      else gen_cell(S, sh.n_b, n_phi, gc.first + l, xi, ws, bw);
      acc = fma(ws, xi, acc);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) wk[((size_t)b * sh.n_z_max + i) * G + g] = acc;
  }
}

// Exact exponent with the positive closure's group sums.  As k_lnphi_exact
// for lam (per-slice partial sums into part), but the cells run group by
// group in the order of the groups' rate buckets, so that after the groups
// with bucket <= j the running sums
//   P1 = sum minv_g Y_g,  P0 = sum tm_g Y_g,  Y_g(k) = sum_{c in g} w_c (e^{i k xi_c} - 1)
// give the rate-weighted sum at latent node j as u_j P1 - P0 (rates
// lambda_g(u_j) = u_j minv_g - tm_g for j >= bucket_g, 0 below); written as
// zpart[j], with ytot = sum_g Y_g, so that sum_g (lambda_g(u_j) - 1) Y_g =
// zpart[j] - ytot.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_lnphi_exact_vp(const double* __restrict__ rows,
                                                        const double* __restrict__ frows,
                                                        const double* __restrict__ cosmo,
                                                        const double* __restrict__ uedge,
                                                        const double* __restrict__ knodes,
                                                        sgl_shape sh, VpLayout V,
                                                        double* __restrict__ vp,
                                                        double* __restrict__ part) {
  __shared__ SliceStage S;
  __shared__ FilStage FS;
  __shared__ double sxi[kCellTile], sws[kCellTile];
  __shared__ double gmv[2 * kMaxMasses], gtm[2 * kMaxMasses];
  __shared__ int gbk[2 * kMaxMasses], order[2 * kMaxMasses];
  const int t = threadIdx.x;
  const int i = blockIdx.y, b = blockIdx.z;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  const int n_k = sh.n_k;
  const int G = vp_groups(sh);
  const int kbase = blockIdx.x * (256 * kKpt);
  double kv[kKpt], lr[kKpt], li[kKpt];
#pragma unroll
  for (int q = 0; q < kKpt; ++q) {
    int n = kbase + t + 256 * q;
    kv[q] = (n < n_k) ? knodes[n] : 0.0;
    lr[q] = li[q] = 0.0;
  }
  const size_t si = (size_t)b * sh.n_z_max + i;
  if (i < n_z) {
    const bool fil = (sh.flags & SGL_FLAG_FILAMENTS) != 0;
    stage_slice(S, rows + si * sh.n_m * kRowStride, uedge, sh);
    if (fil) stage_fil(FS, frows + si * sh.n_m * kFilRowStride, sh);
    const double* tab = vp + V.tab + si * G * 3;
    for (int g = t; g < G; g += 256) {
      gmv[g] = tab[3 * g];
      gtm[g] = tab[3 * g + 1];
      gbk[g] = (int)tab[3 * g + 2];
    }
    __syncthreads();
    // stable order by bucket: rank of g among (bucket, g)
    for (int g = t; g < G; g += 256) {
      const int bg = gbk[g];
      int r = 0;
      for (int h = 0; h < G; ++h) r += (gbk[h] < bg) || (gbk[h] == bg && h < g);
      order[r] = g;
    }
    const int n_phi = (sh.flags & SGL_FLAG_ELL) ? kEllPhi : 1;
    double p1r[kKpt], p1i[kKpt], p0r[kKpt], p0i[kKpt], ytr[kKpt], yti[kKpt];
#pragma unroll
    for (int q = 0; q < kKpt; ++q) p1r[q] = p1i[q] = p0r[q] = p0i[q] = ytr[q] = yti[q] = 0.0;
    double* zp = vp + V.zpart + si * kVpN * (size_t)n_k * 2;
    auto flush = [&](int j) {
#pragma unroll
      for (int q = 0; q < kKpt; ++q) {
        const int n = kbase + t + 256 * q;
        if (n < n_k) {
          const double u = kGhU[j];
          zp[2 * ((size_t)j * n_k + n)] = u * p1r[q] - p0r[q];
          zp[2 * ((size_t)j * n_k + n) + 1] = u * p1i[q] - p0i[q];
        }
      }
    };
    int jw = 0;
    __syncthreads();
    for (int gi = 0; gi < G; ++gi) {
      const int g = order[gi];
      for (; jw < gbk[g] && jw < kVpN; ++jw) flush(jw);
      const GroupCells gc = group_cells(sh, n_phi, g);
      double yr[kKpt], yi[kKpt];
#pragma unroll
      for (int q = 0; q < kKpt; ++q) yr[q] = yi[q] = 0.0;
      for (int c0 = 0; c0 < gc.count; c0 += kCellTile) {
        __syncthreads();
        {
          const int c = c0 + t;
          double xi = 0.0, ws = 0.0, bw;
          if (c < gc.count) {
            if (gc.fil) gen_fil_cell(FS, gc.first + c, xi, ws, bw);
            else gen_cell(S, sh.n_b, n_phi, gc.first + c, xi, ws, bw);
          }
          sxi[t] = xi;
          sws[t] = ws;
        }
        __syncthreads();
        const int nc = min(kCellTile, gc.count - c0);
        for (int c = 0; c < nc; ++c) {
          const double w = sws[c];
          if (w == 0.0) continue;
          const double xi = sxi[c];
#pragma unroll
          for (int q = 0; q < kKpt; ++q) {
            double ph = kv[q] * xi;
            double hs, hc;
            sincos(0.5 * ph, &hs, &hc);
            double sn = 2.0 * hs * hc;
            double cm1 = -2.0 * hs * hs;
            double smp;
            if (fabs(ph) < 1e-2) {
              double p2 = ph * ph;
              smp = -ph * p2 / 6.0 * (1.0 - p2 / 20.0 * (1.0 - p2 / 42.0));
            } else {
              smp = sn - ph;
            }
            lr[q] = fma(w, cm1, lr[q]);
            li[q] = fma(w, smp, li[q]);
            yr[q] = fma(w, cm1, yr[q]);
            yi[q] = fma(w, sn, yi[q]);
          }
        }
      }
      const double mv = gmv[g], tm = gtm[g];
#pragma unroll
      for (int q = 0; q < kKpt; ++q) {
        p1r[q] = fma(mv, yr[q], p1r[q]);
        p1i[q] = fma(mv, yi[q], p1i[q]);
        p0r[q] = fma(tm, yr[q], p0r[q]);
        p0i[q] = fma(tm, yi[q], p0i[q]);
        ytr[q] += yr[q];
        yti[q] += yi[q];
      }
    }
    for (; jw < kVpN; ++jw) flush(jw);
    double* yt = vp + V.ytot + si * (size_t)n_k * 2;
#pragma unroll
    for (int q = 0; q < kKpt; ++q) {
      const int n = kbase + t + 256 * q;
      if (n < n_k) {
        yt[2 * n] = ytr[q];
        yt[2 * n + 1] = yti[q];
      }
    }
  }
  const size_t base = si * n_k;
#pragma unroll
  for (int q = 0; q < kKpt; ++q) {
    int n = kbase + t + 256 * q;
    if (n < n_k) {
      part[2 * (base + n)] = lr[q];
      part[2 * (base + n) + 1] = li[q];
    }
  }
}


__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_lnphi_exact(
    const double* __restrict__ rows, const double* __restrict__ frows,
    const double* __restrict__ cosmo, const double* __restrict__ uedge,
    const double* __restrict__ knodes, sgl_shape sh, double* __restrict__ part,
    double* __restrict__ amp) {
  __shared__ SliceStage S;
  __shared__ double sxi[kCellTile], sws[kCellTile], sbw[kCellTile];
  lnphi_exact_body<false>(rows, frows, cosmo, uedge, knodes, sh, part, amp, S, nullptr, sxi, sws,
                          sbw);
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_lnphi_exact_fil(
    const double* __restrict__ rows, const double* __restrict__ frows,
    const double* __restrict__ cosmo, const double* __restrict__ uedge,
    const double* __restrict__ knodes, sgl_shape sh, double* __restrict__ part,
    double* __restrict__ amp) {
  __shared__ SliceStage S;
  __shared__ FilStage FS;
  __shared__ double sxi[kCellTile], sws[kCellTile], sbw[kCellTile];
  lnphi_exact_body<true>(rows, frows, cosmo, uedge, knodes, sh, part, amp, S, &FS, sxi, sws, sbw);
}

// lam[b][n] = sum over slices of part[b][i][n], in slice order.
__global__ void k_sum_slices(const double* __restrict__ part, int B, int n_z_max, int n_k,
                             double* __restrict__ lam) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)B * n_k) return;
  const int n = (int)(gid % n_k);
  const int b = (int)(gid / n_k);
  double re = 0.0, im = 0.0;
  for (int i = 0; i < n_z_max; ++i) {
    size_t o = 2 * (((size_t)b * n_z_max + i) * n_k + n);
    re += part[o];
    im += part[o + 1];
  }
  lam[2 * gid] = re;
  lam[2 * gid + 1] = im;
}

// ----------------------------------------------------------- binned path --
// Fine histogram of a slice in LDS, moment-major (a bin's moment at an 8-byte
// stride, so the lanes of a wave spread over all banks), four moments per
// fine bin f, each a 64-bit fixed-point sum:
//   A = sum ws g_f,  Bm = sum ws g_f d,  C = sum ws g_f d^2,  E = sum bw g_f,
// ws the tilted weight w e^{-xi}, bw = beta ws the two-point weight, d the
// cell's place in the bin (|d| <= 0.99, below), g_f a tilt bound: 1 on the
// positive side, where ws <= w, and a power of two <= e^{-|xi|_max} on the
// negative side, where ws = w e^{|xi|}; so ws g_f <= w.
//
// Why fixed point: an FP64 sum of atomics depends on the order in which the
// lanes and waves arrive, which changes from run to run.  Here every deposit
// is rounded once, to an integer on a per-slice power-of-two scale, and the
// integer LDS adds are associative: the histogram, and with it every PDF of
// the batched path, is the same on every run and in every graph replay.  The
// scales are S_w = 2^(49 - ilogb(W)) and S_b = 2^(49 - ilogb(D)), W a bound
// on the slice's sum of w (u_max^2 sum_j warea_j over its mass rows) and D the
// same with |beta_j| warea_j, so every deposit and every sum stays below
// 2^51 in magnitude: one FP64 fma with 1.5 * 2^52 rounds a deposit to the
// integer in its low mantissa bits.  A deposit is resolved to 2^-50 W
// (~1e-15 of the slice's weight; the strong-lensing bins of a slice, xi >
// 0.1, hold ~1e-6 of it).  The flush turns the sums back into the FP64
// moments the rest of the path reads: w_f = A / (S_w g_f), the centroid from
// Bm / A and the in-bin variance from C / A - (Bm / A)^2 -- in centred form,
// free of the cancellation of (sum w xi^2) / w - m^2 -- and
// beta w = E / (S_b g_f).  A cell on the negative side beyond the last bin
// (xi < -30, detA > e^30) is dropped, as the positive side drops xi > 30.
//
// The slice's coarse two-point moments are formed at the flush, kCoarse fine
// bins per coarse bin, each fine bin's beta w placed at the bin's w centroid
// m_f with its w variance v_f: sum beta w, sum beta w m_f,
// sum beta w (m_f^2 + v_f).  Within one fine bin (3.1 % relative width) beta
// and xi are taken as uncorrelated, which moves a coarse record's centroid by
// at most 1.6 % of xi.  A cell costs four 64-bit integer LDS atomics on fine
// bins, which neighbouring lanes rarely share.
constexpr int kFineMom = 4;
struct FineHist {
  unsigned long long d[kFineMom * kFineBins];
};

// Place of a cell in its bin: d = |xi| / base_f * C2 - C1 C2, with base_f = 2^e
// a power of two taken from the bin index alone -- 2^floor(log2 lo_f) for bin
// q >= 1 of a side, whose |xi| / base_f then lies in [1, 2R) (R = hi / lo,
// the bin ratio, at most kNegR), and 2^ceil(log2 hi_0) for bin 0, which
// reaches down to 0 and lies in (0, 1] -- and C1 = kNegR, C2 = 1 / (1.01
// kNegR) the centre and reciprocal half-width of [0, 2 kNegR) widened by 1 %,
// so |d| <= 0.99 on every bin with room for the FP32 log's choice at an
// edge.  Scaling by a power of two is exact and needs no transcendental; one
// range for all bins keeps per-bin constants out of the cell loop (the
// in-bin spread of d is then narrower than [-1, 1], by ~40x on the positive
// side, which the 2^-50 resolution absorbs).  The tilt bound of negative bin
// q is g = 2^-k with k = max(1, 2^(e + 2)) >= |xi|_max log2 e, as
// |xi|_max < 2 kNegR 2^e < 2^(e + 2) / log2 e.
constexpr float kPosScale = (float)(kFinePos / 31.032218497590705);   // ln(30 / 1e-12)
constexpr float kNegScale = (float)(kFineNeg / 21.82187812561452);    // ln(30 / 1e-8)
constexpr float kLnLo = -27.631021115928547f;                          // ln 1e-12
constexpr float kLnNegLo = -18.420680743952367f;                       // ln 1e-8
// log2 of the lowest edge (+1e-4, so floor() of an edge that rounds a hair
// low in FP32 still gives its own exponent) and the log2 step of each side
constexpr float kPosLo2 = (float)(-27.631021115928547 * 1.4426950408889634 + 1e-4);
constexpr float kPosStep2 = (float)(31.032218497590705 / kFinePos * 1.4426950408889634);
constexpr float kNegLo2 = (float)(-18.420680743952367 * 1.4426950408889634 + 1e-4);
constexpr float kNegStep2 = (float)(21.82187812561452 / kFineNeg * 1.4426950408889634);
constexpr double kNegR = 1.185877993301057;    // e^(21.82187812561452 / 128)
constexpr int kPosE0 = -39;                    // ceil(log2(1e-12 e^(31.03 / 1024)))
constexpr int kNegE0 = -26;                    // ceil(log2(1e-8 kNegR))
constexpr double kPlaceC2 = 1.0 / (1.01 * kNegR);
constexpr double kPlaceC12 = kNegR * kPlaceC2;

// A cell's fine bin f, its base exponent e and tilt shift k (g = 2^-k).
struct FineSlot {
  int f, e, k;
};
__device__ __forceinline__ int pos_bin_exp(int q) {
  return (q == 0) ? kPosE0 : (int)floorf(fmaf((float)q, kPosStep2, kPosLo2));
}
__device__ __forceinline__ int neg_bin_exp(int q) {
  return (q == 0) ? kNegE0 : (int)floorf(fmaf((float)q, kNegStep2, kNegLo2));
}
__device__ __forceinline__ int neg_tilt(int e) { return (e >= -2) ? (1 << min(e + 2, 30)) : 1; }
// Bin from xi with the FP32 hardware log: which of two neighbouring bins an
// edge case lands in does not matter, every bin keeps its exact centroid and
// variance.
__device__ __forceinline__ FineSlot fine_slot(double xi) {
  FineSlot s;
  if (xi > 0.0) {
    const float lx = log_hw(fmaxf((float)xi, 1e-37f));
    const int q = min(max((int)((lx - kLnLo) * kPosScale), 0), kFinePos - 1);
    s.f = q;
    s.e = pos_bin_exp(q);
    s.k = 0;
  } else {
    const float lx = log_hw(fmaxf((float)(-xi), 1e-37f));
    const int q = min(max((int)((lx - kLnNegLo) * kNegScale), 0), kFineNeg - 1);
    s.f = kFinePos + q;
    s.e = neg_bin_exp(q);
    s.k = neg_tilt(s.e);
  }
  return s;
}
// The same constants from the bin index (the flush).
__device__ __forceinline__ FineSlot bin_slot(int f) {
  FineSlot s;
  s.f = f;
  if (f < kFinePos) {
    s.e = pos_bin_exp(f);
    s.k = 0;
  } else {
    s.e = neg_bin_exp(f - kFinePos);
    s.k = neg_tilt(s.e);
  }
  return s;
}

__device__ __forceinline__ void fine_hist_clear(FineHist& H, int t, int nt) {
  for (int q = t; q < kFineMom * kFineBins; q += nt) H.d[q] = 0ull;
}

// x (|x| < 2^51) rounded to the nearest integer, as the two's-complement bits
// the LDS adds accumulate: 1.5 * 2^52 + x holds the integer in its low
// mantissa bits.
constexpr double kMagic = 6755399441055744.0;            // 1.5 * 2^52
__device__ __forceinline__ unsigned long long magic_bits(double t) {
  return (unsigned long long)__double_as_longlong(t) -
         (unsigned long long)__double_as_longlong(kMagic);
}

// One cell into H: sw, sb the log2 of the slice's scales S_w, S_b.
__device__ __forceinline__ void deposit4(FineHist& H, double xi, double ws, double bw, int sw,
                                         int sb) {
  if (xi < -kXiNegHi) return;
  const FineSlot s = fine_slot(xi);
  const double d = fma(fabs(xi), ldexp(kPlaceC2, -s.e), -kPlaceC12);
  const double a = ldexp(ws, sw - s.k);          // ws g S_w
  const double ad = a * d;
  unsigned long long* q = H.d + s.f;
  atomicAdd(q, magic_bits(a + kMagic));
  atomicAdd(q + kFineBins, magic_bits(ad + kMagic));
  atomicAdd(q + 2 * kFineBins, magic_bits(fma(ad, d, kMagic)));
  atomicAdd(q + 3 * kFineBins, magic_bits(ldexp(bw, sb - s.k) + kMagic));
}

// log2 of the power-of-two scale 2^(49 - ilogb(bound)) of a slice's
// fixed-point sums (0 for an empty slice).
__device__ __forceinline__ int fixed_scale(double bound) {
  return (bound > 0.0) ? 49 - ilogb(bound) : 0;
}

// Sum of v over a wave, the same value in every lane (a butterfly: each step
// adds the same two partial sums in either lane of a pair).
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum of v over the workgroup, the same value in every thread and on every
// run (butterfly within each wave, then the waves in order).  red holds one
// double per wave.
__device__ __forceinline__ double block_sum(double v, double* red, int t, int nt) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < nt / 64; ++w) s += red[w];
  return s;
}

// Bin f's FP64 moments from the fixed-point sums: w, centroid m, in-bin
// variance v, two-point weight bw (w = 0: empty bin).
__device__ __forceinline__ void fine_moments(const FineHist& H, int f, int sw, int sb, double& w,
                                             double& m, double& v, double& bw) {
  const long long A = (long long)H.d[f];
  if (A == 0) { w = m = v = bw = 0.0; return; }
  const FineSlot s = bin_slot(f);
  const double a = (double)A, ra = 1.0 / a;
  const double e1 = (double)(long long)H.d[kFineBins + f] * ra;
  const double e2 = (double)(long long)H.d[2 * kFineBins + f] * ra;
  const double h = ldexp(1.0 / kPlaceC2, s.e);  // |xi| per unit of d
  const double mx = (kPlaceC12 + e1) * h;
  w = ldexp(a, s.k - sw);
  m = (f >= kFinePos) ? -mx : mx;
  v = fmax(e2 - e1 * e1, 0.0) * h * h;
  bw = ldexp((double)(long long)H.d[3 * kFineBins + f], s.k - sb);
}

// Add (or store) the slice's fine partials (w, w m, w (m^2 + v)) and coarse
// moments from H.  Each bin's FP64 moments are formed once and left in H in
// place of its sums (the bits of w, m, v, beta w), which the coarse pass then
// reads.
__device__ inline void fine_hist_flush(FineHist& H, int sw, int sb, double* __restrict__ P,
                                       double* __restrict__ Q, bool add, int t, int nt) {
  auto put = [&](int q, double x) { H.d[q] = (unsigned long long)__double_as_longlong(x); };
  auto get = [&](int q) { return __longlong_as_double((long long)H.d[q]); };
  for (int f = t; f < kFineBins; f += nt) {
    double w, m, v, bw;
    fine_moments(H, f, sw, sb, w, m, v, bw);
    put(f, w);
    put(kFineBins + f, m);
    put(2 * kFineBins + f, v);
    put(3 * kFineBins + f, bw);
    const double wx = w * m, wxx = w * (m * m + v);
    if (add) {
      if (w == 0.0) continue;
      P[f * kMoments] += w;
      P[f * kMoments + 1] += wx;
      P[f * kMoments + 2] += wxx;
    } else {
      P[f * kMoments] = w;
      P[f * kMoments + 1] = wx;
      P[f * kMoments + 2] = wxx;
    }
  }
  __syncthreads();
  for (int cb = t; cb < kSliceBins; cb += nt) {
    double m0 = 0.0, m1 = 0.0, m2 = 0.0;
    for (int f = cb * kCoarse; f < (cb + 1) * kCoarse; ++f) {
      if (get(f) == 0.0) continue;
      const double m = get(kFineBins + f), v = get(2 * kFineBins + f), bw = get(3 * kFineBins + f);
      m0 += bw;
      m1 += bw * m;
      m2 += bw * (m * m + v);
    }
    if (add) {
      Q[cb * 3] += m0;
      Q[cb * 3 + 1] += m1;
      Q[cb * 3 + 2] += m2;
    } else {
      Q[cb * 3] = m0;
      Q[cb * 3 + 1] = m1;
      Q[cb * 3 + 2] = m2;
    }
  }
}

// Per impact-parameter column (shared by every instance): geometric bin
// centre u_g in FP32 with its log and reciprocal, and u1^2 - u0^2 in FP64.
struct ColF {
  float ug, lug, rug, pad;
  double du2, pad2;
};
static_assert(sizeof(ColF) == 32, "ColF is one 32-byte record");

__global__ void k_cols(const double* __restrict__ uedge, sgl_shape sh, int B,
                       double* __restrict__ hist) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= sh.n_b) return;
  const double u0 = uedge[l], u1 = uedge[l + 1];
  const double ug = sqrt(u0 * u1);
  ColF c;
  c.ug = (float)ug;
  c.lug = (float)log(ug);
  c.rug = (float)(1.0 / ug);
  c.pad = 0.0f;
  c.du2 = u1 * u1 - u0 * u0;
  c.pad2 = 0.0;
  reinterpret_cast<ColF*>(hist + hist_layout(sh, B).cols)[l] = c;
}

// One thread per (instance, slice, mass) row: its RowF record.
__global__ __launch_bounds__(256) void k_rowf(const double* __restrict__ rows,
                                              const double* __restrict__ cosmo,
                                              const double* __restrict__ uedge, sgl_shape sh,
                                              int B, double* __restrict__ hist) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t nrow = (size_t)B * sh.n_z_max * sh.n_m;
  if (gid >= nrow) return;
  const int b = (int)(gid / ((size_t)sh.n_z_max * sh.n_m));
  const int i = (int)((gid / sh.n_m) % sh.n_z_max);
  if (i >= (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10]) return;
  reinterpret_cast<RowF*>(hist + hist_layout(sh, B).rowf)[gid] =
      make_rowf(row_const(rows + gid * kRowStride));
}

// Binning pass: every cell, FP32 lens profile (trunc_kg_row_f32), FP32 log
// for the jump.  A thread owns one mass row (its RowF record in registers) and
// walks the impact-parameter columns; the lanes of a wave are consecutive
// masses at the same column, which keeps the x >= x_t and x < 1 branches
// nearly uniform (x / x_t = u r_vir / r_t hardly depends on the mass) and
// spreads the lanes over more histogram bins than consecutive columns do.
// bin_slice takes one slice (instance b, slice i) into the workgroup's LDS
// histogram, then into its fine partials and coarse moments.
template <int kBinThreads, bool kEll>
__device__ __forceinline__ void bin_slice(const double* __restrict__ cosmo,
                                          const double* __restrict__ uedge,
                                          const double* __restrict__ rows, const sgl_shape& sh,
                                          const HistLayout& HL, int b, int i, FineHist& H,
                                          double* __restrict__ hist) {
  const int t = threadIdx.x;
  fine_hist_clear(H, t, kBinThreads);
  __syncthreads();
  const RowF* RF = reinterpret_cast<const RowF*>(hist + HL.rowf) +
                   ((size_t)b * sh.n_z_max + i) * sh.n_m;
  const ColF* CF = reinterpret_cast<const ColF*>(hist + HL.cols);
  const int n_b = sh.n_b;
  // fixed-point scales of the slice: sum_l du2_l <= u_max^2 bounds a row's
  // weight by warea u_max^2 (with ellipticity, its kEllPhi orientations
  // share warea).  Every wave sums all the rows itself, in the same order, so
  // each holds the same bound without a barrier.
  double sw = 0.0, sb = 0.0;
  for (int j = t & 63; j < sh.n_m; j += 64) {
    const double wa = fabs(RF[j].warea);
    sw += wa;
    sb += fabs(RF[j].beta) * wa;
  }
  sw = wave_sum(sw);
  sb = wave_sum(sb);
  const double umax2 = uedge[n_b] * uedge[n_b];
  const int lsw = fixed_scale(umax2 * sw);
  const int lsb = fixed_scale(umax2 * sb);
  const int rpp = min(sh.n_m, kBinThreads);     // rows per pass
  const int parts = kBinThreads / rpp;          // column phases
  const int tj = t % rpp, part = t / rpp;
  auto cell = [&](const RowF& rc, const ColF& cf) {
    float kf, gf;
    trunc_kg_row_f32(cf.ug * rc.s, cf.lug + rc.log_s, cf.rug * rc.inv_s, rc, kf, gf);
    double xi, ws;
    cell_jump_f32log(kf, gf, rc.warea * cf.du2, xi, ws);
    if (ws != 0.0) deposit4(H, xi, ws, rc.beta * ws, lsw, lsb);
  };
  // Halo ellipticity: orientation node q of a row is the row record with its
  // x scaled to the elliptical radius, x_e = x f, f^2 = 1 - eps cos 2phi_q
  // (and 1/kEllPhi of the weight); the pseudo-elliptical kappa, gamma of
  // kg_eps_row follow from the circular profile at x_e.
  auto cell_ell = [&](const RowF& rq, const ColF& cf, float eps, float c2, float f) {
    const float xe = cf.ug * rq.s, rxe = cf.rug * rq.inv_s;
    float k0, g0, kf, gf;
    trunc_kg_row_f32(xe, cf.lug + rq.log_s, rxe, rq, k0, g0);
    if (xe >= rq.xt) {                           // monopole shear at the circular x
      const float rx = rxe * f;
      kf = 0.0f;
      gf = rq.mut4ks * rx * rx;
    } else {
      if (xe < 1e-4f) g0 = rq.ks;
      const float cg = c2 * g0;
      kf = k0 + eps * cg;
      gf = sqrt_hw(fmaxf(g0 * g0 + 2.0f * eps * cg * k0 + eps * eps * (k0 * k0 - cg * cg), 0.0f));
    }
    double xi, ws;
    cell_jump_f32log(kf, gf, rq.warea * cf.du2, xi, ws);
    if (ws != 0.0) deposit4(H, xi, ws, rq.beta * ws, lsw, lsb);
  };
  // every cell of row j: cols(visit) walks this thread's columns
  auto row_cells = [&](const RowF& rc, int j, auto&& cols) {
    if constexpr (!kEll) {
      cols([&](const ColF& cf) { cell(rc, cf); });
    } else {
      const float eps =
          (float)rows[(((size_t)b * sh.n_z_max + i) * sh.n_m + j) * kRowStride + 6];
      for (int q = 0; q < kEllPhi; ++q) {
        const float cos2 = (float)ell_cos2(q), f2 = 1.0f - eps * cos2;
        const float f = sqrt_hw(f2), c2 = (cos2 - eps) * rcp_hw(f2);
        RowF rq = rc;
        rq.s = rc.s * f;
        rq.log_s = rc.log_s + 0.5f * log_hw(f2);
        rq.inv_s = rc.inv_s * rsq_hw(f2);
        rq.warea = rc.warea * (1.0 / kEllPhi);
        cols([&](const ColF& cf) { cell_ell(rq, cf, eps, c2, f); });
      }
    }
  };
  if (rpp % 64 == 0) {
    // a wave's lanes share the column phase: the column index is wave-uniform
    // and the column record comes through the scalar cache
    const int wpart = __builtin_amdgcn_readfirstlane(part);
    for (int j = tj; j < sh.n_m; j += rpp) {
      const RowF rc = RF[j];
      if (rc.warea == 0.0) continue;
      row_cells(rc, j, [&](auto&& visit) {
        for (int l = wpart; l < n_b; l += parts) visit(CF[l]);
      });
    }
  } else if (part < parts) {
    for (int j = tj; j < sh.n_m; j += rpp) {
      const RowF rc = RF[j];
      if (rc.warea == 0.0) continue;
      row_cells(rc, j, [&](auto&& visit) {
        ColF nxt = CF[min(part, n_b - 1)];
        for (int l = part; l < n_b; l += parts) {
          const ColF cf = nxt;                   // the next column is fetched a cell ahead
          nxt = CF[min(l + parts, n_b - 1)];
          visit(cf);
        }
      });
    }
  }
  __syncthreads();
  fine_hist_flush(H, lsw, lsb,
                  hist + HL.partial + ((size_t)b * sh.n_z_max + i) * kFineBins * kMoments,
                  hist + HL.cpart + ((size_t)b * sh.n_z_max + i) * kSliceBins * 3, false, t,
                  kBinThreads);
}

// Two schedules of the binning pass, chosen by the SGL_TUNE_BIN_SCHEDULE knob
// (the autotuner times both; each slice is binned whole by one workgroup
// either way, so both give the same bits):
// - k_bin_cells: one workgroup per (instance, slice); workgroups of slices
//   past the instance's own n_z return at once;
// - k_bin_cells_queue: persistent -- as many workgroups as the device holds
//   at once, each taking (instance, slice) items from the work queue that
//   k_bin_items wrote (one item per slice below the instance's n_z) until it
//   is empty.
// On MI355X the two tie at B = 4096 (73.7 and 74.1 ms; the autotuner's timings
// of sgl_bin_cells differ by 0.1 %), and the first wins on the surrogate's
// cheap lattice of 16 cells per thread (8.5 against 14.6 ms): there the
// per-item queue fetch and its barrier cost more than the hardware
// dispatcher's refill of finished workgroups.
template <int kBinThreads, bool kEll>
__global__ __launch_bounds__(kBinThreads) __attribute__((amdgpu_waves_per_eu(kBinThreads == 256 ? 4 : (kEll ? 7 : 8)))) void k_bin_cells(const double* __restrict__ cosmo,
                                                           const double* __restrict__ uedge,
                                                           const double* __restrict__ rows,
                                                           sgl_shape sh, double* __restrict__ hist) {
  __shared__ FineHist H;
  const int i = blockIdx.x, b = blockIdx.y;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  if (i >= n_z) return;                         // slices past n_z are never read
  bin_slice<kBinThreads, kEll>(cosmo, uedge, rows, sh, hist_layout(sh, gridDim.y), b, i, H, hist);
}

template <int kBinThreads, bool kEll>
__global__ __launch_bounds__(kBinThreads) __attribute__((amdgpu_waves_per_eu(kBinThreads == 256 ? 4 : (kEll ? 7 : 8)))) void k_bin_cells_queue(
    const double* __restrict__ cosmo, const double* __restrict__ uedge,
    const double* __restrict__ rows, sgl_shape sh, int B, double* __restrict__ hist) {
  __shared__ FineHist H;
  __shared__ int s_item;
  const int t = threadIdx.x;
  const HistLayout HL = hist_layout(sh, B);
  int* queue = reinterpret_cast<int*>(hist + HL.items);
  const int n_items = queue[1];                 // written by k_bin_items before this kernel
  for (;;) {
    if (t == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int item = s_item;
    if (item >= n_items) break;                 // uniform: every thread read the same item
    const int packed = queue[2 + item];
    bin_slice<kBinThreads, kEll>(cosmo, uedge, rows, sh, HL, packed >> 7, packed & 127, H, hist);
    __syncthreads();                            // the flush is done before the next item
  }
}

// Work items of the queue schedule: item b * 128 + i for every slice
// i < n_z[b] (n_z <= 96), in instance order (exclusive scan of the slice
// counts in one workgroup), and the queue counter reset -- by a kernel, so a
// captured graph replays it.
constexpr int kItemThreads = 1024;
__global__ __launch_bounds__(kItemThreads) void k_bin_items(const double* __restrict__ cosmo,
                                                            sgl_shape sh, int B,
                                                            double* __restrict__ hist) {
  __shared__ int scan[kItemThreads];
  const int t = threadIdx.x;
  const HistLayout HL = hist_layout(sh, B);
  int* queue = reinterpret_cast<int*>(hist + HL.items);
  const int per = (B + kItemThreads - 1) / kItemThreads;
  const int b0 = min(B, t * per), b1 = min(B, b0 + per);
  int cnt = 0;
  for (int b = b0; b < b1; ++b)
    cnt += min(max((int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10], 0), sh.n_z_max);
  scan[t] = cnt;
  __syncthreads();
  for (int o = 1; o < kItemThreads; o <<= 1) {  // inclusive Hillis-Steele scan
    const int v = (t >= o) ? scan[t - o] : 0;
    __syncthreads();
    scan[t] += v;
    __syncthreads();
  }
  int pos = scan[t] - cnt;
  for (int b = b0; b < b1; ++b) {
    const int nz = min(max((int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10], 0), sh.n_z_max);
    for (int i = 0; i < nz; ++i) queue[2 + pos++] = b * 128 + i;
  }
  if (t == kItemThreads - 1) {
    queue[0] = 0;
    queue[1] = scan[t];
  }
}

// Filament pass (SGL_FLAG_FILAMENTS): every filament cell of the slice --
// kappa = k0 fk(u), gamma = k0 fg(u) in FP32 (no cancellation anywhere on the
// cylinder profile), FP64 weights and detA, fixed-point accumulation -- into a
// moment-major LDS histogram, then added into the slice's fine partials and
// coarse moments (the main pass has written them; this workgroup is the only
// one touching the slice now).  Lanes run over orientation nodes and u bins
// of one row, whose jumps spread over many bins.
constexpr int kFilThreads = 1024;
__global__ __launch_bounds__(kFilThreads) void k_bin_fil(const double* __restrict__ frows,
                                                         const double* __restrict__ cosmo,
                                                         sgl_shape sh, double* __restrict__ hist) {
  __shared__ FineHist H;
  __shared__ float sfk[kFilU], sfg[kFilU], skf[kFilPhi];
  __shared__ double sdu2[kFilU], swp[kFilPhi];
  const int t = threadIdx.x, i = blockIdx.x, b = blockIdx.y;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  if (i >= n_z) return;
  const HistLayout HL = hist_layout(sh, gridDim.y);
  fine_hist_clear(H, t, kFilThreads);
  if (t < kFilU) {
    double uc, du2;
    fil_ubin(t, uc, du2);
    const double sq = sqrt(1.0 + uc * uc);
    sfk[t] = (float)(2.0 * kPi / sq);
    sfg[t] = (float)(4.0 * kPi / (sq + 1.0) - 2.0 * kPi / sq);
    sdu2[t] = du2;
  }
  if (t < kFilPhi) {
    double kf, wp;
    fil_orientation(t, kf, wp);
    skf[t] = (float)kf;
    swp[t] = wp;
  }
  __syncthreads();
  const double* FR = frows + ((size_t)b * sh.n_z_max + i) * sh.n_m * kFilRowStride;
  const int per = kFilU * kFilPhi;
  const int n_cells = sh.n_m * per;
  // fixed-point scales: a row's cells weigh wrow pi r^2 sum_l du2_l sum_n wp_n
  __shared__ double red[kFilThreads / 64];
  double sw = 0.0, sb = 0.0;
  for (int j = t; j < sh.n_m; j += kFilThreads) {
    const double* R = FR + (size_t)j * kFilRowStride;
    const double wr = fabs(R[2]) * kPi * R[1] * R[1];
    sw += wr;
    sb += fabs(R[3]) * wr;
  }
  const double su = block_sum(t < kFilU ? sdu2[t] : 0.0, red, t, kFilThreads);
  const double sp = block_sum(t < kFilPhi ? swp[t] : 0.0, red, t, kFilThreads);
  const int lsw = fixed_scale(su * sp * block_sum(sw, red, t, kFilThreads));
  const int lsb = fixed_scale(su * sp * block_sum(sb, red, t, kFilThreads));
  for (int c = t; c < n_cells; c += kFilThreads) {
    const int j = c / per, r = c - j * per, l = r / kFilPhi, n = r - l * kFilPhi;
    const double* R = FR + (size_t)j * kFilRowStride;
    const double wrow = R[2];
    if (wrow == 0.0) continue;
    const float k0 = (float)(R[0] / R[1]) * skf[n];
    double xi, ws;
    cell_jump_f32log(k0 * sfk[l], k0 * sfg[l], wrow * kPi * R[1] * R[1] * sdu2[l] * swp[n], xi, ws);
    if (ws == 0.0) continue;
    deposit4(H, xi, ws, R[3] * ws, lsw, lsb);
  }
  __syncthreads();
  fine_hist_flush(H, lsw, lsb,
                  hist + HL.partial + ((size_t)b * sh.n_z_max + i) * kFineBins * kMoments,
                  hist + HL.cpart + ((size_t)b * sh.n_z_max + i) * kSliceBins * 3, true, t,
                  kFilThreads);
}

// Per instance: add the slice partials in slice order and compact the
// non-empty fine bins; compact every slice's coarse bins.
__global__ __launch_bounds__(256) void k_reduce_fine(const double* __restrict__ cosmo,
                                                     sgl_shape sh, double* __restrict__ hist) {
  __shared__ int flag[kFineBins];
  __shared__ int pos[kFineBins];
  __shared__ double acc[kFineBins * kMoments];
  const int b = blockIdx.x, t = threadIdx.x;
  const HistLayout HL = hist_layout(sh, gridDim.x);
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  for (int q = t; q < kFineBins * kMoments; q += 256) {
    double s = 0.0;
    for (int i = 0; i < n_z; ++i)
      s += hist[HL.partial + ((size_t)b * sh.n_z_max + i) * kFineBins * kMoments + q];
    acc[q] = s;
  }
  __syncthreads();
  for (int f = t; f < kFineBins; f += 256) flag[f] = (acc[f * kMoments] != 0.0) ? 1 : 0;
  __syncthreads();
  if (t == 0) {
    int n = 0;
    for (int f = 0; f < kFineBins; ++f) { pos[f] = n; n += flag[f]; }
    hist[HL.fcnt + b] = (double)n;
  }
  __syncthreads();
  double* R = hist + HL.frec + (size_t)b * kFineBins * kRec;
  for (int f = t; f < kFineBins; f += 256) {
    if (!flag[f]) continue;
    const double w = acc[f * kMoments];
    const double m = acc[f * kMoments + 1] / w;
    const double v = acc[f * kMoments + 2] / w - m * m;
    double* r = R + (size_t)pos[f] * kRec;
    r[0] = m * (0.25 / kPi);
    r[1] = m;
    r[2] = w;
    r[3] = (v > 0.0) ? v : 0.0;
  }
  double* FP = hist + HL.fpos + (size_t)b * (kFineBins + 1);
  for (int f = t; f < kFineBins; f += 256) FP[f] = (double)pos[f];
  if (t == 0) FP[kFineBins] = hist[HL.fcnt + b];
  // prefix sums of w mu_p (p = 2 .. 9; mu_p the raw moments of N(m, v)) over
  // the positive records in bin order: each thread sums a chunk of
  // kFinePos / 256 bins, the chunk totals are scanned in order by kTay
  // threads, then every record R gets the sum over records 0 .. R-1 (entry
  // np is the total).
  {
    constexpr int kChunk = kFinePos / 256;
    static_assert(kFinePos % 256 == 0 && 257 * kTay <= kFineBins * kMoments, "scan fits in acc");
    double* csum = acc + 0;                      // acc is free again: reuse as [256][kTay]
    double mom[kTay];
    for (int p = 0; p < kTay; ++p) mom[p] = 0.0;
    double wm[kChunk], mm[kChunk], vm[kChunk];
    for (int q = 0; q < kChunk; ++q) {
      const int f = t * kChunk + q;
      wm[q] = flag[f] ? acc[f * kMoments] : 0.0;
      mm[q] = flag[f] ? acc[f * kMoments + 1] / wm[q] : 0.0;
      const double v = flag[f] ? acc[f * kMoments + 2] / wm[q] - mm[q] * mm[q] : 0.0;
      vm[q] = (v > 0.0) ? v : 0.0;
    }
    __syncthreads();                             // every thread has read acc
    // w mu_p for p = 2 .. 9 by mu_{p+1} = m mu_p + p v mu_{p-1}
    auto gmoments = [&](int q, double* out) {
      double mu_prev = 1.0, mu = mm[q];
      for (int p = 1; p <= kTay + 1; ++p) {
        const double nx = mm[q] * mu + p * vm[q] * mu_prev;
        mu_prev = mu;
        mu = nx;                                  // mu_{p+1}
        if (p >= 1) out[p - 1] = wm[q] * mu;      // out[0] = w mu_2
      }
    };
    for (int q = 0; q < kChunk; ++q) {
      double g[kTay + 1];
      gmoments(q, g);
      for (int p = 0; p < kTay; ++p) mom[p] += g[p];
    }
    for (int p = 0; p < kTay; ++p) csum[t * kTay + p] = mom[p];
    __syncthreads();
    if (t < kTay) {
      double run = 0.0;
      for (int c = 0; c < 256; ++c) {
        const double v = csum[c * kTay + t];
        csum[c * kTay + t] = run;                // exclusive prefix of the chunk totals
        run += v;
      }
      csum[256 * kTay + t] = run;
    }
    __syncthreads();
    double* PF = hist + HL.fpfx + (size_t)b * (kFinePos + 1) * kTay;
    for (int p = 0; p < kTay; ++p) mom[p] = csum[t * kTay + p];
    for (int q = 0; q < kChunk; ++q) {
      const int f = t * kChunk + q;
      if (!flag[f]) continue;
      double* e = PF + (size_t)pos[f] * kTay;
      double g[kTay + 1];
      gmoments(q, g);
      for (int p = 0; p < kTay; ++p) { e[p] = mom[p]; mom[p] += g[p]; }
    }
    if (t < kTay) PF[(size_t)pos[kFinePos] * kTay + t] = csum[256 * kTay + t];
  }
  // coarse records, one slice per thread (72 bins each, order kept), with
  // the bin -> first-record map and the prefix Gaussian moments of the
  // positive records
  for (int i = t; i < sh.n_z_max; i += 256) {
    const size_t si = (size_t)b * sh.n_z_max + i;
    const double* Q = hist + HL.cpart + si * kSliceBins * 3;
    double* Rs = hist + HL.srec + si * kSliceBins * kRec;
    double* CPo = hist + HL.cpos + si * (kCoarsePos + 1);
    double* CPf = hist + HL.cpfx + si * (kCoarsePos + 1) * kTayC;
    double run[kTayC];
    for (int p = 0; p < kTayC; ++p) run[p] = 0.0;
    int n = 0;
    if (i < n_z) {
      for (int cb = 0; cb < kSliceBins; ++cb) {
        if (cb <= kCoarsePos) {
          CPo[cb] = (double)n;
          if (cb == kCoarsePos)
            for (int p = 0; p < kTayC; ++p) CPf[(size_t)n * kTayC + p] = run[p];
        }
        const double w = Q[cb * 3];
        if (w == 0.0) continue;
        const double m = Q[cb * 3 + 1] / w;
        double v = Q[cb * 3 + 2] / w - m * m;
        v = (v > 0.0) ? v : 0.0;
        Rs[n * kRec + 0] = m * (0.25 / kPi);
        Rs[n * kRec + 1] = m;
        Rs[n * kRec + 2] = w;
        Rs[n * kRec + 3] = v;
        if (cb < kCoarsePos) {
          // raw moments of N(m, v): mu_1 = m, mu_2 = m^2 + v, mu_{n+1} = m mu_n + n v mu_{n-1}
          double mu_prev = 1.0, mu = m;
          for (int p = 0; p < kTayC; ++p) {
            CPf[(size_t)n * kTayC + p] = run[p];
            run[p] += w * mu;
            const double nx = m * mu + (p + 1) * v * mu_prev;
            mu_prev = mu;
            mu = nx;
          }
        }
        ++n;
      }
    }
    hist[HL.scnt + si] = (double)n;
  }
}

struct BinRec {
  double x4pi, xi, w, v;
};

// e^{i ph - k^2 v/2} - 1 = re + i im with ph = k xibar, and smpg = im - ph.
__device__ __forceinline__ void bin_phase(double k, const BinRec& r, double& re, double& im,
                                          double& smpg) {
  double ph = k * r.xi;
  double rv = k * r.x4pi;               // half phase in revolutions
  double f = rv - rint(rv);             // in [-1/2, 1/2]
  float ff = (float)f;
  float hs = __builtin_amdgcn_sinf(ff);  // sin(2 pi f) = sin(ph / 2)
  float hc = __builtin_amdgcn_cosf(ff);
  float cm1 = -2.0f * hs * hs;
  float sn = 2.0f * hs * hc;
  double smp;
  if (fabs(ph) < 0.5) {
    double p2 = ph * ph;
    smp = -ph * p2 * (1.0 / 6.0) *
          (1.0 - p2 / 20.0 * (1.0 - p2 / 42.0 * (1.0 - p2 / 72.0 * (1.0 - p2 / 110.0))));
  } else {
    smp = (double)sn - ph;
  }
  float x = (float)(-0.5 * k * k * r.v);
  float gm1 = (x > -1e-3f) ? x * (1.0f + 0.5f * x * (1.0f + x * (1.0f / 3.0f))) : __expf(x) - 1.0f;
  re = (double)(cm1 + gm1 + cm1 * gm1);
  im = (double)(sn + sn * gm1);
  smpg = smp + (double)(sn * gm1);
}

template <int kBinTile>
__global__ __launch_bounds__(kBinTile) void k_lnphi_binned(const double* __restrict__ hist,
                                                      const double* __restrict__ cosmo,
                                                      const double* __restrict__ knodes,
                                                      sgl_shape sh, double* __restrict__ lam) {
  __shared__ BinRec sb[kBinTile];
  const int t = threadIdx.x, b = blockIdx.y;
  const int n = blockIdx.x * kBinTile + t;
  const int n_k = sh.n_k;
  const HistLayout HL = hist_layout(sh, gridDim.y);
  const double k = knodes[min(n, n_k - 1)];
  const int nf = (int)hist[HL.fcnt + b];
  const double* fr = hist + HL.frec + (size_t)b * kFineBins * kRec;
  // Records of positive bins whose upper edge is <= kTayEps / (1.05 k) have
  // k (|xi| + 3 sd) <= 1/2 (a bin's content spreads over at most half its
  // relative width of 3.1 %): their sum of w (e^{i k xi - k^2 v/2} - 1 - i k xi)
  // is the series through n = 9 on the prefix raw Gaussian moments
  // (truncation < 1e-9 relative); the rest are summed directly.  The
  // workgroup's k nodes are increasing, so its last node has the lowest
  // first direct record.
  const double* fpos = hist + HL.fpos + (size_t)b * (kFineBins + 1);
  auto first_direct = [&](double kk) {
    const double z = (log(kTayEps / (1.05 * kk)) - log(kXiLo)) * (kFinePos / log(kXiHi / kXiLo));
    const int F = (z <= 0.0) ? 0 : (z >= (double)kFinePos ? kFinePos : (int)z);
    return (int)fpos[F];
  };
  const int r0 = first_direct(k);
  const int r_lo = first_direct(knodes[min((int)(blockIdx.x + 1) * kBinTile, n_k) - 1]);
  double lr, li;
  {
    const double* M = hist + HL.fpfx + ((size_t)b * (kFinePos + 1) + r0) * kTay;
    const double q = k * k;
    // cos - 1 = sum_m (-1)^m (k xi)^{2m} / (2m)!,  sin - k xi = sum_m (-1)^m (k xi)^{2m+1} / (2m+1)!
    lr = q * (-M[0] * 0.5 + q * (M[2] * (1.0 / 24) + q * (-M[4] * (1.0 / 720) + q * (M[6] * (1.0 / 40320)))));
    li = k * q * (-M[1] * (1.0 / 6) + q * (M[3] * (1.0 / 120) + q * (-M[5] * (1.0 / 5040) + q * (M[7] * (1.0 / 362880)))));
  }
  for (int c0 = r_lo; c0 < nf; c0 += kBinTile) {
    __syncthreads();
    if (c0 + t < nf) {
      const double* r = fr + (size_t)(c0 + t) * kRec;
      sb[t] = BinRec{r[0], r[1], r[2], r[3]};
    }
    __syncthreads();
    const int nb = min(kBinTile, nf - c0);
    const int c_from = max(r0 - c0, 0);
    for (int c = c_from; c < nb; ++c) {
      const BinRec r = sb[c];
      if (k * k * r.v > 80.0) {                // e^{-k^2 v / 2} < e^{-40}: w (-1 - i k xi)
        lr -= r.w;
        li -= r.w * k * r.xi;
        continue;
      }
      double re, im, smpg;
      bin_phase(k, r, re, im, smpg);
      lr = fma(r.w, re, lr);
      li = fma(r.w, smpg, li);
    }
  }
  if (n < n_k) {
    lam[2 * ((size_t)b * n_k + n)] = lr;
    lam[2 * ((size_t)b * n_k + n) + 1] = li;
  }
}

// Slice amplitudes a_i(k) = sum_c beta w_c (e^{i k xi_c - k^2 v_c/2} - 1)
// over the slice's coarse Gaussian records, at the two-point sub-grid nodes
// (node jc at k node jc S of the log grid): workgroup = (instance, kAmpTile
// sub-grid nodes), one node per thread.  FP32 output: the two-point term
// consumes the amplitudes in FP32.
constexpr int kAmpTile = 128;
__global__ __launch_bounds__(kAmpTile) void k_amp_binned(const double* __restrict__ hist,
                                                         const double* __restrict__ cosmo,
                                                         sgl_shape sh, float* __restrict__ amp) {
  static_assert(kAmpTile >= kSliceBins, "a slice's coarse records are staged in one pass");
  __shared__ BinRec sb[kAmpTile];
  const int t = threadIdx.x, b = blockIdx.y;
  const int nc = psi2_nodes(sh), S = psi2_stride(sh);
  const int jc = blockIdx.x * kAmpTile + t;
  const HistLayout HL = hist_layout(sh, gridDim.y);
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  const double k = knode_at(sh, min(jc, nc - 1) * S);
  // Coarse records of positive bins below coarse bin cb have mean + 2 sd
  // <= 2 x (upper edge) (the spread of a bin's content is at most half its
  // width), so with 2 k edge <= kTayEps their e^{i k xi - k^2 v/2} - 1 is the
  // moment series through n = 9 (truncation < 1e-10 relative).  The
  // workgroup's last node has the lowest first direct record.
  auto coarse_cut = [&](double kk) {
    const double z = (log(0.5 * kTayEps / kk) - log(kXiLo)) * (kFinePos / log(kXiHi / kXiLo));
    const int F = (z <= 0.0) ? 0 : (z >= (double)kFinePos ? kFinePos : (int)z);
    return F / kCoarse;
  };
  const int cb0 = coarse_cut(k);
  const int cb_lo = coarse_cut(knode_at(sh, (min((int)(blockIdx.x + 1) * kAmpTile, nc) - 1) * S));
  const double q = k * k;
  for (int i = 0; i < n_z; ++i) {                 // n_z is uniform over the workgroup
    const size_t si = (size_t)b * sh.n_z_max + i;
    const int ns = (int)hist[HL.scnt + si];
    const double* sr = hist + HL.srec + si * kSliceBins * kRec;
    const double* cpo = hist + HL.cpos + si * (kCoarsePos + 1);
    const int c0 = (int)cpo[cb0], c_lo = (int)cpo[cb_lo];
    double ar, ai;
    {
      const double* M = hist + HL.cpfx + (si * (kCoarsePos + 1) + c0) * kTayC;
      ar = q * (-M[1] * 0.5 + q * (M[3] * (1.0 / 24) + q * (-M[5] * (1.0 / 720) + q * (M[7] * (1.0 / 40320)))));
      ai = k * (M[0] + q * (-M[2] * (1.0 / 6) + q * (M[4] * (1.0 / 120) + q * (-M[6] * (1.0 / 5040) + q * (M[8] * (1.0 / 362880))))));
    }
    __syncthreads();
    if (c_lo + t < ns) {
      const double* r = sr + (size_t)(c_lo + t) * kRec;
      sb[t] = BinRec{r[0], r[1], r[2], r[3]};
    }
    __syncthreads();
    for (int c = c0 - c_lo; c < ns - c_lo; ++c) {
      const BinRec r = sb[c];
      if (q * r.v > 80.0) {              // e^{-k^2 v / 2} < e^{-40}: the record adds -w
        ar -= r.w;
        continue;
      }
      double re, im, smpg;
      bin_phase(k, r, re, im, smpg);
      ar = fma(r.w, re, ar);
      ai = fma(r.w, im, ai);
    }
    if (jc < nc) reinterpret_cast<float2*>(amp)[si * nc + jc] = make_float2((float)ar, (float)ai);
  }
}

int launch_lnphi_exact(const double* rows, const double* frows, const double* cosmo,
                       const double* uedge, const double* knodes, const sgl_shape& sh, int B,
                       double* lam, double* amp, double* work, hipStream_t st) {
  dim3 grid((sh.n_k + 256 * kKpt - 1) / (256 * kKpt), sh.n_z_max, B);
  if (sh.flags & SGL_FLAG_FILAMENTS)
    hipLaunchKernelGGL(k_lnphi_exact_fil, grid, dim3(256), 0, st, rows, frows, cosmo, uedge,
                       knodes, sh, work, amp);
  else
    hipLaunchKernelGGL(k_lnphi_exact, grid, dim3(256), 0, st, rows, frows, cosmo, uedge, knodes,
                       sh, work, amp);
  size_t n = (size_t)B * sh.n_k;
  hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, work, B,
                     sh.n_z_max, sh.n_k, lam);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_sum_slices(const double* work, const sgl_shape& sh, int B, double* lam, hipStream_t st) {
  size_t n = (size_t)B * sh.n_k;
  hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, work, B,
                     sh.n_z_max, sh.n_k, lam);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

// pass 0: group sums wk; pass 1: exponent partials (work) and the closure's
// zpart / ytot (needs the rate table of sgl_vp setup).
int launch_vp_cells(const double* rows, const double* frows, const double* cosmo,
                    const double* uedge, const double* knodes, const sgl_shape& sh, int B,
                    const VpLayout& V, double* vp, double* work, int pass, hipStream_t st) {
  if (pass == 0) {
    hipLaunchKernelGGL(k_vp_wk, dim3(sh.n_z_max, B, kVpWkSplit), dim3(256), 0, st, rows, frows,
                       cosmo, uedge, sh, vp + V.wk);
  } else {
    dim3 grid((sh.n_k + 256 * kKpt - 1) / (256 * kKpt), sh.n_z_max, B);
    hipLaunchKernelGGL(k_lnphi_exact_vp, grid, dim3(256), 0, st, rows, frows, cosmo, uedge,
                       knodes, sh, V, vp, work);
  }
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

// Workgroups of `threads` the device holds at once for `kernel` (occupancy x
// compute units), looked up once per kernel, size and device.
static int resident_workgroups(const void* kernel, int threads) {
  struct Entry { const void* k; int dev, threads, n; };
  static Entry cache[32];
  static int used = 0;
  static std::mutex mu;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  for (int e = 0; e < used; ++e)
    if (cache[e].k == kernel && cache[e].dev == dev && cache[e].threads == threads)
      return cache[e].n;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 1;
  const int n = per_cu * cus;
  if (used < 32) cache[used++] = Entry{kernel, dev, threads, n};
  return n;
}

// The binning kernel of one launch shape, on either schedule.
template <int kThreads, bool kEll>
static void launch_bin_cells(const double* cosmo, const double* uedge, const double* rows,
                             const sgl_shape& sh, int B, double* hist, hipStream_t st) {
  if (g_sgl_tune[SGL_TUNE_BIN_SCHEDULE] == 1) {
    hipLaunchKernelGGL(k_bin_items, dim3(1), dim3(kItemThreads), 0, st, cosmo, sh, B, hist);
    const int n = resident_workgroups(reinterpret_cast<const void*>(&k_bin_cells_queue<kThreads, kEll>),
                                      kThreads);
    hipLaunchKernelGGL((k_bin_cells_queue<kThreads, kEll>),
                       dim3((unsigned)std::max(1, std::min(n, B * sh.n_z_max))), dim3(kThreads), 0,
                       st, cosmo, uedge, rows, sh, B, hist);
  } else {
    hipLaunchKernelGGL((k_bin_cells<kThreads, kEll>), dim3(sh.n_z_max, B), dim3(kThreads), 0, st,
                       cosmo, uedge, rows, sh, hist);
  }
}

int launch_bin_pass(const double* rows, const double* frows, const double* cosmo,
                    const double* uedge, const sgl_shape& sh, int B, double* hist,
                    hipStream_t st) {
  const size_t nrow = (size_t)B * sh.n_z_max * sh.n_m;
  hipLaunchKernelGGL(k_rowf, dim3((unsigned)((nrow + 255) / 256)), dim3(256), 0, st, rows, cosmo,
                     uedge, sh, B, hist);
  hipLaunchKernelGGL(k_cols, dim3((sh.n_b + 255) / 256), dim3(256), 0, st, uedge, sh, B, hist);
  const dim3 g(sh.n_z_max, B);
  const bool ell = (sh.flags & SGL_FLAG_ELL) != 0;
  switch (g_sgl_tune[SGL_TUNE_BIN_THREADS]) {
    case 0:
      if (ell) launch_bin_cells<256, true>(cosmo, uedge, rows, sh, B, hist, st);
      else launch_bin_cells<256, false>(cosmo, uedge, rows, sh, B, hist, st);
      break;
    case 2:
      if (ell) launch_bin_cells<1024, true>(cosmo, uedge, rows, sh, B, hist, st);
      else launch_bin_cells<1024, false>(cosmo, uedge, rows, sh, B, hist, st);
      break;
    default:
      if (ell) launch_bin_cells<512, true>(cosmo, uedge, rows, sh, B, hist, st);
      else launch_bin_cells<512, false>(cosmo, uedge, rows, sh, B, hist, st);
      break;
  }
  if (sh.flags & SGL_FLAG_FILAMENTS)
    hipLaunchKernelGGL(k_bin_fil, g, dim3(kFilThreads), 0, st, frows, cosmo, sh, hist);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_bin_reduce(const double* rows, const double* cosmo, const double* uedge,
                      const sgl_shape& sh, int B, double* hist, hipStream_t st) {
  hipLaunchKernelGGL(k_reduce_fine, dim3(B), dim3(256), 0, st, cosmo, sh, hist);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_bin(const double* rows, const double* frows, const double* cosmo, const double* uedge,
               const sgl_shape& sh, int B, double* hist, hipStream_t st) {
  int rc = launch_bin_pass(rows, frows, cosmo, uedge, sh, B, hist, st);
  return rc == SGL_OK ? launch_bin_reduce(rows, cosmo, uedge, sh, B, hist, st) : rc;
}

int launch_lnphi_binned(const double* hist, const double* cosmo, const double* knodes,
                        const sgl_shape& sh, int B, double* lam, float* amp, hipStream_t st) {
  if (g_sgl_tune[SGL_TUNE_LNPHI_BLOCK] == 0) {
    dim3 grid((sh.n_k + 127) / 128, B);
    hipLaunchKernelGGL(k_lnphi_binned<128>, grid, dim3(128), 0, st, hist, cosmo, knodes, sh, lam);
  } else {
    dim3 grid((sh.n_k + 255) / 256, B);
    hipLaunchKernelGGL(k_lnphi_binned<256>, grid, dim3(256), 0, st, hist, cosmo, knodes, sh, lam);
  }
  if (sh.flags & SGL_FLAG_CLUSTERING) {
    dim3 grid((psi2_nodes(sh) + kAmpTile - 1) / kAmpTile, B);
    hipLaunchKernelGGL(k_amp_binned, grid, dim3(kAmpTile), 0, st, hist, cosmo, sh, amp);
  }
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

// Truncated-NFW (kappa, gamma) at n given x for one row (k_s, x_t): FP64
// closed forms (the exact path's trunc_kg_row) or the batched path's FP32
// form (trunc_kg_row_f32 on the row's RowF record).
__global__ void k_trunc_nfw(const double* __restrict__ x, int n, double ks, double xt, int prec,
                            double* __restrict__ kap, double* __restrict__ gam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double R[kRowStride] = {ks, 1.0, xt, 1.0, 1.0, 0.0, 0.0, 0.0};
  const RowConst rc = row_const(R);
  double k, g;
  if (prec == SGL_PREC_FP64) {
    trunc_kg_row(x[i], rc, k, g);
  } else {
    const RowF f = make_rowf(rc);
    float kf, gf;
    trunc_kg_row_f32((float)x[i], (float)log(x[i]), (float)(1.0 / x[i]), f, kf, gf);
    k = kf;
    g = gf;
  }
  kap[i] = k;
  gam[i] = g;
}

int launch_trunc_nfw(const double* x, int n, double ks, double xt, int prec, double* kap,
                     double* gam, hipStream_t st) {
  hipLaunchKernelGGL(k_trunc_nfw, dim3((n + 255) / 256), dim3(256), 0, st, x, n, ks, xt, prec, kap,
                     gam);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_fill_grids(const sgl_shape& sh, double* ue, double* kn, double* m_nodes_h,
                      hipStream_t st) {
  int n = sh.n_b + 1;
  if (sh.n_k > n) n = sh.n_k;
  if (sh.n_m > n) n = sh.n_m;
  hipLaunchKernelGGL(k_fill_uedges, dim3((n + 255) / 256), dim3(256), 0, st, sh, ue, kn,
                     m_nodes_h);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

}  // namespace sgl
