// Compound-Poisson exponent of the lens population.
//
//   Lambda(k) = sum_c w_c (e^{i k xi_c} - 1 - i k xi_c)         (compensated)
//   a_i(k)    = sum_{c in slice i} beta_c w_c (e^{i k xi_c} - 1)
//
// Cells are never stored: every workgroup regenerates the cells of its slice
// from the (slice, mass) rows (truncated NFW, jump xi = -ln detA, tilted
// weight w detA) into LDS and streams them against its tile of k nodes.
//
// Exact path (k_lnphi_exact): every cell, FP64 sincos, the reference's
// cancellation-free forms cos - 1 = -2 sin^2(ph/2), sin - ph by series below
// |ph| = 1e-2 (sgl_full.py:2599-2613).
// Binned path (k_bin_cells + k_lnphi_binned): cells histogrammed on a
// signed-log jump lattice with (w, w xi, w xi^2) per bin; each bin acts as
// w e^{i k xibar - k^2 v / 2}, v the in-bin variance.  Phase in FP64,
// reduced to revolutions, sin/cos on the FP32 hardware path, compensated
// terms and accumulation in FP64.
#include "sgl_amd_internal.h"

namespace sgl {

constexpr int kCellTile = 256;

struct CellGen {
  const double* rows;   // [n_m][kRowStride] of this (instance, slice)
  const double* uedge;  // [n_b + 1]
  int n_b;
  __device__ inline void gen(int c, double& xi, double& ws, double& bws) const {
    const int j = c / n_b, l = c - j * n_b;
    const double* R = rows + (size_t)j * kRowStride;
    const double ks = R[kRowKs], rs = R[kRowRs], xt = R[kRowXt], rvir = R[kRowRvir];
    const double wrow = R[kRowW], beta = R[kRowBeta];
    if (!(wrow > 0.0)) { xi = 0.0; ws = 0.0; bws = 0.0; return; }
    const double b0 = rvir * uedge[l], b1 = rvir * uedge[l + 1];
    const double x = sqrt(b0 * b1) / rs;
    const double area = kPi * (b1 * b1 - b0 * b0);
    double kap, gam;
    trunc_kg(x, ks, xt, kap, gam);
    cell_jump(kap, gam, wrow * area, xi, ws);
    bws = beta * ws;
  }
};

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
  if (m_nodes_h != nullptr && t < sh.n_m) {
    double e0 = log(sh.m_lo_h), e1 = log(sh.m_hi_h);
    double a = e0 + (e1 - e0) * t / sh.n_m, b = e0 + (e1 - e0) * (t + 1) / sh.n_m;
    m_nodes_h[t] = exp(0.5 * (a + b));
  }
}

// ------------------------------------------------------------ exact path --
constexpr int kKpt = 4;  // k nodes per thread

__global__ __launch_bounds__(256) void k_lnphi_exact(const double* __restrict__ rows,
                                                     const double* __restrict__ cosmo,
                                                     const double* __restrict__ uedge,
                                                     const double* __restrict__ knodes,
                                                     sgl_shape sh, double* __restrict__ part,
                                                     double* __restrict__ amp) {
  __shared__ double sxi[kCellTile], sws[kCellTile], sbw[kCellTile];
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
    CellGen g{rows + ((size_t)b * sh.n_z_max + i) * sh.n_m * kRowStride, uedge, sh.n_b};
    const int n_cells = sh.n_m * sh.n_b;
    for (int c0 = 0; c0 < n_cells; c0 += kCellTile) {
      __syncthreads();
      {
        int c = c0 + t;
        double xi = 0.0, ws = 0.0, bw = 0.0;
        if (c < n_cells) g.gen(c, xi, ws, bw);
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
__device__ __forceinline__ int fine_bin(double xi) {
  if (xi > 0.0) {
    if (xi < kXiLo) return 0;
    int k = (int)(log(xi / kXiLo) * (kFinePos / log(kXiHi / kXiLo)));
    return min(max(k, 0), kFinePos - 1);
  }
  double a = -xi;
  int k = (a < kXiNegLo) ? 0 : (int)(log(a / kXiNegLo) * (kFineNeg / log(kXiNegHi / kXiNegLo)));
  return kFinePos + min(max(k, 0), kFineNeg - 1);
}

__device__ __forceinline__ int slice_bin(double xi) {
  if (xi > 0.0) {
    if (xi < kXiLo) return 0;
    int k = (int)(log(xi / kXiLo) * (kSlicePos / log(kXiHi / kXiLo)));
    return min(max(k, 0), kSlicePos - 1);
  }
  double a = -xi;
  int k = (a < kXiNegLo) ? 0 : (int)(log(a / kXiNegLo) * (kSliceNeg / log(kXiNegHi / kXiNegLo)));
  return kSlicePos + min(max(k, 0), kSliceNeg - 1);
}

__global__ void k_zero(double* __restrict__ p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t s = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += s) p[i] = 0.0;
}

__global__ __launch_bounds__(256) void k_bin_cells(const double* __restrict__ rows,
                                                   const double* __restrict__ cosmo,
                                                   const double* __restrict__ uedge, sgl_shape sh,
                                                   double* __restrict__ hfine,
                                                   double* __restrict__ hslice) {
  __shared__ double lf[kFineBins * kMoments];
  __shared__ double ls[kSliceBins * kMoments];
  const int t = threadIdx.x, i = blockIdx.x, b = blockIdx.y;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  for (int q = t; q < kFineBins * kMoments; q += 256) lf[q] = 0.0;
  for (int q = t; q < kSliceBins * kMoments; q += 256) ls[q] = 0.0;
  __syncthreads();
  if (i < n_z) {
    CellGen g{rows + ((size_t)b * sh.n_z_max + i) * sh.n_m * kRowStride, uedge, sh.n_b};
    const int n_cells = sh.n_m * sh.n_b;
    for (int c = t; c < n_cells; c += 256) {
      double xi, ws, bw;
      g.gen(c, xi, ws, bw);
      if (ws == 0.0) continue;
      int f = fine_bin(xi) * kMoments;
      atomicAdd(&lf[f], ws);
      atomicAdd(&lf[f + 1], ws * xi);
      atomicAdd(&lf[f + 2], ws * xi * xi);
      int s = slice_bin(xi) * kMoments;
      atomicAdd(&ls[s], bw);
      atomicAdd(&ls[s + 1], bw * xi);
      atomicAdd(&ls[s + 2], bw * xi * xi);
    }
  }
  __syncthreads();
  double* gf = hfine + (size_t)b * kFineBins * kMoments;
  for (int q = t; q < kFineBins * kMoments; q += 256)
    if (lf[q] != 0.0) unsafeAtomicAdd(&gf[q], lf[q]);
  double* gs = hslice + ((size_t)b * sh.n_z_max + i) * kSliceBins * kMoments;
  for (int q = t; q < kSliceBins * kMoments; q += 256) gs[q] = ls[q];
}

// One bin as a pair (xibar / (4 pi), xibar, w, v).
struct BinRec {
  double x4pi, xi, w, v;
};

__device__ __forceinline__ BinRec load_bin(const double* h) {
  BinRec r;
  double w = h[0];
  if (w == 0.0) { r.x4pi = r.xi = r.w = r.v = 0.0; return r; }
  double m = h[1] / w;
  double v = h[2] / w - m * m;
  r.xi = m;
  r.x4pi = m * (0.25 / kPi);
  r.w = w;
  r.v = (v > 0.0) ? v : 0.0;
  return r;
}

// e^{i ph - k^2 v/2} - 1 split into (cm1g, sng) with ph = k xibar:
//   re = cos(ph) g - 1, im = sin(ph) g, and smpg = sin(ph) g - ph.
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

constexpr int kBinTile = 256;

__global__ __launch_bounds__(256) void k_lnphi_binned(const double* __restrict__ hfine,
                                                      const double* __restrict__ hslice,
                                                      const double* __restrict__ cosmo,
                                                      const double* __restrict__ knodes,
                                                      sgl_shape sh, double* __restrict__ lam,
                                                      double* __restrict__ amp) {
  __shared__ BinRec sb[kBinTile];
  const int t = threadIdx.x, b = blockIdx.y;
  const int n = blockIdx.x * 256 + t;
  const int n_k = sh.n_k;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  const double k = (n < n_k) ? knodes[n] : 0.0;
  double lr = 0.0, li = 0.0;
  const double* hf = hfine + (size_t)b * kFineBins * kMoments;
  for (int c0 = 0; c0 < kFineBins; c0 += kBinTile) {
    __syncthreads();
    sb[t] = (c0 + t < kFineBins) ? load_bin(hf + (size_t)(c0 + t) * kMoments) : BinRec{0, 0, 0, 0};
    __syncthreads();
    for (int c = 0; c < kBinTile; ++c) {
      const BinRec r = sb[c];
      if (r.w == 0.0) continue;
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
  if (!(sh.flags & SGL_FLAG_CLUSTERING)) return;
  for (int i = 0; i < sh.n_z_max; ++i) {
    double ar = 0.0, ai = 0.0;
    if (i < n_z) {
      const double* hs = hslice + ((size_t)b * sh.n_z_max + i) * kSliceBins * kMoments;
      for (int c0 = 0; c0 < kSliceBins; c0 += kBinTile) {
        __syncthreads();
        sb[t] = (c0 + t < kSliceBins) ? load_bin(hs + (size_t)(c0 + t) * kMoments)
                                      : BinRec{0, 0, 0, 0};
        __syncthreads();
        const int nb = min(kBinTile, kSliceBins - c0);
        for (int c = 0; c < nb; ++c) {
          const BinRec r = sb[c];
          if (r.w == 0.0) continue;
          double re, im, smpg;
          bin_phase(k, r, re, im, smpg);
          ar = fma(r.w, re, ar);
          ai = fma(r.w, im, ai);
        }
      }
    }
    if (n < n_k) {
      size_t o = 2 * (((size_t)b * sh.n_z_max + i) * n_k + n);
      amp[o] = ar;
      amp[o + 1] = ai;
    }
  }
}

int launch_lnphi_exact(const double* rows, const double* cosmo, const double* uedge,
                       const double* knodes, const sgl_shape& sh, int B, double* lam, double* amp,
                       double* work, hipStream_t st) {
  dim3 grid((sh.n_k + 256 * kKpt - 1) / (256 * kKpt), sh.n_z_max, B);
  hipLaunchKernelGGL(k_lnphi_exact, grid, dim3(256), 0, st, rows, cosmo, uedge, knodes, sh, work,
                     amp);
  size_t n = (size_t)B * sh.n_k;
  hipLaunchKernelGGL(k_sum_slices, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, work, B,
                     sh.n_z_max, sh.n_k, lam);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_bin(const double* rows, const double* cosmo, const double* uedge, const sgl_shape& sh,
               int B, double* hfine, double* hslice, hipStream_t st) {
  size_t nf = hist_layout_fine(B);
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, st, hfine, nf);
  hipLaunchKernelGGL(k_bin_cells, dim3(sh.n_z_max, B), dim3(256), 0, st, rows, cosmo, uedge, sh,
                     hfine, hslice);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_lnphi_binned(const double* hfine, const double* hslice, const double* cosmo,
                        const double* knodes, const sgl_shape& sh, int B, double* lam,
                        double* amp, hipStream_t st) {
  dim3 grid((sh.n_k + 255) / 256, B);
  hipLaunchKernelGGL(k_lnphi_binned, grid, dim3(256), 0, st, hfine, hslice, cosmo, knodes, sh, lam,
                     amp);
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
