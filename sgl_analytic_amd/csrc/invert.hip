// Two-point term, phi packing, Filon inversion (FP64 on the fly, or a dense
// [n_out x 2 n_q] . [2 n_q x B] contraction on MFMA) and FP64 diagnostics.
//
// Inversion: p(t) = (1/pi) Re int_0^inf phi(k) e^{-ikt} dk (sgl.py:994-1012)
// with phi linear between consecutive k nodes (k_0 = 0, phi_0 = 1): every
// k-interval [a, a+h] contributes h e^{-iat} [E1(ht) phi_{n+1} + (E0 - E1)(ht) phi_n],
// E0(th) = int_0^1 e^{-i th s} ds, E1(th) = int_0^1 s e^{-i th s} ds, so the
// oscillation e^{-ikt} is integrated exactly and only phi is interpolated.
#include "sgl_amd_internal.h"

namespace sgl {

// ------------------------------------------------------------- quadform --
// Psi_2[b][n] = 1/2 sum_ij C_ij a_i a_j (complex square, Lam.Q nord = 1,
// sgl_full.py:3353-3397), FP64 for the plain path.  Workgroup = (instance,
// 32 k nodes): the tile's amplitudes and C in LDS, 8 threads per k node each
// taking every 8th column j of u_j = sum_i C_ij a_i, partial sums combined
// with lane shuffles.
constexpr int kQfTile = 32;

__global__ __launch_bounds__(256) void k_quadform(const double* __restrict__ cov,
                                                  const double* __restrict__ cosmo,
                                                  const double* __restrict__ amp, int n_z_max,
                                                  int n_k, double* __restrict__ psi2) {
  using T = double;
  __shared__ T sC[kMaxSlices * kMaxSlices];
  __shared__ T sa[kMaxSlices][kQfTile][2];
  const int b = blockIdx.y, t = threadIdx.x;
  const int n0 = blockIdx.x * kQfTile;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  const double* C = cov + (size_t)b * n_z_max * n_z_max;
  for (int q = t; q < n_z * n_z; q += 256) sC[q] = (T)C[(q / n_z) * n_z_max + (q % n_z)];
  for (int q = t; q < n_z * kQfTile; q += 256) {
    int i = q / kQfTile, kk = q % kQfTile, n = n0 + kk;
    size_t o = 2 * (((size_t)b * n_z_max + i) * n_k + n);
    sa[i][kk][0] = (n < n_k) ? (T)amp[o] : (T)0;
    sa[i][kk][1] = (n < n_k) ? (T)amp[o + 1] : (T)0;
  }
  __syncthreads();
  const int kk = t >> 3, part = t & 7;
  T sr = 0, si = 0;
  for (int j = part; j < n_z; j += 8) {
    T ux = 0, uy = 0;
    for (int i = 0; i < n_z; ++i) {
      T c = sC[i * n_z + j];
      ux += c * sa[i][kk][0];
      uy += c * sa[i][kk][1];
    }
    T ax = sa[j][kk][0], ay = sa[j][kk][1];
    sr += ux * ax - uy * ay;
    si += ux * ay + uy * ax;
  }
  double dr = (double)sr, di = (double)si;
  for (int m = 4; m > 0; m >>= 1) {
    dr += __shfl_xor(dr, m, 64);
    di += __shfl_xor(di, m, 64);
  }
  const int n = n0 + kk;
  if (part == 0 && n < n_k) {
    psi2[2 * ((size_t)b * n_k + n)] = 0.5 * dr;
    psi2[2 * ((size_t)b * n_k + n) + 1] = 0.5 * di;
  }
}

// FP32 two-point term of the batched path on MFMA (v_mfma_f32_16x16x4_f32):
// per instance U = A C with A[n][i] the FP32 slice amplitudes at k node n (Re and
// Im as two real matrices, K = slices) and C the slice covariance, then
// Psi_2[n] = 1/2 sum_j (Ur + i Ui)[n][j] (ar + i ai)[n][j].  A wave takes 16
// k nodes at a time with the A fragments of every slice in registers and C
// from LDS; 4 waves x kQmTiles x 16 nodes per workgroup.
constexpr int kQmPad = 96;                 // slices padded to a multiple of 16
constexpr int kQmTiles = 4;                // 16-node groups per wave
static_assert(kMaxSlices <= kQmPad, "slice padding");

typedef float f32x4q __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_quadform_mfma(const double* __restrict__ cov,
                                                       const double* __restrict__ cosmo,
                                                       const float* __restrict__ amp,
                                                       int n_z_max, int n_k,
                                                       double* __restrict__ psi2) {
  __shared__ float sC[kQmPad][kQmPad + 1];
  const int b = blockIdx.y, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  const double* C = cov + (size_t)b * n_z_max * n_z_max;
  for (int q = t; q < kQmPad * kQmPad; q += 256) {
    const int i = q / kQmPad, j = q % kQmPad;
    sC[i][j] = (i < n_z && j < n_z) ? (float)C[(size_t)i * n_z_max + j] : 0.0f;
  }
  __syncthreads();
  const int ns = (n_z + 3) / 4, nt = (n_z + 15) / 16;
  const int col = lane & 15, grp = lane >> 4;
  const float2* A2 = (const float2*)amp + (size_t)b * n_z_max * n_k;
  for (int g = 0; g < kQmTiles; ++g) {
    const int n0 = (blockIdx.x * 4 * kQmTiles + wave * kQmTiles + g) * 16;
    if (n0 >= n_k) break;
    // A fragments: lane supplies A[n0 + col][4 s + grp]
    float ar[kQmPad / 4], ai[kQmPad / 4];
#pragma unroll
    for (int s = 0; s < kQmPad / 4; ++s) {
      const int i = 4 * s + grp, n = n0 + col;
      float2 v = make_float2(0.0f, 0.0f);
      if (s < ns && i < n_z && n < n_k) v = A2[(size_t)i * n_k + n];
      ar[s] = v.x;
      ai[s] = v.y;
    }
    float sr[4] = {0.f, 0.f, 0.f, 0.f}, si[4] = {0.f, 0.f, 0.f, 0.f};
    for (int jt = 0; jt < nt; ++jt) {
      f32x4q ur = {0.f, 0.f, 0.f, 0.f}, ui = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < kQmPad / 4; ++s) {
        if (s < ns) {
          const float c = sC[4 * s + grp][16 * jt + col];
          ur = __builtin_amdgcn_mfma_f32_16x16x4f32(ar[s], c, ur, 0, 0, 0);
          ui = __builtin_amdgcn_mfma_f32_16x16x4f32(ai[s], c, ui, 0, 0, 0);
        }
      }
      // lane holds U[n0 + 4 grp + r][16 jt + col]
      const int j = 16 * jt + col;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + 4 * grp + r;
        if (j < n_z && n < n_k) {
          const float2 v = A2[(size_t)j * n_k + n];
          const float xr = v.x, xi = v.y;
          sr[r] += ur[r] * xr - ui[r] * xi;
          si[r] += ur[r] * xi + ui[r] * xr;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double dr = sr[r], di = si[r];
      for (int m = 8; m > 0; m >>= 1) {
        dr += __shfl_xor(dr, m, 64);
        di += __shfl_xor(di, m, 64);
      }
      const int n = n0 + 4 * grp + r;
      if (col == 0 && n < n_k) {
        psi2[2 * ((size_t)b * n_k + n)] = 0.5 * dr;
        psi2[2 * ((size_t)b * n_k + n) + 1] = 0.5 * di;
      }
    }
  }
}

// Split-BF16x3 variant of the same product on v_mfma_f32_16x16x32_bf16: A and
// C are each split into a BF16 head and a BF16 tail (x = hi + lo, lo the
// rounded remainder), U = Ahi Chi + Ahi Clo + Alo Chi with FP32 accumulation
// (the dropped Alo Clo term is 2^-16 of the product).  C is symmetric, so the
// B operand C[k][j] is read as row j of C, 8 consecutive k per lane in one
// 16-byte LDS load.  The epilogue (U against the FP32 amplitudes) is the FP32
// kernel's.
constexpr int kQbK = 32;                    // K per MFMA step
constexpr int kQbSteps = kQmPad / kQbK;     // 3 steps cover 96 slices
constexpr int kQbLd = kQmPad + 8;           // bf16 row stride of the C tiles (shorts)
typedef short bf16x8q __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short bf16_hi(float x) {
  const uint32_t u = __float_as_uint(x);
  return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_val(unsigned short h) {
  return __uint_as_float(((uint32_t)h) << 16);
}

__global__ __launch_bounds__(256) void k_quadform_bf16x3(const double* __restrict__ cov,
                                                         const double* __restrict__ cosmo,
                                                         const float* __restrict__ amp,
                                                         int n_z_max, int n_k,
                                                         double* __restrict__ psi2) {
  __shared__ __align__(16) unsigned short sCh[kQmPad][kQbLd];
  __shared__ __align__(16) unsigned short sCl[kQmPad][kQbLd];
  const int b = blockIdx.y, t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int n_z = (int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10];
  const double* C = cov + (size_t)b * n_z_max * n_z_max;
  for (int q = t; q < kQmPad * kQmPad; q += 256) {
    const int i = q / kQmPad, j = q % kQmPad;
    const float c = (i < n_z && j < n_z) ? (float)C[(size_t)i * n_z_max + j] : 0.0f;
    const unsigned short h = bf16_hi(c);
    sCh[i][j] = h;
    sCl[i][j] = bf16_hi(c - bf16_val(h));
  }
  __syncthreads();
  const int ns = (n_z + kQbK - 1) / kQbK, nt = (n_z + 15) / 16;
  const int col = lane & 15, grp = lane >> 4;
  const float2* A2 = (const float2*)amp + (size_t)b * n_z_max * n_k;
  for (int g = 0; g < kQmTiles; ++g) {
    const int n0 = (blockIdx.x * 4 * kQmTiles + wave * kQmTiles + g) * 16;
    if (n0 >= n_k) break;
    // A fragments: lane supplies A[n0 + col][32 s + 8 grp .. + 7], split
    bf16x8q arh[kQbSteps], arl[kQbSteps], aih[kQbSteps], ail[kQbSteps];
#pragma unroll
    for (int s = 0; s < kQbSteps; ++s) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int i = kQbK * s + 8 * grp + e, n = n0 + col;
        float2 v = make_float2(0.0f, 0.0f);
        if (s < ns && i < n_z && n < n_k) v = A2[(size_t)i * n_k + n];
        const unsigned short hr = bf16_hi(v.x), hi = bf16_hi(v.y);
        arh[s][e] = (short)hr;
        arl[s][e] = (short)bf16_hi(v.x - bf16_val(hr));
        aih[s][e] = (short)hi;
        ail[s][e] = (short)bf16_hi(v.y - bf16_val(hi));
      }
    }
    float sr[4] = {0.f, 0.f, 0.f, 0.f}, si[4] = {0.f, 0.f, 0.f, 0.f};
    for (int jt = 0; jt < nt; ++jt) {
      f32x4q ur = {0.f, 0.f, 0.f, 0.f}, ui = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < kQbSteps; ++s) {
        if (s < ns) {
          // B operand C[32 s + 8 grp .. + 7][16 jt + col] = row (16 jt + col) of C
          const bf16x8q ch = *(const bf16x8q*)&sCh[16 * jt + col][kQbK * s + 8 * grp];
          const bf16x8q cl = *(const bf16x8q*)&sCl[16 * jt + col][kQbK * s + 8 * grp];
          ur = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arl[s], ch, ur, 0, 0, 0);
          ur = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh[s], cl, ur, 0, 0, 0);
          ur = __builtin_amdgcn_mfma_f32_16x16x32_bf16(arh[s], ch, ur, 0, 0, 0);
          ui = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ail[s], ch, ui, 0, 0, 0);
          ui = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih[s], cl, ui, 0, 0, 0);
          ui = __builtin_amdgcn_mfma_f32_16x16x32_bf16(aih[s], ch, ui, 0, 0, 0);
        }
      }
      // lane holds U[n0 + 4 grp + r][16 jt + col]
      const int j = 16 * jt + col;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + 4 * grp + r;
        if (j < n_z && n < n_k) {
          const float2 v = A2[(size_t)j * n_k + n];
          sr[r] += ur[r] * v.x - ui[r] * v.y;
          si[r] += ur[r] * v.y + ui[r] * v.x;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double dr = sr[r], di = si[r];
      for (int m = 8; m > 0; m >>= 1) {
        dr += __shfl_xor(dr, m, 64);
        di += __shfl_xor(di, m, 64);
      }
      const int n = n0 + 4 * grp + r;
      if (col == 0 && n < n_k) {
        psi2[2 * ((size_t)b * n_k + n)] = 0.5 * dr;
        psi2[2 * ((size_t)b * n_k + n) + 1] = 0.5 * di;
      }
    }
  }
}

// Psi_2 from the sub-grid (node jc at k node jc S) onto every k node: cubic
// Lagrange in the node index (ln k is uniform in it) on the four sub-grid
// nodes around n, one-sided in the first and last interval.  At n = jc S the
// weights are exactly (0, 1, 0, 0) or its shifts, so sub-grid values pass
// through unchanged.
__global__ void k_psi2_interp(const double* __restrict__ psi2c, sgl_shape sh, int B,
                              double* __restrict__ psi2) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)B * sh.n_k) return;
  const int n = (int)(gid % sh.n_k);
  const int b = (int)(gid / sh.n_k);
  const int S = psi2_stride(sh), nc = psi2_nodes(sh);
  if (S == 1) {                                  // every node on the sub-grid
    psi2[2 * gid] = psi2c[2 * gid];
    psi2[2 * gid + 1] = psi2c[2 * gid + 1];
    return;
  }
  int j = n / S - 1;
  j = max(0, min(j, nc - 4));
  const double s = (double)(n - j * S) / S;   // in [0, 3]
  const double l0 = -(s - 1) * (s - 2) * (s - 3) / 6.0, l1 = s * (s - 2) * (s - 3) / 2.0;
  const double l2 = -s * (s - 1) * (s - 3) / 2.0, l3 = s * (s - 1) * (s - 2) / 6.0;
  const double* y = psi2c + 2 * ((size_t)b * nc + j);
  psi2[2 * gid] = l0 * y[0] + l1 * y[2] + l2 * y[4] + l3 * y[6];
  psi2[2 * gid + 1] = l0 * y[1] + l1 * y[3] + l2 * y[5] + l3 * y[7];
}

// --------------------------------------------------------------- window --
__device__ __forceinline__ double bspline_window(double k, double K) {
  // cubic B-spline of support [0, K], 1 at k = 0 (sgl_full.py:83-94)
  double x = fabs(k) * 2.0 / K;
  double w;
  if (x < 1.0) w = 2.0 / 3.0 - x * x + 0.5 * x * x * x;
  else if (x < 2.0) w = (2.0 - x) * (2.0 - x) * (2.0 - x) / 6.0;
  else w = 0.0;
  return w * 1.5;
}

// Operand layout: K index kk = 2n (Re phi_n), 2n + 1 (Im phi_n), n = 0..n_q-1,
// padded with zeros to Kp = round_up(2 n_q, 64).
inline int k_padded(const sgl_shape& sh) { return ((2 * (sh.n_k + 1) + 63) / 64) * 64; }

__device__ __forceinline__ uint16_t f32_to_bf16_rn(float x) {
  uint32_t u = __float_as_uint(x);
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__global__ void k_phi_pack(const double* __restrict__ lam, const double* __restrict__ psi2,
                           const double* __restrict__ knodes, sgl_shape sh, int B, int prec,
                           int Kp, void* __restrict__ out) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int nq2 = Kp / 2;  // padded node count
  if (gid >= (size_t)B * nq2) return;
  const int n = (int)(gid % nq2);
  const int b = (int)(gid / nq2);
  double re = 0.0, im = 0.0;
  if (n == 0) {
    re = 1.0;
  } else if (n <= sh.n_k) {
    size_t o = 2 * ((size_t)b * sh.n_k + (n - 1));
    double lr = lam[o], li = lam[o + 1];
    if (psi2 != nullptr) { lr += psi2[o]; li += psi2[o + 1]; }
    double k = knodes[n - 1];
    double w = (sh.flags & SGL_FLAG_WINDOW) ? bspline_window(k, knodes[sh.n_k - 1]) : 1.0;
    double a = w * exp(lr);
    double s, c;
    sincos(li, &s, &c);
    re = a * c;
    im = a * s;
  }
  if (prec == SGL_PREC_FP64) {
    if (n <= sh.n_k) {
      double* o = (double*)out + 2 * ((size_t)b * (sh.n_k + 1) + n);
      o[0] = re;
      o[1] = im;
    }
  } else if (prec == SGL_PREC_FP32) {
    float* o = (float*)out + (size_t)b * Kp + 2 * n;
    o[0] = (float)re;
    o[1] = (float)im;
  } else {
    uint16_t* hi = (uint16_t*)out + (size_t)b * Kp + 2 * n;
    uint16_t* lo = (uint16_t*)out + (size_t)B * Kp + (size_t)b * Kp + 2 * n;
    float fr = (float)re, fi = (float)im;
    uint16_t hr = f32_to_bf16_rn(fr), hi_ = f32_to_bf16_rn(fi);
    hi[0] = hr;
    hi[1] = hi_;
    lo[0] = f32_to_bf16_rn(fr - bf16_to_f32(hr));
    lo[1] = f32_to_bf16_rn(fi - bf16_to_f32(hi_));
  }
}

// ---------------------------------------------------------------- Filon --
__device__ inline void filon_e(double th, double& e0r, double& e0i, double& e1r, double& e1i) {
  if (fabs(th) < 0.5) {
    double tr = 1.0, ti = 0.0, a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
    for (int m = 0; m < 24; ++m) {
      a0r += tr / (m + 1);
      a0i += ti / (m + 1);
      a1r += tr / (m + 2);
      a1i += ti / (m + 2);
      double nr = ti * th / (m + 1), ni = -tr * th / (m + 1);
      tr = nr;
      ti = ni;
    }
    e0r = a0r; e0i = a0i; e1r = a1r; e1i = a1i;
    return;
  }
  double s, c;
  sincos(th, &s, &c);
  double it = 1.0 / th, it2 = it * it;
  e0r = s * it;
  e0i = (c - 1.0) * it;
  e1r = c * it2 + s * it - it2;
  e1i = c * it - s * it2;
}

// Contribution of interval [a, a+h] at t: down weight on node n (dr, di), up
// weight on node n+1 (ur, ui), complex, without the 1/pi.
__device__ inline void filon_interval(double a, double h, double t, double& dr, double& di,
                                      double& ur, double& ui) {
  double e0r, e0i, e1r, e1i;
  filon_e(h * t, e0r, e0i, e1r, e1i);
  double s, c;
  sincos(a * t, &s, &c);
  double cr = c, ci = -s;  // e^{-iat}
  ur = h * (cr * e1r - ci * e1i);
  ui = h * (cr * e1i + ci * e1r);
  double xr = e0r - e1r, xi = e0i - e1i;
  dr = h * (cr * xr - ci * xi);
  di = h * (cr * xi + ci * xr);
}

__device__ __forceinline__ double knode(const double* kn, int n) { return n == 0 ? 0.0 : kn[n - 1]; }

// E[j][kk] (FP64 first, then converted): thread per (j, interval n).
__global__ void k_filon_build(const double* __restrict__ knodes, sgl_shape sh, int Kp, int prec,
                              void* __restrict__ out) {
  const int j = blockIdx.y;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // node index n
  const int nq = sh.n_k + 1;
  if (t >= Kp / 2) return;
  const double tj = sh.out_lo + (sh.out_hi - sh.out_lo) * j / (sh.n_out - 1);
  double wr = 0.0, wi = 0.0;  // complex weight A_n(t)
  if (t < nq) {
    const int n = t;
    if (n + 1 < nq) {  // down half of [k_n, k_{n+1}]
      double a = knode(knodes, n), h = knode(knodes, n + 1) - a, dr, di, ur, ui;
      filon_interval(a, h, tj, dr, di, ur, ui);
      wr += dr; wi += di;
    }
    if (n > 0) {  // up half of [k_{n-1}, k_n]
      double a = knode(knodes, n - 1), h = knode(knodes, n) - a, dr, di, ur, ui;
      filon_interval(a, h, tj, dr, di, ur, ui);
      wr += ur; wi += ui;
    }
  }
  // p = (1/pi) Re(phi A) = (1/pi)(Re phi Re A - Im phi Im A)
  double er = wr / kPi, ei = -wi / kPi;
  if (prec == SGL_PREC_FP32) {
    float* o = (float*)out + (size_t)j * Kp + 2 * t;
    o[0] = (float)er;
    o[1] = (float)ei;
  } else {
    const size_t plane = (size_t)sh.n_out * Kp;
    uint16_t* hi = (uint16_t*)out + (size_t)j * Kp + 2 * t;
    uint16_t* lo = hi + plane;
    float fr = (float)er, fi = (float)ei;
    uint16_t a = f32_to_bf16_rn(fr), c = f32_to_bf16_rn(fi);
    hi[0] = a;
    hi[1] = c;
    lo[0] = f32_to_bf16_rn(fr - bf16_to_f32(a));
    lo[1] = f32_to_bf16_rn(fi - bf16_to_f32(c));
  }
}

// Plain path: one workgroup per (j, b), FP64 Filon weights on the fly; the
// 256 threads take contiguous runs of the k intervals and the partial sums
// meet in a fixed-order tree (deterministic).  Output points from `t` when
// non-null (n_t of them), else the shape's uniform grid.
__global__ __launch_bounds__(256) void k_invert_fp64(const double* __restrict__ phi,
                                                     const double* __restrict__ knodes,
                                                     sgl_shape sh, int B,
                                                     const double* __restrict__ t, int n_t,
                                                     double* __restrict__ p) {
  __shared__ double red[256];
  const int j = blockIdx.x, b = blockIdx.y, tt = threadIdx.x;
  const double tj = t ? t[j] : sh.out_lo + (sh.out_hi - sh.out_lo) * j / (sh.n_out - 1);
  const int nint = sh.n_k;                 // intervals between the n_k + 1 nodes
  const int per = (nint + 255) / 256;
  const int n0 = tt * per, n1 = min(nint, n0 + per);
  const double* ph = phi + 2 * (size_t)b * (nint + 1);
  double acc = 0.0;
  for (int n = n0; n < n1; ++n) {
    double a = knode(knodes, n), h = knode(knodes, n + 1) - a, dr, di, ur, ui;
    filon_interval(a, h, tj, dr, di, ur, ui);
    acc += ph[2 * n] * dr - ph[2 * n + 1] * di;
    acc += ph[2 * (n + 1)] * ur - ph[2 * (n + 1) + 1] * ui;
  }
  red[tt] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tt < o) red[tt] += red[tt + o];
    __syncthreads();
  }
  if (tt == 0) p[(size_t)b * n_t + j] = red[0] / kPi;
}

// ----------------------------------------------------------- MFMA GEMM --
// p[b][j] = sum_kk E[j][kk] Phi[b][kk]; E: [M = n_out][Kp], Phi: [N = B][Kp],
// both K-contiguous.  Workgroup tile 32 MT (j) x 64 (b), 4 waves of
// 16 MT x 32, K step 32 through LDS; MT = 2 or 4 is a tuning knob.  FP32: v_mfma_f32_16x16x4_f32.  Split BF16:
// v_mfma_f32_16x16x32_bf16 on (hi hi + hi lo + lo hi).
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kGBN = 64, kGBK = 32, kGLd = kGBK + 4;

// MT = 16-row MFMA tiles per wave along j: the workgroup covers 32 MT rows
// of E (2 waves along j) by 64 columns of Phi (2 waves of 32 along b).
template <int MT>
__global__ __launch_bounds__(256) void k_gemm_f32(const float* __restrict__ E,
                                                  const float* __restrict__ Phi, int M, int N,
                                                  int Kp, float* __restrict__ P) {
  constexpr int BM = 32 * MT;
  __shared__ float sA[BM][kGLd];
  __shared__ float sB[kGBN][kGLd];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * kGBN;
  const int wm = (wave & 1) * 16 * MT, wn = (wave >> 1) * 32;
  f32x4 acc[MT][2];
  for (int a = 0; a < MT; ++a)
    for (int c = 0; c < 2; ++c) acc[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Kp; k0 += kGBK) {
    __syncthreads();
    for (int q = t; q < BM * kGBK / 4; q += 256) {
      int r = q / (kGBK / 4), c4 = (q % (kGBK / 4)) * 4;
      int gm = m0 + r;
      float4 va = (gm < M) ? *(const float4*)(E + (size_t)gm * Kp + k0 + c4) : float4{0, 0, 0, 0};
      sA[r][c4] = va.x; sA[r][c4 + 1] = va.y; sA[r][c4 + 2] = va.z; sA[r][c4 + 3] = va.w;
    }
    for (int q = t; q < kGBN * kGBK / 4; q += 256) {
      int r = q / (kGBK / 4), c4 = (q % (kGBK / 4)) * 4;
      int gn = n0 + r;
      float4 vb = (gn < N) ? *(const float4*)(Phi + (size_t)gn * Kp + k0 + c4) : float4{0, 0, 0, 0};
      sB[r][c4] = vb.x; sB[r][c4 + 1] = vb.y; sB[r][c4 + 2] = vb.z; sB[r][c4 + 3] = vb.w;
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < kGBK; k4 += 4) {
      float bv[2], av[MT];
      for (int c = 0; c < 2; ++c) bv[c] = sB[wn + 16 * c + (lane & 15)][k4 + (lane >> 4)];
      for (int a = 0; a < MT; ++a) av[a] = sA[wm + 16 * a + (lane & 15)][k4 + (lane >> 4)];
      for (int a = 0; a < MT; ++a)
        for (int c = 0; c < 2; ++c)
          acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[c], acc[a][c], 0, 0, 0);
    }
  }
  // D[i][j] of a 16x16 tile: lane holds i = 4 (lane / 16) + r, j = lane % 16
  for (int a = 0; a < MT; ++a)
    for (int c = 0; c < 2; ++c)
      for (int r = 0; r < 4; ++r) {
        int gm = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
        int gn = n0 + wn + c * 16 + (lane & 15);
        if (gm < M && gn < N) P[(size_t)gn * M + gm] = acc[a][c][r];
      }
}

constexpr int kHLd = kGBK + 8;  // bf16 row stride in LDS (shorts)

template <int MT>
__global__ __launch_bounds__(256) void k_gemm_bf16x3(const uint16_t* __restrict__ E,
                                                     const uint16_t* __restrict__ Phi, int M,
                                                     int N, int Kp, float* __restrict__ P) {
  constexpr int BM = 32 * MT;
  __shared__ uint16_t sAh[BM][kHLd], sAl[BM][kHLd];
  __shared__ uint16_t sBh[kGBN][kHLd], sBl[kGBN][kHLd];
  const uint16_t* Eh = E;
  const uint16_t* El = E + (size_t)M * Kp;
  const uint16_t* Bh = Phi;
  const uint16_t* Bl = Phi + (size_t)N * Kp;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * kGBN;
  const int wm = (wave & 1) * 16 * MT, wn = (wave >> 1) * 32;
  f32x4 acc[MT][2];
  for (int a = 0; a < MT; ++a)
    for (int c = 0; c < 2; ++c) acc[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  const uint4 z = {0, 0, 0, 0};
  for (int k0 = 0; k0 < Kp; k0 += kGBK) {
    __syncthreads();
    // a row of 32 bf16 is 4 uint4; 256 threads load 64 rows per pass
    for (int r = t >> 2; r < BM; r += 64) {
      int c8 = (t & 3) * 8, gm = m0 + r;
      uint4 ah = (gm < M) ? *(const uint4*)(Eh + (size_t)gm * Kp + k0 + c8) : z;
      uint4 al = (gm < M) ? *(const uint4*)(El + (size_t)gm * Kp + k0 + c8) : z;
      *(uint4*)&sAh[r][c8] = ah;
      *(uint4*)&sAl[r][c8] = al;
    }
    {
      int r = t >> 2, c8 = (t & 3) * 8, gn = n0 + r;
      uint4 bh = (gn < N) ? *(const uint4*)(Bh + (size_t)gn * Kp + k0 + c8) : z;
      uint4 bl = (gn < N) ? *(const uint4*)(Bl + (size_t)gn * Kp + k0 + c8) : z;
      *(uint4*)&sBh[r][c8] = bh;
      *(uint4*)&sBl[r][c8] = bl;
    }
    __syncthreads();
    // one 16x16x32 step covers the whole K tile: lane supplies k = 8 (lane/16) .. +7
    const int kq = 8 * (lane >> 4);
    bf16x8 bh[2], bl[2];
    for (int c = 0; c < 2; ++c) {
      bh[c] = *(const bf16x8*)&sBh[wn + 16 * c + (lane & 15)][kq];
      bl[c] = *(const bf16x8*)&sBl[wn + 16 * c + (lane & 15)][kq];
    }
    for (int a = 0; a < MT; ++a) {
      bf16x8 ah = *(const bf16x8*)&sAh[wm + 16 * a + (lane & 15)][kq];
      bf16x8 al = *(const bf16x8*)&sAl[wm + 16 * a + (lane & 15)][kq];
      for (int c = 0; c < 2; ++c) {
        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[c], acc[a][c], 0, 0, 0);
        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[c], acc[a][c], 0, 0, 0);
        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[c], acc[a][c], 0, 0, 0);
      }
    }
  }
  for (int a = 0; a < MT; ++a)
    for (int c = 0; c < 2; ++c)
      for (int r = 0; r < 4; ++r) {
        int gm = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
        int gn = n0 + wn + c * 16 + (lane & 15);
        if (gm < M && gn < N) P[(size_t)gn * M + gm] = acc[a][c][r];
      }
}

// ---------------------------------------------------------- diagnostics --
template <typename T>
__device__ inline double block_sum(T v, double* red) {
  const int t = threadIdx.x;
  red[t] = (double)v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

template <typename PT>
__global__ __launch_bounds__(256) void k_diag(const PT* __restrict__ p, sgl_shape sh,
                                              double* __restrict__ diag) {
  __shared__ double red[256];
  const int b = blockIdx.x, t = threadIdx.x;
  const int n = sh.n_out;
  const double h = (sh.out_hi - sh.out_lo) / (n - 1);
  const PT* pb = p + (size_t)b * n;
  auto xat = [&](int j) { return sh.out_lo + h * j; };
  auto wt = [&](int j) { return (j == 0 || j == n - 1) ? 0.5 * h : h; };
  double m0 = 0, m1 = 0, neg = 0, mu = 0, xd = 0;
  for (int j = t; j < n; j += 256) {
    double pj = (double)pb[j], x = xat(j), w = wt(j);
    m0 += w * pj;
    m1 += w * x * pj;
    neg += w * fmax(-pj, 0.0);
    mu += w * exp(x) * pj;
    xd += w * exp(-0.5 * x) * pj;
  }
  double norm = block_sum(m0, red);
  double mean = block_sum(m1, red) / norm;
  double negm = block_sum(neg, red);
  double mum = block_sum(mu, red) / norm;
  double md = block_sum(xd, red) / norm;
  double c2 = 0, c3 = 0, sd = 0;
  for (int j = t; j < n; j += 256) {
    double pj = (double)pb[j] / norm, x = xat(j) - mean, w = wt(j);
    c2 += w * x * x * pj;
    c3 += w * x * x * x * pj;
    double y = exp(-0.5 * xat(j)) - md;
    sd += w * y * y * pj;
  }
  double var = block_sum(c2, red);
  double skw = block_sum(c3, red) / pow(var, 1.5);
  double sdl = sqrt(block_sum(sd, red)) / md;
  // P(mu > 2): trapezoid over [ln 2, out_hi] with linear interpolation at ln 2
  double tail = 0.0;
  const double x0 = log(2.0);
  int i = (int)floor((x0 - sh.out_lo) / h) + 1;  // first node > x0
  if (i < 1) i = 1;
  if (i <= n - 1) {
    for (int j = i + t; j < n - 1; j += 256)
      tail += 0.5 * h * ((double)pb[j] + (double)pb[j + 1]);
  }
  tail = block_sum(tail, red);
  if (t == 0) {
    if (i <= n - 1) {
      double xa = xat(i - 1), f = (x0 - xa) / h;
      double p0 = (double)pb[i - 1] + f * ((double)pb[i] - (double)pb[i - 1]);
      tail += 0.5 * (p0 + (double)pb[i]) * (xat(i) - x0);
    }
    double* d = diag + (size_t)b * 8;
    d[0] = norm; d[1] = mean; d[2] = var; d[3] = skw;
    d[4] = sdl; d[5] = mum; d[6] = tail / norm; d[7] = negm;
  }
}

// ------------------------------------------------------------- launchers --
int launch_quadform(const double* cov, const double* cosmo, const void* amp, int B,
                    const sgl_shape& sh, int prec, double* psi2, hipStream_t st) {
  if (prec == SGL_PREC_FP64) {
    dim3 grid((sh.n_k + kQfTile - 1) / kQfTile, B);
    hipLaunchKernelGGL(k_quadform, grid, dim3(256), 0, st, cov, cosmo,
                       (const double*)amp, sh.n_z_max, sh.n_k, psi2);
  } else {
    // batched path: amplitudes and Psi_2 on the sub-grid
    const int per_wg = 4 * kQmTiles * 16, n_kc = psi2_nodes(sh);
    dim3 grid((n_kc + per_wg - 1) / per_wg, B);
    if (prec == SGL_PREC_BF16X3)
      hipLaunchKernelGGL(k_quadform_bf16x3, grid, dim3(256), 0, st, cov, cosmo,
                         (const float*)amp, sh.n_z_max, n_kc, psi2);
    else
      hipLaunchKernelGGL(k_quadform_mfma, grid, dim3(256), 0, st, cov, cosmo,
                         (const float*)amp, sh.n_z_max, n_kc, psi2);
  }
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_phi_pack(const double* lam, const double* psi2, const double* knodes,
                    const sgl_shape& sh, int B, int prec, void* out, hipStream_t st) {
  int Kp = k_padded(sh);
  size_t n = (size_t)B * (Kp / 2);
  hipLaunchKernelGGL(k_phi_pack, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, lam, psi2,
                     knodes, sh, B, prec, Kp, out);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_psi2_interp(const double* psi2c, const sgl_shape& sh, int B, double* psi2,
                       hipStream_t st) {
  const size_t n = (size_t)B * sh.n_k;
  hipLaunchKernelGGL(k_psi2_interp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, psi2c, sh,
                     B, psi2);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_filon(const double* knodes, const sgl_shape& sh, int prec, void* out, hipStream_t st) {
  int Kp = k_padded(sh);
  dim3 grid((Kp / 2 + 255) / 256, sh.n_out);
  hipLaunchKernelGGL(k_filon_build, grid, dim3(256), 0, st, knodes, sh, Kp, prec, out);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_invert_fp64(const double* phi, const double* knodes, const sgl_shape& sh, int B,
                       const double* t, int n_t, double* p, hipStream_t st) {
  if (t == nullptr) n_t = sh.n_out;
  if (n_t < 1) return SGL_OK;
  dim3 grid(n_t, B);
  hipLaunchKernelGGL(k_invert_fp64, grid, dim3(256), 0, st, phi, knodes, sh, B, t, n_t, p);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_gemm(const void* E, const void* Phi, const sgl_shape& sh, int B, int prec, float* P,
                hipStream_t st) {
  int Kp = k_padded(sh);
  const bool wide = g_sgl_tune[SGL_TUNE_GEMM_BM] == 1;
  const int BM = wide ? 128 : 64;
  dim3 grid((sh.n_out + BM - 1) / BM, (B + kGBN - 1) / kGBN);
  const float* Ef = (const float*)E;
  const float* Pf = (const float*)Phi;
  const uint16_t* Eb = (const uint16_t*)E;
  const uint16_t* Pb = (const uint16_t*)Phi;
  if (prec == SGL_PREC_FP32) {
    if (wide) hipLaunchKernelGGL(k_gemm_f32<4>, grid, dim3(256), 0, st, Ef, Pf, sh.n_out, B, Kp, P);
    else hipLaunchKernelGGL(k_gemm_f32<2>, grid, dim3(256), 0, st, Ef, Pf, sh.n_out, B, Kp, P);
  } else {
    if (wide) hipLaunchKernelGGL(k_gemm_bf16x3<4>, grid, dim3(256), 0, st, Eb, Pb, sh.n_out, B, Kp, P);
    else hipLaunchKernelGGL(k_gemm_bf16x3<2>, grid, dim3(256), 0, st, Eb, Pb, sh.n_out, B, Kp, P);
  }
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_diag(const void* p, int is_double, const sgl_shape& sh, int B, double* diag,
                hipStream_t st) {
  if (is_double)
    hipLaunchKernelGGL(k_diag<double>, dim3(B), dim3(256), 0, st, (const double*)p, sh, diag);
  else
    hipLaunchKernelGGL(k_diag<float>, dim3(B), dim3(256), 0, st, (const float*)p, sh, diag);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

}  // namespace sgl
