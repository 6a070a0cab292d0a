// Surrogate MLP kernels: one BF16 MFMA GEMM with strided operands and fused
// epilogues (bias + SiLU, bias, SiLU derivative, plain), plus the bias
// gradient reduction and the FP32 -> BF16 conversion.  Forward and both
// backward products of a linear layer are this GEMM with different strides:
//   forward     Y[n][m]  = sum_k X[n][k] W[m][k]           (A = X, B = W)
//   grad input  dX[n][k] = sum_m dZ[n][m] W[m][k]          (A = dZ, B = W^T)
//   grad weight dW[m][k] = sum_n dZ[n][m] X[n][k]          (A = dZ^T, B = X^T)
// Accumulation in FP32 on v_mfma_f32_16x16x32_bf16.
#include "sgl_amd_internal.h"

namespace sgl {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__device__ __forceinline__ float from_bf16(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
  float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

constexpr int kMBM = 64, kMBN = 64, kMBK = 32, kMLd = kMBK + 8;

struct GemmArgs {
  const uint16_t* A;  // bf16
  const uint16_t* B;  // bf16
  long lda_m, lda_k, ldb_n, ldb_k;
  int M, N, K;
  int epi;             // SGL_MLP_EPI_*
  const float* bias;   // [N] (epilogues with bias)
  const float* Z;      // [M][ldz] pre-activation (dsilu epilogue)
  float* Zout;         // [M][ldc] pre-activation out (bias+silu epilogue), may be null
  float* C;            // [M][ldc] fp32 out, may be null
  uint16_t* Cb;        // [M][ldc] bf16 out, may be null
  long ldc, ldz;
  float beta;          // C = beta * C + result (plain epilogue)
};

__device__ __forceinline__ void load_tile(uint16_t (*s)[kMLd], const uint16_t* __restrict__ P,
                                          long ld_r, long ld_k, int r0, int R, int k0, int K) {
  const int t = threadIdx.x;
  if (ld_k == 1 && (k0 + kMBK) <= K && (ld_r % 8) == 0) {
    // 64 rows x 32 bf16 = 256 x uint4
    int r = t >> 2, c8 = (t & 3) * 8;
    uint4 v = {0, 0, 0, 0};
    if (r0 + r < R) v = *(const uint4*)(P + (long)(r0 + r) * ld_r + k0 + c8);
    *(uint4*)&s[r][c8] = v;
  } else {
    for (int q = t; q < kMBM * kMBK; q += 256) {
      int r = q % kMBM, c = q / kMBM;  // consecutive threads along rows (ld_r may be 1)
      uint16_t v = 0;
      if (r0 + r < R && k0 + c < K) v = P[(long)(r0 + r) * ld_r + (long)(k0 + c) * ld_k];
      s[r][c] = v;
    }
  }
}

__global__ __launch_bounds__(256) void k_mlp_gemm(GemmArgs g) {
  __shared__ uint16_t sA[kMBM][kMLd];
  __shared__ uint16_t sB[kMBN][kMLd];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int m0 = blockIdx.x * kMBM, n0 = blockIdx.y * kMBN;
  const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
  f32x4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c) acc[a][c] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += kMBK) {
    __syncthreads();
    load_tile(sA, g.A, g.lda_m, g.lda_k, m0, g.M, k0, g.K);
    load_tile(sB, g.B, g.ldb_n, g.ldb_k, n0, g.N, k0, g.K);
    __syncthreads();
    const int kq = 8 * (lane >> 4);
    bf16x8 a0 = *(const bf16x8*)&sA[wm + (lane & 15)][kq];
    bf16x8 a1 = *(const bf16x8*)&sA[wm + 16 + (lane & 15)][kq];
    bf16x8 b0 = *(const bf16x8*)&sB[wn + (lane & 15)][kq];
    bf16x8 b1 = *(const bf16x8*)&sB[wn + 16 + (lane & 15)][kq];
    acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1, acc[1][1], 0, 0, 0);
  }
  for (int a = 0; a < 2; ++a)
    for (int c = 0; c < 2; ++c)
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
        const int gn = n0 + wn + c * 16 + (lane & 15);
        if (gm >= g.M || gn >= g.N) continue;
        float v = acc[a][c][r];
        const long o = (long)gm * g.ldc + gn;
        if (g.epi == SGL_MLP_EPI_BIAS_SILU) {
          float z = v + g.bias[gn];
          if (g.Zout) g.Zout[o] = z;
          v = silu(z);
        } else if (g.epi == SGL_MLP_EPI_BIAS) {
          v += g.bias[gn];
        } else if (g.epi == SGL_MLP_EPI_DSILU) {
          v *= dsilu(g.Z[(long)gm * g.ldz + gn]);
        } else if (g.beta != 0.0f && g.C) {
          v += g.beta * g.C[o];
        }
        if (g.C) g.C[o] = v;
        if (g.Cb) g.Cb[o] = to_bf16(v);
      }
}

// db[j] = sum_n dZ[n][j] (fixed-order per column; FP32 values, FP64 sum)
__global__ void k_mlp_colsum(const float* __restrict__ dZ, int N, int M, float* __restrict__ db) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= M) return;
  double s = 0.0;
  for (int n = 0; n < N; ++n) s += dZ[(long)n * M + j];
  db[j] = (float)s;
}

__global__ void k_to_bf16(const float* __restrict__ x, size_t n, uint16_t* __restrict__ y) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = to_bf16(x[i]);
}

// Cubic Lagrange in the node index (uniform in ln k) from n_c coarse nodes
// onto the n_k fine nodes; fine node n sits at coarse coordinate
// u = n (n_c - 1) / (n_k - 1).
__global__ void k_lnphi_interp(const float* __restrict__ y, long ldy, int n_c, int n_k, int B,
                               double* __restrict__ lam) {
  const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (size_t)B * n_k) return;
  const int n = (int)(gid % n_k);
  const int b = (int)(gid / n_k);
  const double u = (double)n * (n_c - 1) / (n_k - 1);
  int j = (int)floor(u) - 1;
  if (j < 0) j = 0;
  if (j > n_c - 4) j = n_c - 4;
  const double s = u - j;  // in [0, 3]
  const double l0 = -(s - 1) * (s - 2) * (s - 3) / 6.0, l1 = s * (s - 2) * (s - 3) / 2.0;
  const double l2 = -s * (s - 1) * (s - 3) / 2.0, l3 = s * (s - 1) * (s - 2) / 6.0;
  const float* yb = y + (size_t)b * ldy;
  double re = l0 * yb[2 * j] + l1 * yb[2 * (j + 1)] + l2 * yb[2 * (j + 2)] + l3 * yb[2 * (j + 3)];
  double im = l0 * yb[2 * j + 1] + l1 * yb[2 * (j + 1) + 1] + l2 * yb[2 * (j + 2) + 1] +
              l3 * yb[2 * (j + 3) + 1];
  lam[2 * gid] = re;
  lam[2 * gid + 1] = im;
}

}  // namespace sgl

using namespace sgl;

extern "C" {

sgl_status sgl_mlp_gemm(const void* A, long lda_m, long lda_k, const void* B, long ldb_n,
                        long ldb_k, int32_t M, int32_t N, int32_t K, int32_t epi,
                        const float* bias, const float* Z, long ldz, float* Zout, float* C,
                        void* Cb, long ldc, float beta, void* stream) {
  if (!A || !B || M < 1 || N < 1 || K < 1 || (!C && !Cb)) return SGL_ERR_INVALID;
  if (epi < SGL_MLP_EPI_PLAIN || epi > SGL_MLP_EPI_DSILU) return SGL_ERR_INVALID;
  if ((epi == SGL_MLP_EPI_BIAS_SILU || epi == SGL_MLP_EPI_BIAS) && !bias) return SGL_ERR_INVALID;
  if (epi == SGL_MLP_EPI_DSILU && !Z) return SGL_ERR_INVALID;
  if (ldc < N) return SGL_ERR_INVALID;
  GemmArgs g;
  g.A = (const uint16_t*)A; g.B = (const uint16_t*)B;
  g.lda_m = lda_m; g.lda_k = lda_k; g.ldb_n = ldb_n; g.ldb_k = ldb_k;
  g.M = M; g.N = N; g.K = K; g.epi = epi; g.bias = bias; g.Z = Z; g.Zout = Zout;
  g.C = C; g.Cb = (uint16_t*)Cb; g.ldc = ldc; g.ldz = ldz; g.beta = beta;
  dim3 grid((M + kMBM - 1) / kMBM, (N + kMBN - 1) / kMBN);
  hipLaunchKernelGGL(k_mlp_gemm, grid, dim3(256), 0, (hipStream_t)stream, g);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_mlp_colsum(const float* dZ, int32_t N, int32_t M, float* db, void* stream) {
  if (!dZ || !db || N < 1 || M < 1) return SGL_ERR_INVALID;
  hipLaunchKernelGGL(k_mlp_colsum, dim3((M + 255) / 256), dim3(256), 0, (hipStream_t)stream, dZ, N,
                     M, db);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_to_bf16(const float* x, size_t n, void* y, void* stream) {
  if (!x || !y) return SGL_ERR_INVALID;
  if (n == 0) return SGL_OK;
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, x, n, (uint16_t*)y);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

sgl_status sgl_lnphi_interp(const float* y, long ldy, int32_t n_c, int32_t B,
                            const sgl_shape* shape, double* lam, void* stream) {
  if (!y || !shape || !lam || B < 1 || n_c < 4 || ldy < 2L * n_c) return SGL_ERR_INVALID;
  if (shape->n_k < n_c) return SGL_ERR_INVALID;
  size_t n = (size_t)B * shape->n_k;
  hipLaunchKernelGGL(k_lnphi_interp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, y, ldy, n_c, shape->n_k, B, lam);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

}  // extern "C"
