// Positive ("vp") clustering closure of the plain path: the latent chain's
// transitions, the calibration of its variance scale, and the forward
// recursion that gives ln R(k), the closure's replacement of Psi_2.
//
// The reference's positive closure (pos_closure.py:1-733) draws a latent
// state U_s per slice from a stationary chain on 48 Gauss-Hermite nodes and
// modulates the jump rates of each group g of cells by lambda_g(U_s):
//   ln R(k) = ln E exp( sum_s sum_g (lambda_g(U_s) - 1) Y_sg(k) ).
// Here the groups are the mass rows of a slice (halo rows, then filament
// rows), whose cells share the bias beta_g; the clustering is the slice
// covariance C of the smoothed field, so the loading is phi_g = beta_g
// sqrt(C_ss), the chain correlation rho_s = C_{s,s+1} / sqrt(C_ss C_{s+1,s+1}),
// and eta is fitted so that the chain's variance of sum_g lambda_g wk_g
// equals the two-point term's k^2 coefficient A^T C A (A_s = sum_g beta_g wk_g).
#include "sgl_amd_internal.h"

namespace sgl {

namespace {

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

}  // namespace

// Reversible transition slice s -> s + 1 with the node weights exactly
// stationary (pos_closure.py:251-267): J_ij = w_i w_j exp(Gaussian-copula log
// density ratio at correlation rho), symmetric balancing v <- v sqrt(w / (v J v))
// until every marginal v_i (J v)_i matches w_i to 3e-14, K_ij = v_i J_ij v_j / w_i.
// One wave per (slice pair, instance); lane i < 48 owns row i.
__global__ __launch_bounds__(64) void k_vp_transition(const double* __restrict__ cov,
                                                      const double* __restrict__ cosmo,
                                                      sgl_shape sh, VpLayout V,
                                                      double* __restrict__ vp,
                                                      int* __restrict__ status) {
  __shared__ double J[kVpN][kVpN + 1];
  __shared__ double v[kVpN];
  const int s = blockIdx.x, b = blockIdx.y, i = threadIdx.x;
  const int NZ = sh.n_z_max;
  const int n_z = min((int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10], sh.n_z_max);
  if (s >= n_z - 1) return;
  const double* C = cov + (size_t)b * NZ * NZ;
  const double rho = C[(size_t)s * NZ + s + 1] / sqrt(C[(size_t)s * NZ + s] * C[(size_t)(s + 1) * NZ + s + 1]);
  const double den = 1.0 / (2.0 * (1.0 - rho * rho));
  const bool own = i < kVpN;
  const double ui = own ? kGhU[i] : 0.0, wi = own ? kGhW[i] : 1.0;
  if (own) {
    for (int j = 0; j < kVpN; ++j) {
      const double uj = kGhU[j];
      J[i][j] = wi * kGhW[j] * exp((2.0 * rho * ui * uj - rho * rho * (ui * ui + uj * uj)) * den);
    }
    v[i] = 1.0;
  }
  double err = 1.0;
  for (int it = 0; it < 10000; ++it) {
    __syncthreads();
    double marg = 1.0;
    if (own) {
      double jv = 0.0;
      for (int j = 0; j < kVpN; ++j) jv = fma(J[i][j], v[j], jv);
      marg = v[i] * jv;
    }
    err = wave_max(own ? fabs(marg / wi - 1.0) : 0.0);
    if (err < 3e-14) break;
    __syncthreads();
    if (own) v[i] *= sqrt(wi / marg);
  }
  if (i == 0 && err > 1e-10) atomicOr(status, kStatusVpTransition);
  __syncthreads();
  if (own) {
    double* K = vp + V.K + (((size_t)b * NZ + s) * kVpN + i) * kVpN;
    const double vi = v[i] / wi;
    for (int j = 0; j < kVpN; ++j) K[j] = vi * J[i][j] * v[j];
  }
}

// Calibration and rate table, one workgroup per instance.
//
// Rates (pos_closure.py:273-308): for relative variance v_g = eta f_s phi_g^2,
// lambda_g(u) = (u - t)^+ / E_w[(u - t)^+] with the threshold t solving
// Var_w lambda = v on the node segment [u_{j-1}, u_j) (the quadratic
// q2 t^2 + q1 t + q0 = 0 in the suffix sums A_j = sum_{i>=j} w_i u_i,
// B_j = sum w_i, C_j = sum w_i u_i^2), segment j = #{rv <= v} with
// rv_k = E2/E1^2 - 1 at t = u_k; j = 0 is the linear rate 1 + sqrt(v) u.
// Stored as lambda(u_j) = u_j minv - tm for j >= bucket (minv = 1/m,
// tm = t/m; linear: minv = sqrt(v), tm = -1).
//
// Chain variance (pos_closure.py:389-403) of sum_s m_s(U_s),
// m_s(u_j) = sum_g lambda_sg(u_j) wk_sg; eta from false position (Illinois)
// in ln eta on variance(eta) = target inside the reference's bracket (start
// [1/1.5, 1.5], widened by 1.5 per side, pos_closure.py:575-591), to 4e-14
// in ln eta.
constexpr int kVpGroupsPerLane = 2 * kMaxMasses / 64;

struct VpStage {
  double u[kVpN], w[kVpN], A[kVpN], Bs[kVpN], Cs[kVpN], rv[kVpN - 1];
  double f[kMaxSlices], d[kMaxSlices];
  double ms[kMaxSlices][kVpN];
  double h[2][kVpN];
  double hb[4][2][kVpN];                 // per wave: bucket sums H1, H0 of one slice
  double red[256];
  int cap;
};

__device__ __forceinline__ double block_sum(VpStage& S, double x) {
  const int t = threadIdx.x;
  __syncthreads();
  S.red[t] = x;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) S.red[t] += S.red[t + o];
    __syncthreads();
  }
  const double r = S.red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double group_beta(const double* rows, const double* frows,
                                             const sgl_shape& sh, size_t si, int g) {
  if (g < sh.n_m) return rows[(si * sh.n_m + g) * kRowStride + kRowBeta];
  return frows[(si * sh.n_m + (g - sh.n_m)) * kFilRowStride + 3];
}

// Rate of relative variance v: (minv, tm, bucket); false beyond the cap.
__device__ inline bool vp_rate(const VpStage& S, double v, double& minv, double& tm, int& bk) {
  if (!(v < S.rv[kVpN - 2])) {
    minv = 0.0; tm = -1.0; bk = 0;
    return false;
  }
  int lo = 0, hi = kVpN - 1;            // j = #{k : rv_k <= v} in [0, 47)
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (S.rv[mid] <= v) lo = mid + 1; else hi = mid;
  }
  const int j = lo;
  bk = j;
  if (j == 0) {
    minv = sqrt(v);
    tm = -1.0;
    return true;
  }
  const double a = S.A[j], b = S.Bs[j], c = S.Cs[j];
  const double q2 = b - (1.0 + v) * b * b;
  const double q1 = 2.0 * a * ((1.0 + v) * b - 1.0);
  const double q0 = c - (1.0 + v) * a * a;
  const double lo_u = S.u[j - 1], hi_u = S.u[j];
  double t;
  if (q2 != 0.0) {
    const double disc = sqrt(fmax(q1 * q1 - 4.0 * q2 * q0, 0.0));
    const double r1 = (-q1 - disc) / (2.0 * q2), r2 = (-q1 + disc) / (2.0 * q2);
    t = (r1 >= lo_u && r1 < hi_u) ? r1 : r2;
  } else {
    t = -q0 / q1;
  }
  const double m = a - t * b;            // E_w[(u - t)^+] over the nodes above t
  minv = 1.0 / m;
  tm = t / m;
  return true;
}

__global__ __launch_bounds__(256) void k_vp_setup(const double* __restrict__ rows,
                                                  const double* __restrict__ frows,
                                                  const double* __restrict__ cov,
                                                  const double* __restrict__ cosmo, sgl_shape sh,
                                                  VpLayout V, double* __restrict__ vp,
                                                  int* __restrict__ status) {
  __shared__ VpStage S;
  const int t = threadIdx.x, b = blockIdx.x;
  const int NZ = sh.n_z_max, G = vp_groups(sh);
  const int n_z = min((int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10], sh.n_z_max);
  const double* C = cov + (size_t)b * NZ * NZ;
  const double* wk = vp + V.wk + (size_t)b * NZ * G;
  double* tab = vp + V.tab + (size_t)b * NZ * G * 3;
  const double* K = vp + V.K + (size_t)b * NZ * kVpN * kVpN;
  if (t < kVpN) {
    S.u[t] = kGhU[t];
    S.w[t] = kGhW[t];
  }
  if (t == 0) S.cap = 0;
  __syncthreads();
  if (t == 0) {                          // suffix sums, summed from the top node down
    double a = 0.0, bb = 0.0, c = 0.0;
    for (int j = kVpN - 1; j >= 0; --j) {
      a += S.w[j] * S.u[j];
      bb += S.w[j];
      c += S.w[j] * S.u[j] * S.u[j];
      S.A[j] = a; S.Bs[j] = bb; S.Cs[j] = c;
    }
  }
  for (int s = t; s < n_z; s += 256) S.d[s] = sqrt(C[(size_t)s * NZ + s]);
  __syncthreads();
  if (t < kVpN - 1) {
    const double tb = S.u[t];
    const double E1 = S.A[t + 1] - tb * S.Bs[t + 1];
    const double E2 = S.Cs[t + 1] - 2.0 * tb * S.A[t + 1] + tb * tb * S.Bs[t + 1];
    S.rv[t] = E2 / (E1 * E1) - 1.0;
  }
  for (int s = t; s < n_z; s += 256) {   // f_s = (1 - rbar) / (1 + rbar), rbar the mean adjacent rho
    double sum = 0.0;
    int n = 0;
    if (s >= 1) { sum += C[(size_t)(s - 1) * NZ + s] / (S.d[s - 1] * S.d[s]); ++n; }
    if (s + 1 < n_z) { sum += C[(size_t)s * NZ + s + 1] / (S.d[s] * S.d[s + 1]); ++n; }
    const double rbar = n ? sum / n : 0.0;
    S.f[s] = (1.0 - rbar) / (1.0 + rbar);
  }
  // target = A^T C A, A_s = sum_g beta_g wk_g (A kept in ms[.][0] for the moment)
  for (int s = t; s < n_z; s += 256) {
    double a = 0.0;
    for (int g = 0; g < G; ++g)
      a = fma(group_beta(rows, frows, sh, (size_t)b * NZ + s, g), wk[(size_t)s * G + g], a);
    S.ms[s][0] = a;
  }
  __syncthreads();
  double part = 0.0;
  for (int s = t; s < n_z; s += 256) {
    double r = 0.0;
    for (int j = 0; j < n_z; ++j) r = fma(C[(size_t)s * NZ + j], S.ms[j][0], r);
    part = fma(S.ms[s][0], r, part);
  }
  const double target = block_sum(S, part);

  // variance(eta) - target; also writes the rate table for eta
  auto eval = [&](double le) -> double {
    const double eta = exp(le);
    bool bad = false;
    // One wave per slice: lane l owns the groups g = l + 64 q; the rates'
    // bucket sums H1_j = sum_{bucket g = j} wk_g / m_g, H0_j = sum t_g wk_g / m_g
    // by wave reductions over the buckets present, then
    // m_s(u_i) = u_i sum_{j<=i} H1_j - sum_{j<=i} H0_j.
    const int lane = t & 63, wave = t >> 6;
    for (int s = wave; s < n_z; s += 4) {
      double a1[kVpGroupsPerLane], a0[kVpGroupsPerLane];
      int bq[kVpGroupsPerLane];
      int bmin = kVpN, bmax = -1;
      const double vs = eta * S.f[s] * S.d[s] * S.d[s];
#pragma unroll
      for (int q = 0; q < kVpGroupsPerLane; ++q) {
        const int g = lane + 64 * q;
        a1[q] = a0[q] = 0.0;
        bq[q] = kVpN;
        if (g < G) {
          const double beta = group_beta(rows, frows, sh, (size_t)b * NZ + s, g);
          double minv, tm;
          int bk;
          bad |= !vp_rate(S, vs * beta * beta, minv, tm, bk);
          const size_t o = 3 * ((size_t)s * G + g);
          tab[o] = minv;
          tab[o + 1] = tm;
          tab[o + 2] = (double)bk;
          const double wg = wk[(size_t)s * G + g];
          a1[q] = minv * wg;
          a0[q] = tm * wg;
          bq[q] = bk;
          bmin = min(bmin, bk);
          bmax = max(bmax, bk);
        }
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        bmin = min(bmin, __shfl_xor(bmin, o));
        bmax = max(bmax, __shfl_xor(bmax, o));
      }
      for (int j = lane; j < kVpN; j += 64) {
        S.hb[wave][0][j] = 0.0;
        S.hb[wave][1][j] = 0.0;
      }
      for (int j = bmin; j <= bmax; ++j) {
        double h1 = 0.0, h0 = 0.0;
#pragma unroll
        for (int q = 0; q < kVpGroupsPerLane; ++q)
          if (bq[q] == j) { h1 += a1[q]; h0 += a0[q]; }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          h1 += __shfl_xor(h1, o);
          h0 += __shfl_xor(h0, o);
        }
        if (lane == 0) { S.hb[wave][0][j] = h1; S.hb[wave][1][j] = h0; }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (lane < kVpN) {
        double p1 = 0.0, p0 = 0.0;
        for (int j = 0; j <= lane; ++j) { p1 += S.hb[wave][0][j]; p0 += S.hb[wave][1][j]; }
        S.ms[s][lane] = fma(S.u[lane], p1, -p0);
      }
      __builtin_amdgcn_wave_barrier();
    }
    if (bad) S.cap = 1;
    __syncthreads();
    // chain: var = sum_s w.mu_s^2 + 2 sum_{s>=1} mu_s . (K_{s-1}^T h_{s-1}),
    // h_s = K_{s-1}^T h_{s-1} + w mu_s
    double acc = 0.0;
    for (int s = 0; s < n_z; ++s) {
      const int cur = s & 1;
      if (t < kVpN) {
        double mean = 0.0;
        for (int i = 0; i < kVpN; ++i) mean = fma(S.w[i], S.ms[s][i], mean);
        const double mu = S.ms[s][t] - mean;
        acc = fma(S.w[t] * mu, mu, acc);
        double hp = 0.0;
        if (s > 0) {
          const double* Ks = K + (size_t)(s - 1) * kVpN * kVpN;
          for (int i = 0; i < kVpN; ++i) hp = fma(Ks[(size_t)i * kVpN + t], S.h[cur ^ 1][i], hp);
          acc = fma(2.0 * mu, hp, acc);
        }
        S.h[cur][t] = hp + S.w[t] * mu;
      }
      __syncthreads();
    }
    const double var = block_sum(S, acc);
    const bool capped = S.cap != 0;
    __syncthreads();
    if (t == 0) S.cap = 0;
    __syncthreads();
    return capped ? 1e300 : var - target;
  };

  // bracket from [1/1.5, 1.5], widened by 1.5 per side while it stays inside
  // eta in [1e-3, 1e4]; a root outside that range is an error, as in
  // pos_closure.Closure.calibrate (pos_closure.py:575-591)
  const double step = log(1.5), llo = log(1e-3), lhi = log(1e4);
  double a = -step, bb = step;
  double fa = eval(a), fb = eval(bb);
  int flag = 0;
  while (fa * fb > 0.0) {
    if (fa > 0.0 && a > llo) {
      bb = a; fb = fa;
      a -= step; fa = eval(a);
    } else if (fb < 0.0 && bb < lhi) {
      a = bb; fa = fb;
      bb += step; fb = eval(bb);
    } else {
      flag = kStatusVpRate;
      break;
    }
  }
  // Illinois false position inside the bracket (superlinear for this smooth,
  // monotone variance; a capped value forces a bisection step), to 4e-14 in
  // ln eta
  double x0 = a, f0 = fa, x1 = bb, f1 = fb;
  for (int it = 0; it < 100 && fabs(x1 - x0) > 4e-14; ++it) {
    double c = (f0 >= 1e299 || f1 >= 1e299) ? 0.5 * (x0 + x1) : x1 - f1 * (x1 - x0) / (f1 - f0);
    if (!(c > fmin(x0, x1) && c < fmax(x0, x1))) c = 0.5 * (x0 + x1);
    const double fc = eval(c);
    if (fc == 0.0) { x0 = x1 = c; break; }
    if ((fc > 0.0) != (f1 > 0.0)) { x0 = x1; f0 = f1; } else { f0 *= 0.5; }
    x1 = c;
    f1 = fc;
  }
  const double le = x1;
  if (eval(le) >= 1e299) flag |= kStatusVpRate;
  if (t == 0) {
    vp[V.eta + b] = exp(le);
    vp[V.eta + gridDim.x + b] = target;
    if (flag) atomicOr(status, flag);
  }
}

// ln R(k) by the forward recursion over the slices (pos_closure.py:183-240):
// alpha_0 = w; per slice alpha <- K_{s-1}^T alpha (s > 0), then
// alpha_j *= exp(z_j - max_j Re z_j) with z_j = zpart_j - ytot, renormalised by
// max_j |alpha_j|, the log scale accumulated; ln R = scale + ln sum_j alpha_j.
// One wave per k node, lane j < 48 holds alpha_j; the matrix-vector product
// takes alpha_i from lane i by readlane and K_{s-1}[i][j] from LDS (shared by
// the block's 8 waves, which read neighbouring k nodes of each zpart row).
__device__ __forceinline__ double read_lane(double x, int l) {
  const long long v = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readlane((int)v, l);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

constexpr int kChainWaves = 8;

__global__ __launch_bounds__(64 * kChainWaves) void k_vp_chain(const double* __restrict__ cosmo,
                                                               sgl_shape sh, VpLayout V,
                                                               const double* __restrict__ vp,
                                                               double* __restrict__ lnr) {
  __shared__ double Ks[kVpN * kVpN];     // K_{s-1}[i][j]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, b = blockIdx.y;
  const int n = blockIdx.x * kChainWaves + wave;
  const int NZ = sh.n_z_max, n_k = sh.n_k;
  const int n_z = min((int)cosmo[(size_t)b * SGL_COSMO_STRIDE + 10], sh.n_z_max);
  const bool own = lane < kVpN, live = n < n_k;
  double ar = own ? kGhW[lane] : 0.0, ai = 0.0;
  double scale = 0.0;
  for (int s = 0; s < n_z; ++s) {
    const size_t si = (size_t)b * NZ + s;
    if (s > 0) {
      __syncthreads();
      const double* K = vp + V.K + (si - 1) * kVpN * kVpN;
      for (int q = threadIdx.x; q < kVpN * kVpN; q += 64 * kChainWaves) Ks[q] = K[q];
      __syncthreads();
      double nr = 0.0, ni = 0.0;
      for (int i = 0; i < kVpN; ++i) {
        const double kij = own ? Ks[i * kVpN + lane] : 0.0;
        nr = fma(kij, read_lane(ar, i), nr);
        ni = fma(kij, read_lane(ai, i), ni);
      }
      ar = nr;
      ai = ni;
    }
    double zr = 0.0, zi = 0.0;
    if (own && live) {
      const double2 z = reinterpret_cast<const double2*>(vp + V.zpart + si * kVpN * (size_t)n_k * 2)
          [(size_t)lane * n_k + n];
      const double2 y = reinterpret_cast<const double2*>(vp + V.ytot + si * (size_t)n_k * 2)[n];
      zr = z.x - y.x;
      zi = z.y - y.y;
    }
    const double mx = wave_max(own ? zr : -1e308);
    if (own) {
      const double e = exp(zr - mx);
      double sn, cs;
      sincos(zi, &sn, &cs);
      const double er = e * cs, ei = e * sn;
      const double xr = ar * er - ai * ei;
      ai = ar * ei + ai * er;
      ar = xr;
    }
    const double nrm = wave_max(own ? sqrt(ar * ar + ai * ai) : 0.0);
    const double inv = 1.0 / nrm;
    ar *= inv;
    ai *= inv;
    scale += mx + log(nrm);
  }
  double sr = ar, sn = ai;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    sr += __shfl_xor(sr, o);
    sn += __shfl_xor(sn, o);
  }
  if (live && lane == 0) {
    lnr[2 * ((size_t)b * n_k + n)] = scale + 0.5 * log(sr * sr + sn * sn);
    lnr[2 * ((size_t)b * n_k + n) + 1] = atan2(sn, sr);
  }
}

int launch_vp_setup(const double* rows, const double* frows, const double* cov,
                    const double* cosmo, const sgl_shape& sh, int B, const VpLayout& V, double* vp,
                    int* status, hipStream_t st) {
  if (sh.n_z_max > 1)
    hipLaunchKernelGGL(k_vp_transition, dim3(sh.n_z_max - 1, B), dim3(64), 0, st, cov, cosmo, sh,
                       V, vp, status);
  hipLaunchKernelGGL(k_vp_setup, dim3(B), dim3(256), 0, st, rows, frows, cov, cosmo, sh, V, vp,
                     status);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int launch_vp_chain(const double* cosmo, const sgl_shape& sh, int B, const VpLayout& V,
                    const double* vp, double* lnr, hipStream_t st) {
  hipLaunchKernelGGL(k_vp_chain, dim3((sh.n_k + kChainWaves - 1) / kChainWaves, B),
                     dim3(64 * kChainWaves), 0, st, cosmo, sh, V, vp, lnr);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

}  // namespace sgl
