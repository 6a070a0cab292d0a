// Subhalos on the reference-shape path (SGL_FLAG_SUBHALOS): the exact
// decorated-host sector of the reference in its default configuration
// (pipeline.py:171-183: sector "exact", method "moments", clumps truncated at
// their own r200, circular clumps).
//
//   k_sub_hosts    one workgroup per (shell, mass) row: the JvdB14 subhalo
//                  mass fraction (formation redshift by bisection on the
//                  growth), the clump mass function on 28 log bins of
//                  psi = m / M_vir, the clumps' NFW parameters, the clump
//                  number selected to the host's r200 sphere and its projected
//                  density (SubhaloModel.fs / host, subhalos.py:672-772), the
//                  ray-distance grid and every clump cell's kappa, gamma
//                  (ExactSector._clump_cells, subhalos_exact.py:1234-1304);
//   (ref2d.hip)    the largest host-grid radius within reach of a clump;
//   k_sub_ring     one workgroup per row: the ring average
//                  of the projected clump density and its mean clump-ray
//                  cosine (subhalos.py:287-330);
//   k_sub_coarse   one workgroup per row: the rare-mark cut, the bulk
//                  cumulants of K and S, the 3-node Gauss rules in K and in
//                  S | K, the conditional shear regression, the rare bins and
//                  the host-mass table of the rebuilt host: the point tables
//                  that the cell emission interpolates (host_cells_moments,
//                  subhalos_exact.py:1655-1811).
#include <cmath>

#include "ref_sub.h"

namespace sgl {

namespace {

// Ludlow+16 (sgl.py:590-606, conc_model 16) with its z <= 7 clamp
struct Ludlow {
  double poly, c0, beta, g1, g2;
};
__device__ inline Ludlow ludlow(double z) {
  const double a = 1.0 + fmin(z, 7.0);
  Ludlow L;
  L.poly = 4.135 - 0.564 * a - 0.210 * a * a + 0.0557 * a * a * a - 0.00348 * a * a * a * a;
  L.c0 = 3.395 * pow(a, -0.215);
  L.beta = 0.307 * pow(a, 0.540);
  L.g1 = 0.628 * pow(a, -0.047);
  L.g2 = 0.317 * pow(a, -0.893);
  return L;
}
__device__ inline double conc_of(const Ludlow& L, double sig0) {
  const double q = kDeltaC / (sig0 * L.poly);
  return L.c0 * pow(q, -L.g1) * pow(1.0 + pow(q, 1.0 / L.beta), -L.beta * (L.g2 - L.g1));
}
__device__ inline double mu_nfw(double y) { return log1p(y) - y / (1.0 + y); }

// the growth in delta_c(z) (sgl_full.Dg): the ODE growth, or Carroll-Press-
// Turner under the legacy growth
__device__ inline double sub_dg(const Cosmo& c, const Derived& d, const double* gd, double z) {
  if (c.growth_legacy != 0.0)
    return cpt_growth(z, c.Om, c.Ok, d.Or, d.OL) / cpt_growth(0.0, c.Om, c.Ok, d.Or, d.OL);
  return growth_dense_at(gd, z);
}

__device__ inline double dvir_of(const Cosmo& c, const Derived& d, double z) {
  const double zp1 = 1.0 + z;
  const double dd = c.Om * zp1 * zp1 * zp1 / e2_of_z(c, d.Or, d.OL, z) - 1.0;
  return 18.0 * kPi * kPi + 82.0 * dd - 39.0 * dd * dd;
}

// SubhaloModel.fs (subhalos.py:672-705): z_f from delta_c(z_f) = delta_c(z) +
// w_f sqrt(sigma^2(M/2) - sigma^2(M)) (bisection to the last bit; the
// reference's Brent root to 1e-15), N_tau on 200 midpoints, f_s in (0, 0.95)
__device__ double sub_fs(const Cosmo& c, const Derived& d, const double* spl, const double* gd,
                         double M, double z, double dc0, double wf) {
  const double dcz = dc0 / sub_dg(c, d, gd, z);
  const double s1 = sigma_spl(spl, log(0.5 * M)), s0 = sigma_spl(spl, log(M));
  const double ds2 = s1 * s1 - s0 * s0;
  if (!(ds2 > 0.0)) return 0.0;
  const double rhs = dcz + wf * sqrt(ds2);
  if (dc0 / sub_dg(c, d, gd, 30.0) < rhs) return 0.0;
  double lo = z, hi = 30.0;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (dc0 / sub_dg(c, d, gd, mid) < rhs) lo = mid; else hi = mid;
  }
  const double zf = 0.5 * (lo + hi);
  double acc = 0.0;
  for (int q = 0; q < 200; ++q) {
    const double zz = z + (q + 0.5) * (zf - z) / 200.0;
    acc += 6.006 * sqrt(dvir_of(c, d, zz) / 178.0) / (1.0 + zz);
  }
  const double Ntau = acc * (zf - z) / 200.0;
  if (!(Ntau > 0.0)) return 0.0;
  const double f = 0.3563 / pow(Ntau, 0.6) - 0.075;
  return (f > 0.0 && f < 0.95) ? f : 0.0;
}

__device__ inline double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = v;
  __syncthreads();
  for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ inline double block_max(double v, double* red) {
  const int t = threadIdx.x;
  __syncthreads();
  red[t] = v;
  __syncthreads();
  for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

}  // namespace

// ------------------------------------------------------------------ hosts --
__global__ __launch_bounds__(256) void k_sub_hosts(const Cosmo* __restrict__ cosmo,
                                                   const Derived* __restrict__ der,
                                                   const double* __restrict__ slices,
                                                   const double* __restrict__ mass,
                                                   const double* __restrict__ rows,
                                                   const double* __restrict__ spl,
                                                   const double* __restrict__ gd, double gden,
                                                   double wf, double* __restrict__ sub) {
  __shared__ double red[256];
  __shared__ double shv[16];
  __shared__ double sD[kSubNpsi], sRsc[kSubNpsi], sKsc[kSubNpsi], sCc[kSubNpsi];
  __shared__ double sdc[kSubNd];
  __shared__ double sn3[kSubNR];
  const int j = blockIdx.x, i = blockIdx.y, t = threadIdx.x;
  const int row = i * kRefMasses + j;
  double* H = sub + (size_t)row * kHostStride;
  const Cosmo c = cosmo[0];
  if (i >= (int)c.n_z) {
    if (t == 0) { H[kHrValid] = 0.0; H[kHrNq] = 0.0; }
    return;
  }
  const Derived d = der[0];
  const double* S = slices + (size_t)i * kSliceStride;
  const double* R = rows + (size_t)row * kRowStride;
  const double M = mass[(size_t)j * kMassStride];
  const double z = S[kSlZ];
  const double dc0 = 0.6 * pow(1.5 * kPi, 2.0 / 3.0);
  if (t == 0) {
    double f = (M >= kSubMhostMin) ? sub_fs(c, d, spl, gd, M, z, dc0, wf) : 0.0;
    double valid = 0.0;
    if (f > 0.0) {
      const double C = R[kRowXt];           // x_t = c (edge r200)
      const double Dv = dvir_of(c, d, z);
      const double muC = mu_nfw(C);
      const double eta = eta_rvir(C, Dv, muC, cbrt(200.0 / Dv));
      const double Mvir = M * mu_nfw(eta * C) / muC;
      const double psi_min = kSubMFloor / Mvir;
      if (psi_min < 1.0) {
        valid = 1.0;
        shv[1] = C;
        shv[2] = eta;
        shv[3] = Mvir;
        shv[4] = log(psi_min);
        const double s = (1.0 + kSubAlpha) / kSubOmega;
        shv[5] = kSubOmega * pow(kSubBeta, s) / gden * f;
      }
    }
    shv[0] = valid;
    shv[6] = f;
  }
  __syncthreads();
  if (shv[0] == 0.0) {
    if (t == 0) { H[kHrValid] = 0.0; H[kHrNq] = 0.0; }
    return;
  }
  const double C = shv[1], eta = shv[2], Mvir = shv[3], lpm = shv[4], g = shv[5];
  const double rs = R[kRowRs], ks = R[kRowKs];
  const double E2 = S[kSlE2], rhoc = d.rhoc0 * E2, scr = S[kSlScr];
  const double r200 = rs * C;
  const double x0 = kSubX0 * eta;
  // clump mass bins and the clumps' NFW parameters (M200 = m, truncated at r200)
  if (t < kSubNpsi) {
    const double step = (0.0 - lpm) / kSubNpsi;
    const double le0 = t * step + lpm;
    const double le1 = (t + 1 == kSubNpsi) ? 0.0 : (t + 1) * step + lpm;
    const double psi = exp(0.5 * (le0 + le1));
    const double dN = g * pow(psi, kSubAlpha) * exp(-kSubBeta * pow(psi, kSubOmega)) * (le1 - le0);
    const double m = psi * Mvir;
    const Ludlow L = ludlow(z);
    const double Cc = conc_of(L, sigma_spl(spl, log(m)));
    const double r2c = cbrt(3.0 * m / (4.0 * kPi * 200.0 * rhoc));
    const double rsc = r2c / Cc;
    const double rhos = (200.0 / 3.0) * rhoc * Cc * Cc * Cc / mu_nfw(Cc);
    const double ksc = rhos * rsc / scr;
    H[kHrDN + t] = dN;                     // scaled by the r200 fraction below
    H[kHrMc + t] = m;                      // v_r = 1 for clumps in r200
    H[kHrRsc + t] = rsc;
    H[kHrKsc + t] = ksc;
    H[kHrCc + t] = Cc;
    H[kHrD + t] = Cc * rsc;
    sD[t] = Cc * rsc;
    sRsc[t] = rsc;
    sKsc[t] = ksc;
    sCc[t] = Cc;
  }
  // fraction of the clump profile x B(x) / (1 + c x)^2 inside r200 (_frac_inside:
  // 4096-point midpoint rules on [0, 1] and [0, eta])
  double pa = 0.0, pb = 0.0;
  for (int q = t; q < 4096; q += 256) {
    const double xa = (q + 0.5) * (1.0 / 4096), xb = (q + 0.5) * (eta / 4096);
    const double Ba = 1.0 / sqrt(pow(xa / x0, -kSubBiasExp) + 1.0);
    const double Bb = 1.0 / sqrt(pow(xb / x0, -kSubBiasExp) + 1.0);
    pa += xa * Ba / ((1.0 + C * xa) * (1.0 + C * xa));
    pb += xb * Bb / ((1.0 + C * xb) * (1.0 + C * xb));
  }
  const double ma = block_sum(pa, red) * (1.0 / 4096), mb = block_sum(pb, red) * (eta / 4096);
  const double frac = ma / mb;
  // projected clump density on 256 radii (_n3_proj_nb: 160-node trapezoid)
  {
    const double Rr = (t + 0.5) / kSubNR;
    const double um = sqrt(fmax(1.0 - Rr * Rr, 0.0));
    const double du = 1.0 / (kSubNu - 1);
    double acc = 0.0, prev = 0.0;
    for (int u = 0; u < kSubNu; ++u) {
      const double uu = u * du * um;
      const double X = sqrt(Rr * Rr + uu * uu);
      const double q = X / x0;
      const double B = 1.0 / sqrt(1.0 / (q * q * sqrt(q)) + 1.0);
      const double dd = 1.0 + C * X;
      const double pj = B / (X * dd * dd);
      if (u > 0) acc += 0.5 * (pj + prev) * du;
      prev = pj;
    }
    sn3[t] = acc * um * 2.0;
  }
  __syncthreads();
  double tr = 0.0;
  if (t < kSubNR - 1) {
    const double R0 = (t + 0.5) / kSubNR, R1 = (t + 1.5) / kSubNR;
    tr = (R1 - R0) * 0.5 * (sn3[t] * 2.0 * kPi * R0 + sn3[t + 1] * 2.0 * kPi * R1);
  }
  const double norm = block_sum(tr, red);
  H[kHrN2 + t] = sn3[t] / norm / (r200 * r200);
  if (t < kSubNpsi) H[kHrDN + t] *= frac;
  // ray-distance grid (_dgrid): [0, geomspace(1e-5 Dmax, Dmax, 90)]
  double Dm = 0.0;
  for (int b = 0; b < kSubNpsi; ++b) Dm = fmax(Dm, sD[b]);
  if (t < kSubNd) {
    const double e0 = (t == 0) ? 0.0 : geom_np(1e-5 * Dm, Dm, kSubNd, t - 1);
    const double e1 = geom_np(1e-5 * Dm, Dm, kSubNd, t);
    const double dcv = (e0 > 0.0) ? sqrt(e0 * e1) : 0.5 * e1;
    sdc[t] = dcv;
    H[kHrDc + t] = dcv;
    H[kHrDA + t] = kPi * (e1 * e1 - e0 * e0);
  }
  __syncthreads();
  // clump cells (mass bin b, distance cell dd) within the clump's reach
  for (int cidx = t; cidx < kSubCells; cidx += 256) {
    const int b = cidx / kSubNd, dd = cidx - b * kSubNd;
    double kap = 0.0, gam = 0.0;
    if (sdc[dd] < sD[b]) trunc_kg(sdc[dd] / sRsc[b], sKsc[b], sCc[b], kap, gam);
    H[kHrCk + cidx] = kap;
    H[kHrCg + cidx] = gam;
  }
  if (t == 0) {
    H[kHrValid] = 1.0;
    H[kHrF] = shv[6];
    H[kHrEta] = eta;
    H[kHrMvir] = Mvir;
    H[kHrDmax] = Dm;
    H[kHrXsub] = (C * rs + Dm) / rs * 1.001;
    H[kHrRs] = rs;
    H[kHrM] = M;
    H[kHrXref] = 0.0;
    H[kHrNq] = 0.0;
    H[kHrXt] = C;
    H[kHrR200] = r200;
    H[kHrC] = C;
    H[kHrZ] = z;
    H[kHrKs] = ks;
    H[kHrXin] = 1e-6 * C;
    H[kHrG] = g;
  }
}

// ------------------------------------------------------------------- ring --
// one workgroup per row, threads over the (coarse radius, distance cell)
// pairs; the 48 angle cosines and the density table in LDS
__global__ __launch_bounds__(256) void k_sub_ring(const double* __restrict__ host,
                                                  double* __restrict__ coarse,
                                                  double* __restrict__ ring) {
  __shared__ double cth[kSubNth];
  __shared__ double fp[kSubNR];
  __shared__ double dcs[kSubNd];
  __shared__ double rr[kSubNc];
  const int row = blockIdx.x, t = threadIdx.x;
  const double* H = host + (size_t)row * kHostStride;
  if (H[kHrValid] == 0.0 || !(H[kHrXref] > 0.0)) return;
  const double xmin = 0.5 * H[kHrXin], xmax = H[kHrXref], rs = H[kHrRs], r200 = H[kHrR200];
  if (t < kSubNth) cth[t] = cos((t + 0.5) * kPi / kSubNth);
  fp[t] = H[kHrN2 + t];
  if (t < kSubNd) dcs[t] = H[kHrDc + t];
  if (t < kSubNc) {
    const double xg = geom_np(xmin, xmax, kSubNc, t);
    rr[t] = xg * rs;
    coarse[(size_t)row * kCoarseStride + kCmLx + t] = log(xg);
  }
  __syncthreads();
  const double xp0 = (0.5 / kSubNR) * r200, xpl = ((kSubNR - 0.5) / kSubNR) * r200;
  const double inv_dx = 1.0 / ((xpl - xp0) / (kSubNR - 1));
  double* RG = ring + (size_t)row * kRingStride;
  for (int e = t; e < kSubNc * kSubNd; e += 256) {
    const int r = e / kSubNd, dd = e - r * kSubNd;
    const double rv = rr[r], dc = dcs[dd];
    const double cc = rv * rv + dc * dc, trd = 2.0 * rv * dc;
    double s = 0.0, sc = 0.0;
    for (int k = 0; k < kSubNth; ++k) {
      const double arg = cc - trd * cth[k];
      const double x = (arg > 0.0) ? sqrt(arg) : 0.0;
      double v;
      if (x <= xp0) {
        v = fp[0];
      } else if (x >= xpl) {
        v = (x == xpl) ? fp[kSubNR - 1] : 0.0;
      } else {
        int jj = (int)((x - xp0) * inv_dx);
        jj = min(max(jj, 0), kSubNR - 2);
        const double xj = ((jj + 0.5) / kSubNR) * r200, xj1 = ((jj + 1.5) / kSubNR) * r200;
        const double sl = (fp[jj + 1] - fp[jj]) / (xj1 - xj);
        v = fma(sl, x - xj, fp[jj]);
      }
      s += v;
      sc += v * cth[k];
    }
    RG[e] = s * (1.0 / kSubNth);
    RG[kSubNc * kSubNd + e] = (s > 0.0) ? sc / s : 0.0;
  }
}

// ----------------------------------------------------------------- coarse --
namespace {

// _std_moments_from_cumulants (subhalos_exact.py:571-583)
__device__ inline void std_tau(double k2, double k3, double k4, double k5, double& sd, double& t3,
                               double& t4, double& t5) {
  sd = sqrt(fmax(k2, 0.0));
  const double ss = (sd > 0.0) ? sd : 1.0;
  const double g3 = k3 / (ss * ss * ss), g4 = k4 / (ss * ss * ss * ss),
               g5 = k5 / (ss * ss * ss * ss * ss);
  t3 = g3;
  t4 = g4 + 3.0;
  t5 = g5 + 10.0 * g3;
}

// _gauss3_tau_kernel (subhalos_exact.py:587-660) with no upper bound: the
// 3-node rule (nodes clamped to >= lo within tolerance), else the 2-node rule
// in the 3-node layout, else one node at 0
__device__ inline void gauss3(double t3, double t4, double t5, double lo, double* t, double* p) {
  t[0] = t[1] = t[2] = 0.0;
  p[0] = 1.0;
  p[1] = p[2] = 0.0;
  const double D = t4 - t3 * t3 - 1.0;
  bool ok = false;
  if (fabs(D) > 1e-10 && !isnan(D)) {
    const double al = (t3 * t4 + t3 - t5) / D;
    const double be = -t4 - t3 * al;
    const double ga = -t3 - al;
    const double pp = be - al * al / 3.0;
    const double qq = 2.0 * al * al * al / 27.0 - al * be / 3.0 + ga;
    if (pp < 0.0) {
      const double rp = sqrt(-pp / 3.0);
      double arg = 3.0 * qq / (2.0 * pp * rp);
      if (fabs(arg) <= 1.0 + 1e-12) {
        arg = fmin(fmax(arg, -1.0), 1.0);
        const double ph = acos(arg) / 3.0, al3 = al / 3.0;
        double r0 = 2.0 * rp * cos(ph) - al3;
        double r1 = 2.0 * rp * cos(ph - 2.0 * kPi / 3.0) - al3;
        double r2 = 2.0 * rp * cos(ph - 4.0 * kPi / 3.0) - al3;
        double tmp;
        if (r0 > r1) { tmp = r0; r0 = r1; r1 = tmp; }
        if (r1 > r2) { tmp = r1; r1 = r2; r2 = tmp; }
        if (r0 > r1) { tmp = r0; r0 = r1; r1 = tmp; }
        if (r1 - r0 > 1e-7 && r2 - r1 > 1e-7) {
          const double w0 = (1.0 + r1 * r2) / ((r0 - r1) * (r0 - r2));
          const double w1 = (1.0 + r0 * r2) / ((r1 - r0) * (r1 - r2));
          const double w2 = (1.0 + r0 * r1) / ((r2 - r0) * (r2 - r1));
          const double tt = kSubG3TolT * (1.0 + fmax(fabs(r0), fabs(r2)));
          if (w0 > -kSubG3TolW && w1 > -kSubG3TolW && w2 > -kSubG3TolW && r0 >= lo - tt) {
            t[0] = fmax(r0, lo);
            t[1] = fmax(r1, lo);
            t[2] = fmax(r2, lo);
            p[0] = fmax(w0, 0.0);
            p[1] = fmax(w1, 0.0);
            p[2] = fmax(w2, 0.0);
            ok = true;
          }
        }
      }
    }
  }
  if (!ok) {
    const double gg = t3, rt = sqrt(gg * gg + 4.0);
    const double a1 = 0.5 * (gg + rt), a2 = 0.5 * (gg - rt);
    const double tt = kSubG3TolT * (1.0 + fmax(fabs(a1), fabs(a2)));
    if (a2 >= lo - tt) {
      const double t1c = fmax(a1, lo);
      t[0] = fmax(a2, lo);
      t[1] = t1c;
      t[2] = t1c;
      p[0] = a1 / (a1 - a2);
      p[1] = -a2 / (a1 - a2);
      p[2] = 0.0;
    }
  }
}

}  // namespace

constexpr int kSortN = 4096;
static_assert(kSubCells <= kSortN, "the clump cells fit the sort");
static_assert(kSubNd * 16 + kSubNc * 16 + 6 * kSubNd * kSubNbr <= kSortN + kSortN / 2,
              "LDS buffer");

__global__ __launch_bounds__(256) void k_sub_coarse(const Cosmo* __restrict__ cosmo,
                                                    const Derived* __restrict__ der,
                                                    const double* __restrict__ slices,
                                                    const double* __restrict__ spl,
                                                    const double* __restrict__ sigtab,
                                                    double* __restrict__ host,
                                                    double* __restrict__ coarse,
                                                    const double* __restrict__ ring) {
  // LDS: the sort of the clump cells by kappa, later the bulk sums Q[90][16]
  // and the per-radius sums C[128][16]
  __shared__ double big[kSortN + kSortN / 2];
  __shared__ double colmax[kSubNd], colmean[kSubNd], sdN[kSubNpsi], smc[kSubNpsi], sdA[kSubNd];
  __shared__ double red[256];
  __shared__ double scal[8];
  __shared__ double sWk[kSubNc], sCross[kSubNc];
  __shared__ double sdc[kSubNd], sDb[kSubNpsi];
  __shared__ unsigned char cls[kSubCells];    // 0 no cell, 1 bulk, 2 rare
  __shared__ int cnt[256];
  const int row = blockIdx.x, t = threadIdx.x;
  double* H = host + (size_t)row * kHostStride;
  if (H[kHrValid] == 0.0 || !(H[kHrXref] > 0.0)) return;
  double* Cm = coarse + (size_t)row * kCoarseStride;
  const double* RG = ring + (size_t)row * kRingStride;
  const double* EC = RG + kSubNc * kSubNd;
  const double* ck = H + kHrCk;
  const double* cg = H + kHrCg;
  if (t < kSubNd) {
    double mx = 0.0, sm = 0.0;
    for (int r = 0; r < kSubNc; ++r) {
      const double v = RG[r * kSubNd + t];
      mx = fmax(mx, v);
      sm += v;
    }
    colmax[t] = mx;
    colmean[t] = sm / kSubNc;
    sdA[t] = H[kHrDA + t];
    sdc[t] = H[kHrDc + t];
  }
  if (t < kSubNpsi) {
    sdN[t] = H[kHrDN + t];
    smc[t] = H[kHrMc + t];
    sDb[t] = H[kHrD + t];
  }
  __syncthreads();
  // a cell (b, dd) exists where dc < D_b; rate nu = dN_b dA_dd max_r ring
  auto valid = [&](int cidx) {
    const int b = cidx / kSubNd, dd = cidx - b * kSubNd;
    return sdc[dd] < sDb[b];
  };
  double* skey = big;
  int* sidx = reinterpret_cast<int*>(big + kSortN);
  for (int s = t; s < kSortN; s += 256) {
    const bool v = (s < kSubCells) && valid(s);
    skey[s] = v ? ck[s] : -__builtin_huge_val();
    sidx[s] = s;
  }
  __syncthreads();
  // bitonic sort, descending kappa
  for (int k = 2; k <= kSortN; k <<= 1)
    for (int jj = k >> 1; jj > 0; jj >>= 1) {
      for (int s = t; s < kSortN; s += 256) {
        const int l = s ^ jj;
        if (l > s) {
          const double a = skey[s], b = skey[l];
          const bool desc = (s & k) == 0;
          if ((a < b) == desc) {
            skey[s] = b;
            skey[l] = a;
            const int ti = sidx[s];
            sidx[s] = sidx[l];
            sidx[l] = ti;
          }
        }
      }
      __syncthreads();
    }
  // rare-mark cut (_cut_for): the first mark (from the top) whose cumulative
  // rate exceeds eps_rare sets kappa*; none: kappa* = 0.  The cumulative sum
  // in sorted order as a block scan over runs of 16 marks per thread.
  {
    constexpr int kPerT = kSortN / 256;
    double loc[kPerT];
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kPerT; ++u) {
      const int sp = t * kPerT + u;
      double nu = 0.0;
      if (skey[sp] > -__builtin_huge_val()) {
        const int cidx = sidx[sp], b = cidx / kSubNd, dd = cidx - b * kSubNd;
        nu = sdN[b] * sdA[dd] * colmax[dd];
      }
      loc[u] = nu;
      acc += nu;
    }
    red[t] = acc;
    __syncthreads();
    if (t == 0) {
      double run = 0.0;
      for (int q = 0; q < 256; ++q) { const double v = red[q]; red[q] = run; run += v; }
    }
    __syncthreads();
    double cr = red[t];
    int first = kSortN;
#pragma unroll
    for (int u = 0; u < kPerT; ++u) {
      cr += loc[u];
      if (first == kSortN && loc[u] > 0.0 && cr > kSubEpsRare) first = t * kPerT + u;
    }
    __syncthreads();
    cnt[t] = first;
    __syncthreads();
    if (t == 0) {
      int f = kSortN;
      for (int q = 0; q < 256; ++q) f = min(f, cnt[q]);
      scal[0] = (f < kSortN) ? skey[f] : 0.0;
    }
    __syncthreads();
  }
  const double kstar = scal[0];
  // classify the cells; mass scale, rare count and range
  double pms = 0.0, nr = 0.0, kmn = __builtin_huge_val(), kmx = 0.0;
  for (int cidx = t; cidx < kSubCells; cidx += 256) {
    unsigned char c = 0;
    if (valid(cidx)) {
      const int b = cidx / kSubNd, dd = cidx - b * kSubNd;
      pms += sdN[b] * sdA[dd] * colmean[dd] * smc[b];
      const double k = ck[cidx];
      c = (k <= kstar) ? 1 : 2;
      if (c == 2) {
        nr += 1.0;
        kmn = fmin(kmn, k);
        kmx = fmax(kmx, k);
      }
    }
    cls[cidx] = c;
  }
  const double ms = fmax(block_sum(pms, red), 1e-300);
  const double nrare = block_sum(nr, red);
  const double kr_min = -block_max(-kmn, red), kr_max = block_max(kmx, red);
  const int nbr = (nrare > 0.0) ? kSubNbr : 0;
  const int nq = 9 + 2 * nbr;
  // bulk sums per distance cell, Q[dd][16] (the sort buffer is free now)
  double* Q = big;
  double* Cr = big + kSubNd * 16;           // [kSubNc][16]
  for (int e = t; e < kSubNd * 16; e += 256) {
    const int dd = e / 16, col = e - dd * 16;
    double acc = 0.0;
    for (int b = 0; b < kSubNpsi; ++b) {
      const int cidx = b * kSubNd + dd;
      if (cls[cidx] != 1) continue;
      const double k = ck[cidx];
      const double p = sdN[b] * sdA[dd], g = cg[cidx], mh = smc[b] / ms;
      double v;
      switch (col) {
        case 0: v = p * k; break;
        case 1: v = p * k * k; break;
        case 2: v = p * k * k * k; break;
        case 3: v = p * k * k * k * k; break;
        case 4: v = p * k * k * k * k * k; break;
        case 5: v = p * mh; break;
        case 6: v = p * mh * mh; break;
        case 7: v = p * mh * mh * mh; break;
        case 8: v = p * mh * mh * mh * mh; break;
        case 9: v = p * mh * mh * mh * mh * mh; break;
        case 10: v = p * k * mh; break;
        case 11: v = p * g * g; break;
        case 12: v = p * k * g * g; break;
        case 13: v = p * g; break;
        case 14: v = p * k * g; break;
        default: v = p * mh * g; break;
      }
      acc += v;
    }
    Q[e] = acc;
  }
  __syncthreads();
  if (t < kSubNc) {
    double a[16];
#pragma unroll
    for (int col = 0; col < 16; ++col) a[col] = 0.0;
    for (int dd = 0; dd < kSubNd; ++dd) {
      const double rv = RG[t * kSubNd + dd], re = rv * EC[t * kSubNd + dd];
#pragma unroll
      for (int col = 0; col < 13; ++col) a[col] += rv * Q[dd * 16 + col];
#pragma unroll
      for (int col = 13; col < 16; ++col) a[col] += re * Q[dd * 16 + col];
    }
#pragma unroll
    for (int col = 0; col < 16; ++col) Cr[t * 16 + col] = a[col];
  }
  // the (K, S) determinant without cancellation (_det_KS): one wave per
  // radius, lanes over the clump mass bins
  {
    const int lane = t & 63, wv = t >> 6;
    for (int r = wv; r < kSubNc; r += 4) {
      double n_ = 0.0, a_ = 0.0;
      if (lane < kSubNpsi) {
        for (int dd = 0; dd < kSubNd; ++dd) {
          const int cidx = lane * kSubNd + dd;
          if (cls[cidx] != 1) continue;
          const double v = RG[r * kSubNd + dd] * sdN[lane] * sdA[dd];
          n_ += v;
          a_ += v * ck[cidx];
        }
      }
      const double ab = (n_ > 0.0) ? a_ / n_ : 0.0;
      double w = 0.0;
      if (lane < kSubNpsi) {
        for (int dd = 0; dd < kSubNd; ++dd) {
          const int cidx = lane * kSubNd + dd;
          if (cls[cidx] != 1) continue;
          const double dk = ck[cidx] - ab;
          w += RG[r * kSubNd + dd] * sdN[lane] * sdA[dd] * dk * dk;
        }
      }
      const double mb = (lane < kSubNpsi) ? smc[lane] / ms : 0.0;
      double cr = 0.0;
      for (int b2 = 1; b2 < kSubNpsi; ++b2) {
        const double n2 = __shfl(n_, b2), ab2 = __shfl(ab, b2), m2 = __shfl(mb, b2);
        if (lane < b2 && n_ > 0.0 && n2 > 0.0) {
          const double x = m2 * ab - mb * ab2;
          cr += n_ * n2 * x * x;
        }
      }
      w = wave_sum(w);
      cr = wave_sum(cr);
      if (lane == 0) {
        sWk[r] = w;
        sCross[r] = cr;
      }
    }
  }
  __syncthreads();
  // rare bins: geomspace(kmin (1 - 1e-12), kmax (1 + 1e-12), 4)
  double be[kSubNbr + 1];
  if (nbr) {
#pragma unroll
    for (int q = 0; q <= kSubNbr; ++q)
      be[q] = geom_np(kr_min * (1 - 1e-12), kr_max * (1 + 1e-12), kSubNbr + 1, q);
  }
  auto rare_bin = [&](double k) {
    int bi = 0;
    while (bi < kSubNbr && be[bi + 1] <= k) ++bi;   // searchsorted(be, k, "right") - 1
    return min(max(bi, 0), kSubNbr - 1);
  };
  // per (distance cell, rare bin): rate prefactor sum W, mean kappa mu, the
  // two-pass spread V = sum p (k - mu)^2, and the sums of p m, p g^2, p g.
  // Every rate of a distance cell shares its ring factor, so the per-radius
  // sums are ring-weighted sums of these, and the within-bin kappa variance
  // is sum_d ring (V_d + W_d (mu_d - kbar)^2): both terms >= 0, exact for a
  // one-cell bin (the two-pass form of _rare_bins_var)
  double* ra = big + kSubNd * 16 + kSubNc * 16;       // [kSubNd][kSubNbr][6]
  if (nbr) {
    for (int e = t; e < kSubNd * kSubNbr; e += 256) {
      const int dd = e / kSubNbr, bi = e - dd * kSubNbr;
      double W0 = 0.0, W1 = 0.0, M1 = 0.0, G2 = 0.0, G1 = 0.0;
      for (int b = 0; b < kSubNpsi; ++b) {
        const int cidx = b * kSubNd + dd;
        if (cls[cidx] != 2) continue;
        const double k = ck[cidx];
        if (rare_bin(k) != bi) continue;
        const double p = sdN[b] * sdA[dd], g = cg[cidx];
        W0 += p;
        W1 += p * k;
        M1 += p * smc[b];
        G2 += p * g * g;
        G1 += p * g;
      }
      const double mu = (W0 > 0.0) ? W1 / W0 : 0.0;
      double V = 0.0;
      for (int b = 0; b < kSubNpsi; ++b) {
        const int cidx = b * kSubNd + dd;
        if (cls[cidx] != 2) continue;
        const double k = ck[cidx];
        if (rare_bin(k) != bi) continue;
        const double dk = k - mu;
        V += sdN[b] * sdA[dd] * dk * dk;
      }
      double* o = ra + (size_t)e * 6;
      o[0] = W0;
      o[1] = mu;
      o[2] = V;
      o[3] = M1;
      o[4] = G2;
      o[5] = G1;
    }
  }
  __syncthreads();
  double pmax = 0.0;
  if (t < kSubNc) {
    const int r = t;
    const double* Cv = Cr + r * 16;
    const double c1 = Cv[0], c2 = Cv[1], c3 = Cv[2], c4 = Cv[3], c5 = Cv[4];
    const double s1 = Cv[5], s2 = Cv[6], s3 = Cv[7], s4 = Cv[8], s5 = Cv[9];
    const double c11 = Cv[10], E2b = Cv[11], E1 = Cv[13], KE1 = Cv[14], SE1 = Cv[15];
    double nub[kSubNbr] = {}, k1b[kSubNbr] = {}, m1b[kSubNbr] = {}, g2b[kSubNbr] = {},
           e1b[kSubNbr] = {}, v2b[kSubNbr] = {};
    if (nbr) {
      for (int dd = 0; dd < kSubNd; ++dd) {
        const double rv = RG[r * kSubNd + dd], re = rv * EC[r * kSubNd + dd];
#pragma unroll
        for (int bi = 0; bi < kSubNbr; ++bi) {
          const double* o = ra + (dd * kSubNbr + bi) * 6;
          nub[bi] += rv * o[0];
          k1b[bi] += rv * (o[0] * o[1]);
          m1b[bi] += rv * o[3];
          g2b[bi] += rv * o[4];
          e1b[bi] += re * o[5];
        }
      }
      double kbg[kSubNbr];
#pragma unroll
      for (int q = 0; q < kSubNbr; ++q) kbg[q] = k1b[q] / ((nub[q] > 0.0) ? nub[q] : 1.0);
      for (int dd = 0; dd < kSubNd; ++dd) {
        const double rv = RG[r * kSubNd + dd];
#pragma unroll
        for (int bi = 0; bi < kSubNbr; ++bi) {
          const double* o = ra + (dd * kSubNbr + bi) * 6;
          const double dm = o[1] - kbg[bi];
          v2b[bi] += rv * (o[2] + o[0] * dm * dm);
        }
      }
    }
    // K rule, S | K rule (regression on K plus the residual with the marginal
    // shape), conditional shear regression (subhalos_exact.py:1733-1776)
    double sdK, t3, t4, t5, tK[3], pK[3];
    std_tau(c2, c3, c4, c5, sdK, t3, t4, t5);
    gauss3(t3, t4, t5, (0.0 - c1) / ((sdK > 0.0) ? sdK : 1.0), tK, pK);
    double sdS, u3, u4, u5, tS[3], pS[3];
    std_tau(s2, s3, s4, s5, sdS, u3, u4, u5);
    gauss3(u3, u4, u5, (0.0 - s1) / ((sdS > 0.0) ? sdS : 1.0), tS, pS);
    const double c2s = (c2 > 0.0) ? c2 : 1.0;
    const double bK = (c2 > 0.0) ? c11 / c2s : 0.0;
    const double vperp = (c2 > 0.0) ? (sCross[r] + s2 * sWk[r]) / c2s : s2;
    const double sres = sqrt(vperp);
    const double b1 = (c2 > 0.0) ? KE1 / c2s : 0.0;
    const double cov = SE1 - bK * KE1;
    const double den = vperp + kSubKsRidge * s2;
    const double b2 = (den > 0.0) ? cov / den : 0.0;
    const double expl = b1 * b1 * c2 + b2 * b2 * vperp;
    const double G2n = fmax(E2b - expl, 0.0);
    double nur = 0.0;
#pragma unroll
    for (int q = 0; q < kSubNbr; ++q) nur += nub[q];
    const double p0r = nbr ? exp(-nur) : 1.0;
    const double sc = (nbr && nur > 0.0) ? -expm1(-nur) / nur : 0.0;
    double* Kp = Cm + kCmK + r * kSubNq;
    double* Sp = Cm + kCmS + r * kSubNq;
    double* Wp = Cm + kCmW + r * kSubNq;
    double* A1 = Cm + kCmA1 + r * kSubNq;
    double* A2 = Cm + kCmA2 + r * kSubNq;
    for (int a = 0; a < 3; ++a) {
      const double Kn = c1 + sdK * tK[a];
      for (int b = 0; b < 3; ++b) {
        const int q = 3 * a + b;
        const double Sm = fmax(s1 + bK * (Kn - c1) + sres * tS[b], 0.0) * ms;
        const double dK = Kn - c1, dS = Sm / ms - s1;
        const double e1n = E1 + b1 * dK + b2 * (dS - bK * dK);
        Kp[q] = Kn;
        Sp[q] = Sm;
        Wp[q] = pK[a] * pS[b] * p0r;
        A1[q] = e1n;
        A2[q] = G2n + e1n * e1n;
      }
    }
    if (nbr) {
      const double E2 = E2b + E1 * E1;
      for (int q = 0; q < kSubNbr; ++q) {
        const double nz = (nub[q] > 0.0) ? nub[q] : 1.0;
        const double kb = k1b[q] / nz, skb = sqrt(v2b[q] / nz);
        const double mb = m1b[q] / nz, g2 = g2b[q] / nz, e1 = e1b[q] / nz;
        for (int sgn2 = 0; sgn2 < 2; ++sgn2) {
          const int qq = 9 + q + sgn2 * kSubNbr;
          Kp[qq] = c1 + (sgn2 ? kb + skb : kb - skb);
          Sp[qq] = s1 * ms + mb;
          Wp[qq] = 0.5 * nub[q] * sc;
          A1[qq] = E1 + e1;
          A2[qq] = E2 + g2 + 2.0 * E1 * e1;
        }
      }
    }
    for (int q = 0; q < nq; ++q)
      if (Wp[q] > kSubMtWfloor) pmax = fmax(pmax, Sp[q]);
  }
  // host-mass table of the rebuilt host: geomspace(M_lo, M, 17) and M, unique
  const double Smax = block_max(pmax, red);
  const double M = H[kHrM];
  const double Mlo = fmax(M - Smax, kSubMmin);
  const int nmt = (Mlo < M) ? kSubNtab : 1;
  if (t < nmt) {
    const double Mt = (nmt == 1) ? M : geom_np(Mlo, M, kSubNtab, t);
    const Cosmo c = cosmo[0];
    const Derived d = der[0];
    const int i = row / kRefMasses;
    const double* S = slices + (size_t)i * kSliceStride;
    const double z = S[kSlZ], rhoc = d.rhoc0 * S[kSlE2];
    const Ludlow L = ludlow(z);
    const double C = conc_of(L, sigma_spl(spl, log(Mt)));
    const double r2 = cbrt(3.0 * Mt / (4.0 * kPi * 200.0 * rhoc));
    const double rs = r2 / C;
    const double rhos = (200.0 / 3.0) * rhoc * C * C * C / mu_nfw(C);
    const double ksv = rhos * rs / S[kSlScr];
    Cm[kCmLMt + t] = log(Mt);
    Cm[kCmLrs + t] = log(rs);
    Cm[kCmLks + t] = log(ksv);
    double eps = 0.0;
    if (sigtab) {
      const double dc0 = 0.6 * pow(1.5 * kPi, 2.0 / 3.0);
      const double lmstar = ln_mstar(sigtab, dc0 / S[kSlDg]);
      const double s = 0.54 * exp(-0.05 * (log(Mt) - lmstar));
      eps = fmax(0.0, (1.0 - s) / (1.0 + s));
    }
    Cm[kCmEpt + t] = eps;
    (void)c;
  }
  if (t == 0) {
    H[kHrNq] = nq;
    H[kHrNmt] = nmt;
    H[kHrMs] = ms;
    H[kHrKstar] = kstar;
  }
}

}  // namespace sgl

using namespace sgl;

namespace {
// Gamma(s, x) - Gamma(s, 50) for s = (1 + alpha) / omega = 0.045 and
// x = beta psi_res^omega = 5e-15 (the normalisation of the clump mass
// function, subhalos.py:656-659): Gamma(s) minus the series of gamma(s, x)
// (x^s / s to 1e-30 relative at this x), minus the leading asymptotic term of
// Gamma(s, 50) (about 1e-25 of the total).
double sub_gden() {
  const double s = (1.0 + kSubAlpha) / kSubOmega;
  const double x = kSubBeta * std::pow(kSubPsiRes, kSubOmega);
  double term = 1.0, lower = 0.0;
  for (int n = 0; n < 8; ++n) {
    lower += term / (s + n);
    term *= -x / (n + 1);
  }
  lower *= std::pow(x, s);
  const double upper50 = std::pow(kSubBeta, s - 1.0) * std::exp(-kSubBeta) *
                         (1.0 + (s - 1.0) / kSubBeta + (s - 1.0) * (s - 2.0) / (kSubBeta * kSubBeta));
  return std::tgamma(s) - lower - upper50;
}
}  // namespace

namespace sgl {

int sub_hosts_launch(const double* cosmo, const double* derived, const double* slices,
                     const double* mass, const double* rows, const double* spl,
                     const double* gdense, double* host, hipStream_t st) {
  const double wf = std::sqrt(2.0 * std::log(0.815 * std::exp(-0.25) / std::pow(0.5, 0.707) + 1.0));
  hipLaunchKernelGGL(k_sub_hosts, dim3(kRefMasses, kRefMaxShells), dim3(256), 0, st,
                     reinterpret_cast<const Cosmo*>(cosmo), reinterpret_cast<const Derived*>(derived),
                     slices, mass, rows, spl, gdense, sub_gden(), wf, host);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

int sub_coarse_launch(const double* cosmo, const double* derived, const double* slices,
                      const double* spl, const double* sigtab, double* host, double* coarse,
                      double* ring, hipStream_t st) {
  hipLaunchKernelGGL(k_sub_ring, dim3(kSubRows), dim3(256), 0, st, host, coarse, ring);
  hipLaunchKernelGGL(k_sub_coarse, dim3(kSubRows), dim3(256), 0, st,
                     reinterpret_cast<const Cosmo*>(cosmo), reinterpret_cast<const Derived*>(derived),
                     slices, spl, sigtab, host, coarse, ring);
  return hipGetLastError() == hipSuccess ? SGL_OK : SGL_ERR_HIP;
}

}  // namespace sgl
