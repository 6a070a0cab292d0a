// extern "C" entry points of libsgl_amd.so (declared in include/sgl_amd.h).
// Argument checks happen here, on the host, before any launch: a shape or
// count the kernels do not support returns SGL_ERR_INVALID and launches
// nothing.
#include <string.h>

#include <cmath>

#include "sgl_amd_internal.h"

using namespace sgl;

namespace sgl {
int g_sgl_tune[SGL_TUNE_COUNT] = {2, 0, 0, 0};
}

namespace {

bool shape_ok(const sgl_shape* s) {
  if (s == nullptr) return false;
  if (s->n_z_max < 1 || s->n_z_max > kMaxSlices) return false;
  if (s->n_m < 1 || s->n_m > kMaxMasses) return false;
  if (s->n_b < 1 || s->n_b > kMaxB) return false;
  if (s->n_k < 2 || s->n_k > 1 << 20) return false;
  if (s->n_out < 2 || s->n_out > 1 << 16) return false;
  if (!(s->m_lo_h > 0 && s->m_hi_h > s->m_lo_h)) return false;
  if (!(s->u_lo > 0 && s->u_hi > s->u_lo)) return false;
  if (!(s->k_lo > 0 && s->k_hi > s->k_lo)) return false;
  if (!(s->out_hi > s->out_lo)) return false;
  return true;
}

bool batch_ok(int32_t B) { return B >= 1 && B <= (1 << 20); }

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

__global__ void k_copy(const double* __restrict__ src, double* __restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_zero_int(int* p) { *p = 0; }

}  // namespace

// Scratch of sgl_lnphi_exact_noalloc: the per-slice partial sums of the
// exponent (B * n_z_max * n_k complex), at most kScratchDoubles doubles
// (32 MiB).  One copy per GPU, held until the process ends; its device address
// is looked up once per device (hipGetSymbolAddress) and only that address is
// passed to kernels.  The kernels overwrite every slot they later read, so no
// clearing is needed between calls.  Calls that overlap on one GPU race on it.
constexpr size_t kScratchDoubles = (size_t)1 << 22;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchDevices = 64;
static double* g_scratch_ptr[kScratchDevices];

static double* scratch_ptr() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kScratchDevices) return nullptr;
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) return nullptr;
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  return g_scratch_ptr[dev];
}

extern "C" {

const char* sgl_version(void) { return "sgl_amd 0.1 (gfx950)"; }

sgl_status sgl_set_tuning(int32_t knob, int32_t value) {
  static const int kMax[SGL_TUNE_COUNT] = {2, 1, 1, 1};
  if (knob < 0 || knob >= SGL_TUNE_COUNT || value < 0 || value > kMax[knob]) return SGL_ERR_INVALID;
  g_sgl_tune[knob] = value;
  return SGL_OK;
}

int32_t sgl_get_tuning(int32_t knob) {
  if (knob < 0 || knob >= SGL_TUNE_COUNT) return -1;
  return g_sgl_tune[knob];
}

size_t sgl_tables_bytes(const sgl_shape* shape, int32_t B) {
  if (!shape_ok(shape) || !batch_ok(B)) return 0;
  return tables_layout(*shape, B).total * sizeof(double);
}

size_t sgl_hist_bytes(const sgl_shape* shape, int32_t B) {
  if (!shape_ok(shape) || !batch_ok(B)) return 0;
  return hist_layout(*shape, B).total * sizeof(double);
}

size_t sgl_filon_bytes(const sgl_shape* shape, int32_t precision) {
  if (!shape_ok(shape)) return 0;
  size_t n = (size_t)shape->n_out * k_padded_host(*shape);
  if (precision == SGL_PREC_FP32) return n * sizeof(float);
  if (precision == SGL_PREC_BF16X3) return 2 * n * sizeof(uint16_t);
  return 0;
}

size_t sgl_phi_bytes(const sgl_shape* shape, int32_t B, int32_t precision) {
  if (!shape_ok(shape) || !batch_ok(B)) return 0;
  if (precision == SGL_PREC_FP64) return (size_t)B * (shape->n_k + 1) * 2 * sizeof(double);
  size_t n = (size_t)B * k_padded_host(*shape);
  if (precision == SGL_PREC_FP32) return n * sizeof(float);
  if (precision == SGL_PREC_BF16X3) return 2 * n * sizeof(uint16_t);
  return 0;
}

sgl_status sgl_tables_build(const double* cosmo, int32_t B, const sgl_shape* shape, void* tables,
                            int32_t* status, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || cosmo == nullptr || tables == nullptr ||
      status == nullptr)
    return SGL_ERR_INVALID;
  const sgl_shape& sh = *shape;
  hipStream_t st = as_stream(stream);
  TablesLayout L = tables_layout(sh, B);
  double* T = static_cast<double*>(tables);
  size_t nc = (size_t)B * SGL_COSMO_STRIDE;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, st, cosmo,
                     T + L.cosmo, nc);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, status);
  int rc = launch_fill_grids(sh, T + L.uedge, T + L.knodes, T + L.m_nodes, st);
  if (rc != SGL_OK) return (sgl_status)rc;
  rc = tables_build(T + L.cosmo, B, sh, T + L.m_nodes, T + L.derived, T + L.slices, T + L.mass,
                    T + L.rows, T + L.frows, T + L.cov, T + L.sigtab, status, st);
  return (sgl_status)rc;
}

sgl_status sgl_lnphi_exact(const void* tables, int32_t B, const sgl_shape* shape, double* lam,
                           double* amp, double* work, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !lam || !amp || !work)
    return SGL_ERR_INVALID;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_lnphi_exact(T + L.rows, T + L.frows, T + L.cosmo, T + L.uedge,
                                        T + L.knodes, *shape, B, lam, amp, work,
                                        as_stream(stream));
}

sgl_status sgl_lnphi_exact_noalloc(const void* tables, int32_t B, const sgl_shape* shape,
                                   double* lam, double* amp, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !lam || !amp) return SGL_ERR_INVALID;
  if ((size_t)B * shape->n_z_max * shape->n_k * 2 > kScratchDoubles) return SGL_ERR_INVALID;
  double* work = scratch_ptr();
  if (work == nullptr) return SGL_ERR_HIP;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_lnphi_exact(T + L.rows, T + L.frows, T + L.cosmo, T + L.uedge,
                                        T + L.knodes, *shape, B, lam, amp, work,
                                        as_stream(stream));
}

size_t sgl_vp_bytes(const sgl_shape* shape, int32_t B) {
  if (!shape_ok(shape) || !batch_ok(B)) return 0;
  return vp_layout(*shape, B).total * sizeof(double);
}

sgl_status sgl_lnphi_exact_vp(const void* tables, int32_t B, const sgl_shape* shape, double* lam,
                              double* lnr, double* work, void* vpwork, int32_t* status,
                              void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !lam || !lnr || !work || !vpwork || !status)
    return SGL_ERR_INVALID;
  if (!(shape->flags & SGL_FLAG_CLUSTERING)) return SGL_ERR_INVALID;
  const sgl_shape& sh = *shape;
  hipStream_t st = as_stream(stream);
  TablesLayout L = tables_layout(sh, B);
  VpLayout V = vp_layout(sh, B);
  const double* T = static_cast<const double*>(tables);
  double* W = static_cast<double*>(vpwork);
  int rc = launch_vp_cells(T + L.rows, T + L.frows, T + L.cosmo, T + L.uedge, T + L.knodes, sh, B,
                           V, W, work, 0, st);
  if (rc == SGL_OK)
    rc = launch_vp_setup(T + L.rows, T + L.frows, T + L.cov, T + L.cosmo, sh, B, V, W, status, st);
  if (rc == SGL_OK)
    rc = launch_vp_cells(T + L.rows, T + L.frows, T + L.cosmo, T + L.uedge, T + L.knodes, sh, B,
                         V, W, work, 1, st);
  if (rc == SGL_OK) rc = launch_sum_slices(work, sh, B, lam, st);
  if (rc == SGL_OK) rc = launch_vp_chain(T + L.cosmo, sh, B, V, W, lnr, st);
  return (sgl_status)rc;
}

sgl_status sgl_bin_cells(const void* tables, int32_t B, const sgl_shape* shape, void* hist,
                         void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !hist) return SGL_ERR_INVALID;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_bin(T + L.rows, T + L.frows, T + L.cosmo, T + L.uedge, *shape, B,
                                static_cast<double*>(hist), as_stream(stream));
}

sgl_status sgl_bin_pass(const void* tables, int32_t B, const sgl_shape* shape, void* hist,
                        void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !hist) return SGL_ERR_INVALID;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_bin_pass(T + L.rows, T + L.frows, T + L.cosmo, T + L.uedge, *shape,
                                     B, static_cast<double*>(hist), as_stream(stream));
}

sgl_status sgl_bin_reduce(const void* tables, int32_t B, const sgl_shape* shape, void* hist,
                          void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !hist) return SGL_ERR_INVALID;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_bin_reduce(T + L.rows, T + L.cosmo, T + L.uedge, *shape, B,
                                       static_cast<double*>(hist), as_stream(stream));
}

sgl_status sgl_lnphi_binned(const void* tables, const void* hist, int32_t B,
                              const sgl_shape* shape, double* lam, float* amp, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !hist || !lam) return SGL_ERR_INVALID;
  if ((shape->flags & SGL_FLAG_CLUSTERING) && !amp) return SGL_ERR_INVALID;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_lnphi_binned(static_cast<const double*>(hist), T + L.cosmo,
                                         T + L.knodes, *shape, B, lam, amp, as_stream(stream));
}

sgl_status sgl_quadform(const void* tables, const void* amp, int32_t B, const sgl_shape* shape,
                        int32_t precision, double* psi2, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !amp || !psi2) return SGL_ERR_INVALID;
  if (precision != SGL_PREC_FP64 && precision != SGL_PREC_FP32 && precision != SGL_PREC_BF16X3)
    return SGL_ERR_UNSUPPORTED;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_quadform(T + L.cov, T + L.cosmo, amp, B, *shape, precision, psi2,
                                     as_stream(stream));
}

sgl_status sgl_trunc_nfw_kg(const double* x, int32_t n, double ks, double xt, int32_t precision,
                            double* kap, double* gam, void* stream) {
  if (!x || !kap || !gam || n < 1 || !(xt > 0.0) || !std::isfinite(ks)) return SGL_ERR_INVALID;
  if (precision != SGL_PREC_FP64 && precision != SGL_PREC_FP32) return SGL_ERR_UNSUPPORTED;
  return (sgl_status)launch_trunc_nfw(x, n, ks, xt, precision, kap, gam, as_stream(stream));
}

int32_t sgl_psi2_nodes(const sgl_shape* shape) {
  if (!shape_ok(shape)) return 0;
  return psi2_nodes(*shape);
}

sgl_status sgl_psi2_interp(const double* psi2c, int32_t B, const sgl_shape* shape, double* psi2,
                           void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !psi2c || !psi2) return SGL_ERR_INVALID;
  return (sgl_status)launch_psi2_interp(psi2c, *shape, B, psi2, as_stream(stream));
}

sgl_status sgl_phi_pack(const double* lam, const double* psi2, const void* tables, int32_t B,
                        const sgl_shape* shape, int32_t precision, void* phi_op, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !lam || !tables || !phi_op) return SGL_ERR_INVALID;
  if (precision < SGL_PREC_FP64 || precision > SGL_PREC_BF16X3) return SGL_ERR_UNSUPPORTED;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_phi_pack(lam, psi2, T + L.knodes, *shape, B, precision, phi_op,
                                     as_stream(stream));
}

sgl_status sgl_filon_build(const void* tables, int32_t B, const sgl_shape* shape,
                             int32_t precision, void* E_op, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !tables || !E_op) return SGL_ERR_INVALID;
  if (precision != SGL_PREC_FP32 && precision != SGL_PREC_BF16X3) return SGL_ERR_UNSUPPORTED;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_filon(T + L.knodes, *shape, precision, E_op, as_stream(stream));
}

sgl_status sgl_invert_gemm(const void* E_op, const void* phi_op, int32_t B, const sgl_shape* shape,
                           int32_t precision, float* p, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !E_op || !phi_op || !p) return SGL_ERR_INVALID;
  if (precision != SGL_PREC_FP32 && precision != SGL_PREC_BF16X3) return SGL_ERR_UNSUPPORTED;
  return (sgl_status)launch_gemm(E_op, phi_op, *shape, B, precision, p, as_stream(stream));
}

sgl_status sgl_invert_fp64(const double* phi, const void* tables, int32_t B,
                           const sgl_shape* shape, double* p, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !phi || !tables || !p) return SGL_ERR_INVALID;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_invert_fp64(phi, T + L.knodes, *shape, B, nullptr, 0, p,
                                        as_stream(stream));
}

sgl_status sgl_invert_fp64_at(const double* phi, const void* tables, int32_t B,
                              const sgl_shape* shape, const double* t, int32_t n_t, double* p,
                              void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !phi || !tables || !t || !p) return SGL_ERR_INVALID;
  if (n_t < 1 || n_t > (1 << 24)) return SGL_ERR_INVALID;
  TablesLayout L = tables_layout(*shape, B);
  const double* T = static_cast<const double*>(tables);
  return (sgl_status)launch_invert_fp64(phi, T + L.knodes, *shape, B, t, n_t, p,
                                        as_stream(stream));
}

sgl_status sgl_diagnostics(const void* p, int32_t p_is_double, int32_t B, const sgl_shape* shape,
                           double* diag, void* stream) {
  if (!shape_ok(shape) || !batch_ok(B) || !p || !diag) return SGL_ERR_INVALID;
  return (sgl_status)launch_diag(p, p_is_double, *shape, B, diag, as_stream(stream));
}

}  // extern "C"
