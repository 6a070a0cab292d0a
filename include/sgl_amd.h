/*
 * sgl_amd.h -- C ABI of the MI355X magnification-PDF hot path.
 *
 * Plain pointers and sizes only.  Pointers are DEVICE pointers unless the
 * parameter says "host".  Every call is stream-ordered on `stream` (a
 * hipStream_t passed as void*, NULL = the default stream), never synchronises,
 * never allocates: the caller owns every buffer (the *_bytes queries give
 * the sizes).  Every entry point returns an sgl_status; nothing throws across
 * the ABI.
 *
 * The reference (sgl_analytic, pure Python + Numba) has no C ABI; each entry
 * point below names the reference function whose job it takes over, so a
 * ctypes binding can sit where that function is called.
 */
#ifndef SGL_AMD_H
#define SGL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sgl_status {
  SGL_OK = 0,
  SGL_ERR_INVALID = 1,     /* bad argument (shape, count, null pointer) */
  SGL_ERR_HIP = 2,         /* a HIP launch or runtime call failed */
  SGL_ERR_UNSUPPORTED = 3, /* precision / mode not built */
} sgl_status;

/* shape flags */
#define SGL_FLAG_CLUSTERING 1 /* add the two-point term Psi_2 */
#define SGL_FLAG_WINDOW 2     /* cubic B-spline window on phi(k) */
#define SGL_FLAG_FILAMENTS 4  /* add the cylinder-filament population: per (slice, mass)
                                 120 u bins x 13 orientation nodes (the "fil" ingredient) */
#define SGL_FLAG_EDGE_RVIR 8  /* truncate the halos at r_vir (x_t = r_vir / r_s) instead of
                                 r200 (x_t = c): the reference's edge="rvir" */
#define SGL_FLAG_ELL 16       /* halo ellipticity (the "ell" ingredient): eps(M, z) of
                                 Allgood+06 with M*(z), every halo cell split into 12
                                 orientation nodes with the pseudo-elliptical kappa, gamma
                                 (sgl_full.py:551-613, 1659, 2038-2064); rows carry eps in
                                 slot 6, the tables a sigma(M) table for M* */
#define SGL_FLAG_REF_GRID 32  /* the reference's own grids (set by sgl_ref_tables_build):
                                 shells with the engine's 103-node log-z grid below z_s as
                                 edges plus 0 and z_s, one midpoint node each
                                 (sgl_full.py:480-518, zint="midpoint"), and the engine's
                                 mass nodes 1..99 of exp(linspace(ln 1e7, ln 1e17, 100))
                                 Msun (sgl_full.py:471-474, 1864) */
#define SGL_FLAG_SUBHALOS 64  /* reference-shape path only: subhalos on their hosts (the
                                 "sub" ingredient, exact decorated-host sector, moments
                                 method, clumps truncated at their r200; subhalos.py:
                                 540-772, subhalos_exact.py:1531-1931); the tables carry
                                 the reference's sigma(M) spline and a dense growth table */

/* Batched path: slice amplitudes and Psi_2 at every SGL_PSI2_STRIDE-th k node
 * (sub-grid node jc = k node jc * stride of the log grid, the last one up to
 * stride - 1 spacings past k_hi), then cubic Lagrange in ln k onto every
 * node (sgl_psi2_interp).  Grids with n_k - 1 < 3 * SGL_PSI2_STRIDE use
 * every node. */
#define SGL_PSI2_STRIDE 8

/* precision of the MFMA contractions */
#define SGL_PREC_FP64 0   /* plain path: FP64 vector ALU */
#define SGL_PREC_FP32 1   /* FP32 MFMA, FP32 accumulation */
#define SGL_PREC_BF16X3 2 /* split-BF16 (hi+lo) x 3 products on MFMA, FP32 accumulation */

/* Launch-shape knobs (sgl_set_tuning): chosen per device/shape by the
 * project's autotuner; every value gives identical results up to the
 * floating-point order of the histogram adds. */
#define SGL_TUNE_BIN_THREADS 0   /* binning block: 0 = 256, 1 = 512, 2 = 1024 threads */
#define SGL_TUNE_LNPHI_BLOCK 1   /* binned exponent block: 0 = 128, 1 = 256 threads */
#define SGL_TUNE_GEMM_BM 2       /* inversion GEMM tile rows: 0 = 64, 1 = 128 */
#define SGL_TUNE_BIN_SCHEDULE 3  /* binning schedule: 0 = one workgroup per (instance, slice),
                                    1 = persistent workgroups over a work queue of slices */
#define SGL_TUNE_COUNT 4

/* Grids of one solve (host struct, passed by pointer). */
typedef struct sgl_shape {
  int32_t n_z_max; /* slices per instance (max over the batch), <= 96 */
  int32_t n_m;     /* masses, log-uniform bins over [m_lo_h, m_hi_h] Msun/h */
  int32_t n_b;     /* impact-parameter bins, log-uniform over [u_lo, u_hi] r_vir */
  int32_t n_k;     /* log-spaced k nodes over [k_lo, k_hi]; a k = 0 node is added */
  int32_t n_out;   /* output ln(mu) nodes, uniform over [out_lo, out_hi] */
  int32_t flags;   /* SGL_FLAG_* */
  double m_lo_h, m_hi_h, u_lo, u_hi, k_lo, k_hi, out_lo, out_hi;
} sgl_shape;

/* Per-instance input record: 16 doubles
 *   h, Om, sigma8, Ob, ns, zeq, Ok, w0, wa, z_s, n_z, radiation, growth, zint, zint_n, 0
 * (the cosmology arguments of generate_pdf_lnmu, api.py:26-31, plus the
 * source redshift and the slice count of this instance).
 *   radiation (slot SGL_REC_RADIATION): 0 = "engine", Omega_r =
 *     Omega_m / (1 + z_eq) in E(z) and the growth ODE (the pipeline default);
 *     1 = None, no radiation (sgl.py:93-104).
 *   growth (slot SGL_REC_GROWTH): 0 = "ode", the growth ODE for D(z)
 *     everywhere (the pipeline default); 1 = "legacy", LambdaCDM only: D(z)
 *     from the Heath integral without radiation (sgl.py:270-306) in the mass
 *     function, and the Carroll-Press-Turner growth Dg(z) in the halo and
 *     filament bias and in beta (sgl_full.py:521-548).  The caller resolves
 *     "auto" (legacy for LambdaCDM, ode otherwise) and keeps legacy to
 *     w0 = -1, wa = 0.
 *   zint (slot SGL_REC_ZINT, read by sgl_ref_tables_build only): 0 =
 *     "midpoint", the engine nodes below z_s as shell edges plus 0 and z_s,
 *     one node at each shell's middle, n_z = sgl_ref_shell_count(z_s) (the
 *     pipeline default); 1 = "engine", the shells [z_{j-1}, z_j] of the
 *     engine grid with z_j < z_s, j >= 1, one node at the upper edge with
 *     width z_j - z_{j-1}, n_z = sgl_ref_shell_count(z_s) - 2
 *     (sgl_full.py:480-518, 1858-1859); 2 = "gauss", the midpoint shells at
 *     zint_n (slot SGL_REC_ZINT_N) Gauss-Legendre nodes each, shell-major,
 *     n_z = zint_n sgl_ref_shell_count(z_s) <= SGL_REF_MAX_SHELLS.  The
 *     kernel-shape path ignores both. */
#define SGL_COSMO_STRIDE 16
#define SGL_REC_RADIATION 11
#define SGL_REC_GROWTH 12
#define SGL_REC_ZINT 13
#define SGL_REC_ZINT_N 14

/* Buffer sizes in bytes for a batch of B instances. */
size_t sgl_tables_bytes(const sgl_shape* shape, int32_t B);     /* grids+derived+slices+mass+rows+cov */
size_t sgl_hist_bytes(const sgl_shape* shape, int32_t B);       /* FP64 jump histograms */
size_t sgl_filon_bytes(const sgl_shape* shape, int32_t precision); /* inversion operand E */
size_t sgl_phi_bytes(const sgl_shape* shape, int32_t B, int32_t precision); /* operand phi */

/* Library version string (host). */
const char* sgl_version(void);

/* Set / read a launch-shape knob (host-side, process-wide; not for
 * concurrent use from several threads).  No reference counterpart: the
 * reference has no kernels to tune. */
sgl_status sgl_set_tuning(int32_t knob, int32_t value);
int32_t sgl_get_tuning(int32_t knob);

/*
 * Lens-population tables: cosmology (E, chi, D_A, Sigma_cr, growth ODE,
 * EH98 transfer, sigma(M), Sheth-Tormen, Ludlow+16 c(M), r_vir, bias) and
 * per-(slice, mass) rows, plus the slice covariance C_ij when
 * SGL_FLAG_CLUSTERING is set.
 * Replaces: pipeline.make_cosmology (sgl_analytic/pipeline.py:36-72) and the
 * per-(z, M) loop of sgl_full._build_shell_into (sgl_analytic/sgl_full.py:1848-1885),
 * shell_covariance (sgl_full.py:2514-2538).
 * tables: sgl_tables_bytes(shape, B) bytes; status: one int32, OR-ed flags.
 */
sgl_status sgl_tables_build(const double* cosmo, int32_t B, const sgl_shape* shape,
                            void* tables, int32_t* status, void* stream);

/*
 * Exact FP64 compensated exponent over every cell (plain path):
 *   lam[b][n]    = sum_c w_c (e^{i k_n xi_c} - 1 - i k_n xi_c)
 *   amp[b][i][n] = sum_{c in slice i} beta_c w_c (e^{i k_n xi_c} - 1)
 * as interleaved (re, im) doubles, cells generated in-kernel from the rows.
 * work: B * n_z_max * n_k * 2 doubles (per-slice partial sums of lam, added
 * in slice order).
 * Replaces: Lam.__call__ / _sep_raw / _sep_lam_kernel
 * (sgl_full.py:3156-3160, 3040-3086, 2586-2619) on the 1-D jump chain of
 * sgl.Lambda_of_k (sgl.py:961-980).
 */
sgl_status sgl_lnphi_exact(const void* tables, int32_t B, const sgl_shape* shape,
                           double* lam, double* amp, double* work, void* stream);

/*
 * sgl_lnphi_exact without the caller's work buffer: the per-slice partial sums
 * live in the library's own device array of at most 4,194,304 doubles (32 MiB,
 * one per GPU, held until the process ends), so B * n_z_max * n_k * 2 must not
 * exceed that (SGL_ERR_INVALID otherwise).  No allocation of any kind happens
 * in the call.  Calls that overlap on one GPU share the array and race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 */
sgl_status sgl_lnphi_exact_noalloc(const void* tables, int32_t B, const sgl_shape* shape,
                                   double* lam, double* amp, void* stream);

/*
 * Exact path with the positive ("vp") clustering closure in place of Psi_2:
 *   lam[b][n] as sgl_lnphi_exact, and
 *   lnr[b][n] = ln E exp( sum_s sum_g (lambda_sg(U_s) - 1) Y_sg(k_n) ),
 *   Y_sg(k) = sum_{c in slice s, group g} w_c (e^{i k xi_c} - 1),
 * interleaved (re, im) doubles; pass lnr as the psi2 argument of
 * sgl_phi_pack.  Groups g: the halo mass rows of a slice, then (with
 * SGL_FLAG_FILAMENTS) its filament rows.  U_s: a stationary chain on 48
 * Gauss-Hermite nodes with correlation rho_s = C_{s,s+1}/sqrt(C_ss C_{s+1,s+1})
 * between neighbouring slices; lambda_g: rectified rates with E lambda = 1 and
 * Var lambda = eta f_s (beta_g^2 C_ss), f_s = (1 - rbar_s)/(1 + rbar_s); eta
 * fitted so that the chain variance of the k^2 term equals A^T C A,
 * A_s = sum_g beta_g sum_{c in g} w_c xi_c.  Needs SGL_FLAG_CLUSTERING (the
 * slice covariance C of the tables).
 * work: as sgl_lnphi_exact; vpwork: sgl_vp_bytes(shape, B) bytes, its first
 * B doubles hold eta after the call, the next B the target A^T C A; status: OR-ed with 4 (a transition
 * balancing that did not converge) or 8 (a rate variance beyond the 48-node
 * cap, or no bracket for eta).
 * Replaces: pos_closure.install / Closure (calibrate, population)
 * (sgl_analytic/pos_closure.py:453-591, 708-733), the transition and rates
 * (pos_closure.py:251-308), the fused group sums and the chain recursion
 * (pos_closure_fused.py:51-140, pos_closure.py:183-240).
 */
size_t sgl_vp_bytes(const sgl_shape* shape, int32_t B);
sgl_status sgl_lnphi_exact_vp(const void* tables, int32_t B, const sgl_shape* shape, double* lam,
                              double* lnr, double* work, void* vpwork, int32_t* status,
                              void* stream);

/*
 * Convergence and shear kap[i], gam[i] of an NFW profile of amplitude ks
 * truncated at x_t, at n points x[i] (device arrays, x = R / r_s > 0):
 * SGL_PREC_FP64, the exact path's
 * closed forms; SGL_PREC_FP32, the batched path's FP32 form (the one
 * sgl_bin_pass evaluates per cell).  Outside x_t: kap = 0, gam = the
 * monopole shear of the enclosed mass.
 * Replaces: nfw_trunc_kg (sgl_full.py:660-669).
 */
sgl_status sgl_trunc_nfw_kg(const double* x, int32_t n, double ks, double xt, int32_t precision,
                            double* kap, double* gam, void* stream);

/*
 * Batched path, step 1: histogram every cell into a fine signed-log jump
 * lattice (w, w xi, w xi^2 per bin; FP64 sums in LDS per slice) and a
 * per-slice coarse lattice (beta w, beta w xi, beta w xi^2), then reduce
 * the slices in order into compacted Gaussian bin records with their prefix
 * moments.  sgl_bin_cells = sgl_bin_pass (every cell, FP32 lens profile in
 * a form without cancellation at x = 1, FP32 jump log) followed by
 * sgl_bin_reduce (the slice-order reduction); the two halves are exported
 * so that each can be timed on its own.  The histogram sums are 64-bit
 * fixed point (a per-slice power-of-two scale), so the result is the same
 * bits on every call for the same inputs.
 * Replaces: Cells.add / _add_direct_kernel (sgl_full.py:1314-1383, 971-1046).
 */
sgl_status sgl_bin_cells(const void* tables, int32_t B, const sgl_shape* shape, void* hist,
                         void* stream);
sgl_status sgl_bin_pass(const void* tables, int32_t B, const sgl_shape* shape, void* hist,
                        void* stream);
sgl_status sgl_bin_reduce(const void* tables, int32_t B, const sgl_shape* shape, void* hist,
                          void* stream);

/*
 * Batched path, step 2: exponent from the histograms (each bin a Gaussian
 * record: centroid and in-bin variance), FP64 range reduction, FP32
 * sin/cos, FP64 compensated small-phase terms and FP64 accumulation.  lam:
 * FP64 [B][n_k][2]; amp (with SGL_FLAG_CLUSTERING): the slice amplitudes as
 * FP32 [B][n_z_max][n_kc][2] on the two-point sub-grid, n_kc =
 * sgl_psi2_nodes(shape), written for the instance's own slices only (the
 * two-point term consumes them in FP32).
 * Replaces: Lam._sep_raw (sgl_full.py:3040-3086) on binned cells.
 */
sgl_status sgl_lnphi_binned(const void* tables, const void* hist, int32_t B,
                            const sgl_shape* shape, double* lam, float* amp, void* stream);

/*
 * Two-point term Psi_2[b][n] = 1/2 sum_ij C_ij amp_i amp_j (complex square,
 * no conjugate), result FP64.  precision SGL_PREC_FP64: amp is FP64 (as
 * written by sgl_lnphi_exact), FP64 sums, psi2 [B][n_k][2] on every node;
 * SGL_PREC_FP32: amp is FP32 on the sub-grid (as written by
 * sgl_lnphi_binned), U = A C on FP32 MFMA with FP32 accumulation, psi2
 * [B][n_kc][2] on the sub-grid (expand it with sgl_psi2_interp);
 * SGL_PREC_BF16X3: the same with A and C split into BF16 head + tail and
 * U = Ahi Chi + Ahi Clo + Alo Chi on BF16 MFMA, FP32 accumulation.
 * Replaces: Lam.Q (sgl_full.py:3353-3397), nord = 1.
 */
sgl_status sgl_quadform(const void* tables, const void* amp, int32_t B, const sgl_shape* shape,
                        int32_t precision, double* psi2, void* stream);

/* Node count n_kc of the batched two-point sub-grid (host). */
int32_t sgl_psi2_nodes(const sgl_shape* shape);

/*
 * psi2[b][n] (FP64, every k node) from psi2c[b][jc] on the sub-grid by
 * cubic Lagrange in ln k (four nearest sub-grid nodes, one-sided at the
 * ends); exact at the sub-grid nodes.
 * Replaces: Lam._Q_interpolated (sgl_full.py:2953-3022), the reference's
 * spline of Psi_2 from its coarse grid onto the inversion grid.
 */
sgl_status sgl_psi2_interp(const double* psi2c, int32_t B, const sgl_shape* shape, double* psi2,
                           void* stream);

/*
 * phi = window(k) exp(lam + psi2) packed as the inversion operand with the
 * k = 0 node phi = 1: FP64 [B][n_k+1][2], FP32 [B][Kp], or split BF16
 * (hi plane then lo plane, [B][Kp] each); Kp = round_up(2 (n_k+1), 64).
 * psi2 may be NULL (no two-point term).
 */
sgl_status sgl_phi_pack(const double* lam, const double* psi2, const void* tables, int32_t B,
                        const sgl_shape* shape, int32_t precision, void* phi_op, void* stream);

/*
 * Filon inversion operand E[n_out][Kp]: p(t_j) = sum_n E[j][2n] Re phi_n
 * + E[j][2n+1] Im phi_n, phi piecewise linear between the k nodes (FP32, or
 * split BF16 hi plane then lo plane).  tables/B: any tables buffer built
 * with this shape (only its k grid is read).
 * Replaces: the k-sum of sgl.P_of_xi (sgl.py:994-1012) / P_vector_joint
 * (sgl.py:1127-1156).
 */
sgl_status sgl_filon_build(const void* tables, int32_t B, const sgl_shape* shape,
                           int32_t precision, void* E_op, void* stream);

/* p[b][j] (float) = E @ phi on MFMA (FP32 or split-BF16x3). */
sgl_status sgl_invert_gemm(const void* E_op, const void* phi_op, int32_t B, const sgl_shape* shape,
                           int32_t precision, float* p, void* stream);

/* Plain path: p[b][j] (double) with the Filon weights evaluated on the fly
 * in FP64; phi from sgl_phi_pack(..., SGL_PREC_FP64, ...) as [B][n_k+1] complex. */
sgl_status sgl_invert_fp64(const double* phi, const void* tables, int32_t B,
                           const sgl_shape* shape, double* p, void* stream);

/* As sgl_invert_fp64 at n_t caller-given ln(mu) points t (device array):
 * p[b][j] for j < n_t.  The drop-in generate_pdf_lnmu evaluates the user's
 * grid with it (api.py:26-84 accepts any strictly increasing grid). */
sgl_status sgl_invert_fp64_at(const double* phi, const void* tables, int32_t B,
                              const sgl_shape* shape, const double* t, int32_t n_t, double* p,
                              void* stream);

/*
 * FP64 diagnostics of p on the output grid, 8 doubles per instance:
 * grid mass, mean, variance, skewness, sigma_DL/D_L, <mu>, P(mu > 2),
 * negative mass.  p_is_double selects double or float input.
 * Replaces: the moment block of invert (sgl_full.py:362-371) and the
 * diagnostics of generate_pdf_lnmu (api.py:71-79).
 */
sgl_status sgl_diagnostics(const void* p, int32_t p_is_double, int32_t B, const sgl_shape* shape,
                           double* diag, void* stream);

/* ------------------------------------------------------------------------
 * Reference-shape path: the reference's own 2-D (k_kappa, k_gamma) chain on one
 * instance, with the reference's grids (shells of sgl_full.z_shells, the
 * engine mass grid, adaptive impact-parameter cells, the 961 x 300 fine
 * lattice, width-relative k and (kappa, gamma) grids).  The host drives it in
 * the order of pipeline.run (sgl_analytic/pipeline.py:165-241): tables ->
 * critical curves -> subhalo tables (with SGL_FLAG_SUBHALOS) -> cells (count,
 * then emit) -> binning -> reach probes (sgl_ref_lam_axis) ->
 * Phi on the tensor grid -> P3 -> arc integrals and normalisation.  Grid
 * sizes depend on device results (cell count, sigma, reach), so the host reads
 * those scalars between calls.
 */
#define SGL_REF_MAX_SHELLS 208            /* lens nodes: (103 engine nodes below z_s + the last
                                             shell) x 2 Gauss-Legendre nodes */
#define SGL_REF_MASSES 99                 /* engine mass nodes 1..99 */
#define SGL_REF_FINE_BINS (961 * 300)     /* fine (kappa, gamma) lattice, sgl_full.py:456-468 */

/* Shell count of z_s (engine nodes below z_s, plus one); the caller stores it in
 * the record's n_z slot.  Host function. */
int32_t sgl_ref_shell_count(double z_s);
/* (row, orientation) tasks of the cell kernels: 104 x 99 x (12 with SGL_FLAG_ELL, else 1). */
int32_t sgl_ref_tasks(int32_t flags);
size_t sgl_ref_tables_bytes(int32_t flags);    /* flags: SGL_FLAG_ELL | SGL_FLAG_FILAMENTS |
                                                  SGL_FLAG_EDGE_RVIR (x_t = r_vir / r_s) |
                                                  SGL_FLAG_SUBHALOS */
size_t sgl_ref_roots_bytes(void);

/*
 * Tables of one instance on the reference grid (cosmo: one 16-double record,
 * device; status zeroed, then 2 if the record's z_s or shell count is invalid).
 * Replaces: pipeline.make_cosmology (pipeline.py:36-72), z_shells
 * (sgl_full.py:480-518, zint="midpoint") and the per-(z, M) quantities of
 * _build_shell_into (sgl_full.py:1848-1884).
 */
sgl_status sgl_ref_tables_build(const double* cosmo, int32_t flags, void* tables, int32_t* status,
                                void* stream);

/*
 * Critical curves: per (shell, mass) row the radii where detA of the
 * untruncated profile changes sign (45 bisections on sgl._XG) and, with
 * SGL_FLAG_ELL, the roots of every orientation (roots: sgl_ref_roots_bytes).
 * status |= 16 when a row has more roots than its slots.
 * Replaces: _refine_edges / _refine_edges_eps_batch (sgl_full.py:2342-2372,
 * 2279-2339).
 */
sgl_status sgl_ref_roots(const void* tables, int32_t flags, double* roots, int32_t* status,
                         void* stream);

/*
 * Subhalo tables (flags must hold SGL_FLAG_SUBHALOS, and not SGL_FLAG_EDGE_RVIR;
 * sub: sgl_ref_sub_bytes(), after sgl_ref_roots): per host row the clump law
 * (f_s, the 28 clump mass bins, their NFW parameters and reach, the projected
 * clump density, the clump cells' kappa and gamma), the coarse host radii up
 * to the largest host-grid radius within reach of a clump, and on them the
 * decorated-host point tables (K offset, removed mass, weight, shear terms;
 * 9 + 2 x 3 points) and the rebuilt host's mass table.
 * Replaces: SubhaloModel.fs / host (subhalos.py:672-772), ExactSector.
 * _clump_cells and host_cells_moments up to the emission (subhalos_exact.py:
 * 1553-1622, 1647-1811) and _ring_and_pair_angle_mean_cos (subhalos.py:287-330).
 */
size_t sgl_ref_sub_bytes(void);
sgl_status sgl_ref_sub_build(const void* tables, int32_t flags, const double* roots, void* sub,
                             int32_t* status, void* stream);

/*
 * Cells, first pass: per (row, orientation) task the cell count; counts
 * (int64[tasks + 1]) is left holding the exclusive offsets, counts[tasks] =
 * total cells.  status |= 32 (more edges than a task holds).  sub: the
 * subhalo tables when flags hold SGL_FLAG_SUBHALOS, else NULL (a host cell
 * within reach of a clump counts 9 + 2 x 3 points, or 9).
 * Replaces: the grid of _truncated_host (sgl_full.py:2025-2050, 2088-2127).
 */
sgl_status sgl_ref_cells_count(const void* tables, int32_t flags, const double* roots,
                               const void* sub, int64_t* counts, int32_t* status, void* stream);
/*
 * Cells, second pass: kappa, gamma (clipped below at 1.0000001e-9), weight and
 * the fine-lattice key (bin, or SGL_REF_FINE_BINS outside) of every cell:
 * per row the inner cells orientation by orientation (on a host with
 * subhalos the points of the decorated cells first, then the cells beyond
 * reach), the monopole tail (likewise), the filament cells.
 * Replaces: _truncated_host (sgl_full.py:2025-2138), the emission of the
 * decorated host (subhalos_exact.py:802-925), the filament block of
 * _build_shell_into (sgl_full.py:1984-2007) and the bin search of
 * _add_direct_kernel (sgl_full.py:971-1046).
 */
sgl_status sgl_ref_cells_emit(const void* tables, int32_t flags, const double* roots,
                              const void* sub, const int64_t* offsets, double* kap, double* gam,
                              double* w, int32_t* key, int32_t* status, void* stream);

/*
 * The xi_lin (Limber) two-point term, the "bias" ingredient with
 * clustering="xilin" and the Gaussian closure (pipeline.py:135-207).  flags
 * must hold SGL_FLAG_CLUSTERING on every call below and on sgl_ref_cells_count,
 * whose counts then has 2 (tasks + 1) slots: the second half holds the
 * exclusive offsets of the lens cells (distinct lens radii per task) and
 * counts[2 tasks + 1] their total.  Order: sgl_ref_xi_init -> count ->
 * sgl_ref_cells_emit_xi -> sgl_ref_bin -> sgl_ref_xi_sort -> sgl_ref_xi_rows ->
 * sgl_ref_xi_axis (reach probes) / sgl_ref_xi_grid (psi of sgl_ref_phi).
 * xi_state (sgl_ref_xi_state_bytes) is the per-PDF state: the k_perp rule, the
 * coarse lattice, the first moments, the compressed rows of every (shell, m)
 * group and the coarse-grid Psi_2 with its spline coefficients.
 */
size_t sgl_ref_xi_state_bytes(void);
/* k_perp rule (42 Gauss-Legendre nodes in ln k, weights k^2 P0(k) w / 2 pi) and
 * the coarse (kappa, gamma) edges.  Replaces: xi_kperp_grid
 * (sgl_full.py:2483-2511) and the coarse edges of Cells (sgl_full.py:457-468). */
sgl_status sgl_ref_xi_init(const void* tables, int32_t flags, void* xi_state, void* stream);
/*
 * sgl_ref_cells_emit plus, per point, its lens-cell id and coarse key (shell *
 * 341 * 110 + coarse bin, or past every key outside the lattice) in xi_pts
 * (int32[2 n_points]: ids, then keys) and per lens cell (R [Mpc], beta) in
 * xi_cells (double[2 n_cells]; R < 0 for points without tangential shear).
 * offsets: the counts of sgl_ref_cells_count (both halves).
 * Replaces: the point-to-radius bookkeeping of Cells._xi_deposit
 * (sgl_full.py:1397-1441) done while the cells are emitted.
 */
sgl_status sgl_ref_cells_emit_xi(const void* tables, int32_t flags, const double* roots,
                                 const void* sub, const int64_t* offsets, int64_t n_points,
                                 double* kap, double* gam, double* w, int32_t* key,
                                 const void* xi_state, int32_t* xi_pts, double* xi_cells,
                                 int32_t* status, void* stream);
/* out (device, 8 doubles) = [runs, parts, shells, kappa bins nP, gamma bins nQ,
 * clustering variance, largest singular value, 0]: the sizes the host needs
 * between sgl_ref_xi_sort, sgl_ref_xi_rows and sgl_ref_xi_grid. */
sgl_status sgl_ref_xi_info(const void* xi_state, double* out, void* stream);
/*
 * Stable sort of the points by coarse key (emission order kept inside a key),
 * runs of one key, their parts of at most 2048 points, and the J_{2m}(k_j R)
 * table of every lens cell.  ws: sgl_ref_xi_sort_bytes(n_points, n_cells),
 * read again by sgl_ref_xi_rows.
 * Replaces: the grouping of _xi_add_grouped and _bessel_pack_nb
 * (sgl_full.py:1080-1184).
 */
size_t sgl_ref_xi_sort_bytes(int64_t n_points, int64_t n_cells);
sgl_status sgl_ref_xi_sort(const void* tables, int32_t flags, int64_t n_points,
                           const int32_t* xi_pts, int64_t n_cells, const double* xi_cells,
                           void* ws, void* xi_state, void* stream);
/*
 * Deposit and compression: H = sum beta w J_{2m}(k_j R) per (shell, coarse
 * bin, m, j) summed run by run, the exact first moments and the halo rows'
 * analytic shear tail, then per (shell, m) group the eigenvectors of A^T A
 * (cyclic Jacobi in LDS) with the reference's singular-value cut and global
 * prune, the rotated axis tables and the clustering variance.
 * rows_ws: sgl_ref_xi_rows_bytes(n_runs, n_parts) (from sgl_ref_xi_info),
 * kept for sgl_ref_xi_grid.
 * Replaces: Cells._xi_deposit, xi_add_shear_tail and xi_end
 * (sgl_full.py:1397-1523), the prune of Lam (sgl_full.py:2801-2812) and the
 * clustering part of Lam.sigma (sgl_full.py:3437-3446).
 */
size_t sgl_ref_xi_rows_bytes(int64_t n_runs, int64_t n_parts);
sgl_status sgl_ref_xi_rows(const void* tables, int32_t flags, int64_t n_points, const double* kap,
                           const double* gam, const double* w, const int32_t* xi_pts,
                           int64_t n_cells, const double* xi_cells, const void* sort_ws,
                           int64_t n_runs, int64_t n_parts, void* rows_ws, void* xi_state,
                           void* stream);
/* Psi_2 on an axis at n probe points k (out: [n][2], complex): axis 0
 * Psi_2(k, 0), axis 1 Psi_2(0, k).  Replaces: _Qxi_gauss on the reach probes
 * (sgl_full.py:2869-2951, 38-47). */
sgl_status sgl_ref_xi_axis(const void* tables, int32_t flags, const void* xi_state,
                           const double* k, int32_t n, int32_t axis, double* out, void* stream);
/*
 * Psi_2 on the tensor grid linspace(0, kk_max, nk) x linspace(0, kg_max, ng)
 * (psi: complex [nk][ng], the psi of sgl_ref_phi): on the 101 x 101 coarse
 * grid k = k_max s^2 as batched FP64 MFMA contractions (gamma bins x J_m,
 * kappa bins x [cos; sin]), then the not-a-knot bicubic spline in
 * s = sqrt(k / k_max).  n_z, nP, nQ: from sgl_ref_xi_info;
 * grid_ws: sgl_ref_xi_grid_bytes(n_z, nP, nQ).
 * Replaces: _Qxi_gauss and Lam._Q_interpolated (sgl_full.py:2869-3022,
 * 3114-3128).
 */
size_t sgl_ref_xi_grid_bytes(int32_t n_z, int32_t nP, int32_t nQ);
sgl_status sgl_ref_xi_grid(const void* tables, int32_t flags, const void* xi_state,
                           const void* rows_ws, int64_t n_runs, int32_t n_z, int32_t nP,
                           int32_t nQ, void* grid_ws, double kk_max, double kg_max, int32_t nk,
                           int32_t ng, double* psi, void* stream);

/*
 * The positive ("vp") clustering closure on the xi_lin population
 * (xi_closure="vp", pipeline.py:154-207; pos_closure.install,
 * pos_closure.py:708-733): ln R = ln E_U exp(sum_s sum_c W_sc (lam_sc(U_s) - 1)
 * F_c) - shear_var k_gamma^2 / 4 replaces Psi_2.  Order, after
 * sgl_ref_cells_emit_xi: sgl_ref_vp_runs (before the cells are freed) ->
 * sgl_ref_bin -> the xi sort and rows -> sgl_ref_vp_setup -> sgl_ref_vp_variance
 * (as often as the eta root search needs) -> sgl_ref_vp_rates ->
 * sgl_ref_vp_lnr + sgl_ref_vp_qxi (reach probes, coarse grid) ->
 * sgl_ref_vp_spline (psi of sgl_ref_phi) -> sgl_ref_window when the reach did
 * not converge.
 */
/* The emitted cells keyed by (shell, fine bin), stably sorted and summed per
 * run (w, w kappa, w gamma) in fixed parts: the per-shell fine-lattice pieces.
 * offsets: the point offsets of sgl_ref_cells_count; *n_cells (device int64)
 * receives the number of (shell, bin) cells.  ws: sgl_ref_vp_runs_bytes,
 * read by sgl_ref_vp_setup.
 * Replaces: the per-shell Cells of build_population and their pos_pieces
 * (sgl_full.py:1693-1699, 1790-1802). */
size_t sgl_ref_vp_runs_bytes(int64_t n_points);
sgl_status sgl_ref_vp_runs(const void* tables, int32_t flags, int64_t n_points,
                           const int64_t* offsets, const int32_t* key, const double* kap,
                           const double* gam, const double* w, void* ws, int64_t* n_cells,
                           void* stream);
/*
 * Closure set-up: each cell's pooled bin and centroid (pop, npop of
 * sgl_ref_bin), its own centroid's coarse parent and that parent's loading
 * (from the xi rows: the m = 0 group's kappa mode over the lens rate), groups
 * of one shell and parent, the line-of-sight correlation of consecutive
 * shells on 30001 k_par nodes, f_s and the stationary transitions.  state:
 * sgl_ref_vp_state_bytes(n_cells); meta_out (device int64[16]): [0] cells
 * (bins with nonzero sums), [1] groups, [2] shells, [5] status bits (1 bins
 * disagree with pop, 2 a shell without a kappa mode, 8 a transition not
 * balanced).
 * Replaces: pos_closure.population and Closure.__init__ up to the
 * calibration (pos_closure.py:388-479), los_rho and transition
 * (pos_closure.py:251-267, 372-385).
 */
size_t sgl_ref_vp_state_bytes(int64_t n_cells);
sgl_status sgl_ref_vp_setup(const void* tables, int32_t flags, int64_t n_points,
                            const void* runs_ws, int64_t n_cells, const double* pop,
                            const int32_t* npop, const void* xi_state, const void* xi_rows_ws,
                            int64_t n_xi_runs, int64_t n_xi_parts, void* state, int64_t* meta_out,
                            void* stream);
/* *out (device) = the chain's kappa clustering variance at eta (status bit 4
 * in the state's record when a rate variance exceeds the 48-state cap).
 * Replaces: Closure.variance (pos_closure.py:560-573) with chain_variance
 * (pos_closure.py:352-369). */
sgl_status sgl_ref_vp_variance(void* state, int64_t n_cells, int32_t n_z, double eta, double* out,
                               void* stream);
/* The rates at the calibrated eta in interval form and the visiting order of
 * the groups (meta_out as sgl_ref_vp_setup's, [3] visited groups, [4] visited
 * cells).  Replaces: vp_rates / group_rates and rate_intervals
 * (pos_closure.py:273-318, pos_closure_fused.py:18-49). */
sgl_status sgl_ref_vp_rates(void* state, int64_t n_cells, int32_t n_z, double eta,
                            int64_t* meta_out, void* stream);
/* ln R on the tensor grid kv[nk] x qv[nq] (out: complex [nk][nq]; pop as in
 * sgl_ref_phi, n_bins its occupied count); ws: sgl_ref_vp_lnr_bytes.
 * Replaces: FusedVP.lnR_blocked / _fused_groups (pos_closure_fused.py:52-140)
 * and _chain_recursion (pos_closure.py:183-240). */
size_t sgl_ref_vp_lnr_bytes(int32_t n_z, int32_t n_bins, int32_t nk, int32_t nq);
sgl_status sgl_ref_vp_lnr(const void* state, int64_t n_cells, int32_t n_z, const double* pop,
                          int32_t n_bins, const double* kv, int32_t nk, const double* qv,
                          int32_t nq, void* ws, double* out, void* stream);
/* In place on ln R [nk][nq]: Im unwrapped along k_kappa at the first k_gamma,
 * then along k_gamma (numpy.unwrap), shear_var qv^2 / 4 subtracted from Re.
 * Replaces: Closure.qxi (pos_closure.py:690-705). */
sgl_status sgl_ref_vp_qxi(double* lnr, int32_t nk, int32_t nq, const double* qv,
                          double shear_var, void* stream);
/* psi [nk][ng] (complex) on k_kappa = linspace(0, K, nk) x k_gamma =
 * linspace(0, Q, ng) from coarse [nkc][nc] on k = K s_k^2 x Q s^2 (s uniform
 * on [0, 1]): the not-a-knot bicubic interpolant in (s_k, s).
 * ws: sgl_ref_vp_spline_bytes.  Replaces: Lam._Q_interpolated with a
 * separate k_kappa node count (sgl_full.py:2953-3022, 226-238). */
size_t sgl_ref_vp_spline_bytes(int32_t nkc, int32_t nc, int32_t nk);
sgl_status sgl_ref_vp_spline(const double* coarse, int32_t nkc, int32_t nc, int32_t nk,
                             int32_t ng, void* ws, double* psi, void* stream);
/* phi [2 nk][ng] (as sgl_ref_phi writes it) times the positive-definite window
 * (cubic B-spline in k_kappa x Wendland phi_{3,1} in k_gamma on the box).
 * Replaces: _window_pd and its use in invert (sgl_full.py:83-94, 328-333). */
sgl_status sgl_ref_window(double* phi, int32_t nk, int32_t ng, void* stream);

/*
 * Binning: stable sort by key, per-bin sums of w, w kappa, w gamma in the cells'
 * order, the occupied bins compacted in bin order into pop = [w][kappa
 * centroid][gamma centroid] (SGL_REF_FINE_BINS doubles each), npop = their
 * count, scal = [sum w kappa^2 over the bins, sum W kappa over every cell]
 * (sigma_kappa^2 without clustering, and kbar).  ws: sgl_ref_bin_bytes(n_cells).
 * Replaces: Cells.add (sgl_full.py:1314-1383) and the flattening of Lam.__init__
 * (sgl_full.py:2742-2755), Lam.sigma (sgl_full.py:3437-3452) without clustering.
 */
size_t sgl_ref_bin_bytes(int64_t n_cells);
sgl_status sgl_ref_bin(int64_t n_cells, const int32_t* key, const double* kap, const double* gam,
                       const double* w, void* ws, double* pop, int32_t* npop, double* scal,
                       void* stream);

/*
 * The Gaussian large-scale-structure term of clustering="limber" with the
 * "bias" ingredient (pipeline.py:126-134, 195-202): the population is built
 * without SGL_FLAG_CLUSTERING, and ln Phi gains -var/2 (k_kappa^2 +
 * k_gamma^2 / 2), sigma_kappa^2 gains var.
 * sgl_ref_lss_var: var of the instance in tables (after sgl_ref_tables_build
 * with the same flags) into *var (device, one double); ws:
 * sgl_ref_lss_bytes() of device memory.  Replaces: limber_lss_var
 * (sgl_full.py:2541-2572).
 * sgl_ref_lss_psi: the term on the tensor grid k_k x k_g into psi [nk][ng][2]
 * (the psi input of sgl_ref_phi).  sgl_ref_lss_axis: the term added to out
 * [n][2] of sgl_ref_lam_axis (axis 0: k_kappa = k, axis 1: k_gamma = k).
 * Replaces: the lss_var line of Lam.__call__ (sgl_full.py:3104-3105) on the
 * grid and on the reach probes (sgl_full.py:38-47).
 */
size_t sgl_ref_lss_bytes(void);
sgl_status sgl_ref_lss_var(const void* tables, int32_t flags, void* ws, double* var, void* stream);
sgl_status sgl_ref_lss_psi(double var, const double* k_k, int32_t nk, const double* k_g,
                           int32_t ng, double* psi, void* stream);
sgl_status sgl_ref_lss_axis(double var, const double* k, int32_t n, int32_t axis, double* out,
                            void* stream);

/*
 * Lambda on an axis at n probe points k (out: [n][2]): axis 0 Lambda(k, 0),
 * axis 1 Lambda(0, k).  Replaces: the Lam calls of _reach (sgl_full.py:38-47).
 */
sgl_status sgl_ref_lam_axis(const double* pop, const int32_t* npop, const double* k, int32_t n,
                            int32_t axis, double* out, void* stream);

/*
 * Phi = exp(Lambda + psi) on the tensor grid kk[nk] x kg[ng], written as
 * phi[2 nk][ng] (row 2n real, 2n + 1 imaginary part of k_kappa row n); psi
 * (complex [nk][ng], may be NULL) is the two-point term.  Lambda's dense part
 * is an FP64 MFMA contraction.  nbins: the host's copy of *npop;
 * ws: sgl_ref_phi_bytes(nk, ng, nbins).
 * Replaces: Lam.__call__ / _direct_call / _sep_raw (sgl_full.py:3156-3160,
 * 3088-3131, 3040-3086), sgl_direct.direct_real (sgl_direct.py:40-51) and the
 * exp of invert (sgl_full.py:272).
 */
size_t sgl_ref_phi_bytes(int32_t nk, int32_t ng, int32_t nbins);
sgl_status sgl_ref_phi(const double* pop, const int32_t* npop, int32_t nbins, const double* kk,
                       int32_t nk, const double* kg, int32_t ng, const double* psi, void* ws,
                       double* phi, void* stream);

/*
 * P3(kappa, gamma) on kap = linspace(kap_min, kap_max, n_kap) x gam =
 * linspace(0, gam_max, n_gam) from phi on k_k = linspace(0, kk_max, nk) x
 * k_g = linspace(0, kg_max, ng): Hankel then Fourier transform as FP64 MFMA
 * contractions, plus the Euler-Maclaurin endpoint terms of the k_gamma
 * trapezoid.  ws: sgl_ref_p3_bytes(nk, ng, n_kap, n_gam).
 * Replaces: sgl.P_vector_joint (sgl.py:1127-1156) and sgl_full.py:277-298.
 */
size_t sgl_ref_p3_bytes(int32_t nk, int32_t ng, int32_t n_kap, int32_t n_gam);
sgl_status sgl_ref_p3(const double* phi, int32_t nk, int32_t ng, double kk_max, double kg_max,
                      double kap_min, double kap_max, int32_t n_kap, double gam_max, int32_t n_gam,
                      void* ws, double* p3, void* stream);

/*
 * P_S(xi) at n_xi points xi (device): pi / mu^2 times the integral of the
 * bilinear interpolant of P3 along the constant-mu arc (2400 kappa points for
 * gamma <= g_def, the tail-box segment by gamma at half the gamma step),
 * divided by *norm when norm is not NULL.
 * Replaces: P_i of invert (sgl_full.py:299-360).
 */
sgl_status sgl_ref_arcs(const double* p3, double kap_min, double kap_max, int32_t n_kap,
                        double gam_max, int32_t n_gam, double g_def, const double* xi, int32_t n_xi,
                        const double* norm, double* out, void* stream);

/* mom = [norm, mean, variance, skewness, sigma_DL/D_L] of pw on xw (trapezoids).
 * Replaces: sgl_full.py:362-371. */
sgl_status sgl_ref_moments(const double* xw, const double* pw, int32_t n, double* mom,
                           void* stream);

/* C[M][N] = A[M][K] B[K][N], FP64 on MFMA (row-major, leading dimensions given).
 * The contraction of sgl_ref_phi and sgl_ref_p3; replaces the float64 GEMMs of
 * sgl_direct.direct_real (sgl_direct.py:46-50) and P_vector_joint. */
sgl_status sgl_ref_gemm_f64(int32_t M, int32_t N, int32_t K, const double* A, int64_t lda,
                            const double* B, int64_t ldb, double* C, int64_t ldc, void* stream);

/* ------------------------------------------------------------------------
 * Surrogate MLP (no reference counterpart: the reference is a solver, not an
 * emulator, README.md:8).  A linear layer's forward and both backward
 * products are one BF16 MFMA GEMM with strided operands:
 *   C[m][n] = sum_k A(m,k) B(n,k),  A(m,k) = A[m lda_m + k lda_k] (bf16),
 *   B(n,k) = B[n ldb_n + k ldb_k] (bf16), FP32 accumulation, and epilogue
 *   PLAIN      C = beta C + acc
 *   BIAS       C = acc + bias[n]
 *   BIAS_SILU  Zout = acc + bias[n]; C = silu(Zout)
 *   DSILU      C = acc * silu'(Z[m][n])
 * C (FP32) and/or Cb (BF16) receive the result, row stride ldc.
 */
#define SGL_MLP_EPI_PLAIN 0
#define SGL_MLP_EPI_BIAS 1
#define SGL_MLP_EPI_BIAS_SILU 2
#define SGL_MLP_EPI_DSILU 3
sgl_status sgl_mlp_gemm(const void* A, long lda_m, long lda_k, const void* B, long ldb_n,
                        long ldb_k, int32_t M, int32_t N, int32_t K, int32_t epi,
                        const float* bias, const float* Z, long ldz, float* Zout, float* C,
                        void* Cb, long ldc, float beta, void* stream);
/* db[j] = sum_n dZ[n][j] over an [N][M] FP32 matrix (bias gradient). */
sgl_status sgl_mlp_colsum(const float* dZ, int32_t N, int32_t M, float* db, void* stream);
/* y = bf16(x) (round to nearest even), n elements. */
sgl_status sgl_to_bf16(const float* x, size_t n, void* y, void* stream);
/* Surrogate output -> exponent on the solver's k grid: y[b][2 j + (0|1)] is
 * Re/Im ln phi at n_c coarse nodes (every (n_k-1)/(n_c-1)-th k node, first and
 * last coinciding); cubic Lagrange in ln k onto all n_k nodes, written as
 * lam[b][n] (re, im) for sgl_phi_pack (psi2 = NULL). */
sgl_status sgl_lnphi_interp(const float* y, long ldy, int32_t n_c, int32_t B,
                            const sgl_shape* shape, double* lam, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* SGL_AMD_H */
