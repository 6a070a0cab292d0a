"""ctypes binding of libsgl_amd.so (the C ABI declared in include/sgl_amd.h).

The library is built in-tree by ``python -m sgl_analytic_amd.build``.  There
is no fallback: if the library is missing or a GPU is absent, loading raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgl_amd.so")

SGL_OK, SGL_ERR_INVALID, SGL_ERR_HIP, SGL_ERR_UNSUPPORTED = 0, 1, 2, 3
FLAG_CLUSTERING, FLAG_WINDOW, FLAG_FILAMENTS, FLAG_EDGE_RVIR, FLAG_ELL = 1, 2, 4, 8, 16
FLAG_SUBHALOS = 64
ELL_PHI, MSTAR_N = 12, 1200   # orientation nodes per elliptical cell; M* sigma table
PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
COSMO_STRIDE = 16
REC_RADIATION, REC_GROWTH, REC_ZINT, REC_ZINT_N = 11, 12, 13, 14  # SGL_REC_*: option slots
STATUS_VP_TRANSITION = 4   # vp closure: transition balancing not converged
STATUS_VP_RATE = 8         # vp closure: rate variance beyond the 48-node cap, or no eta bracket

EXPORTS = (
    "sgl_version", "sgl_tables_bytes", "sgl_hist_bytes", "sgl_filon_bytes", "sgl_phi_bytes",
    "sgl_tables_build", "sgl_lnphi_exact", "sgl_bin_cells", "sgl_lnphi_binned",
    "sgl_quadform", "sgl_phi_pack", "sgl_filon_build", "sgl_invert_gemm", "sgl_invert_fp64",
    "sgl_diagnostics", "sgl_invert_fp64_at", "sgl_mlp_gemm", "sgl_mlp_colsum", "sgl_to_bf16",
    "sgl_lnphi_interp", "sgl_set_tuning", "sgl_get_tuning", "sgl_vp_bytes", "sgl_lnphi_exact_vp",
    "sgl_bin_pass", "sgl_bin_reduce", "sgl_psi2_nodes", "sgl_psi2_interp",
    "sgl_trunc_nfw_kg", "sgl_ref_shell_count", "sgl_ref_tasks", "sgl_ref_tables_bytes",
    "sgl_ref_roots_bytes", "sgl_ref_tables_build", "sgl_ref_cells_count", "sgl_ref_cells_emit",
    "sgl_ref_bin_bytes", "sgl_ref_bin", "sgl_ref_lam_axis", "sgl_ref_phi_bytes", "sgl_ref_phi",
    "sgl_ref_p3_bytes", "sgl_ref_p3", "sgl_ref_arcs", "sgl_ref_moments", "sgl_ref_gemm_f64",
    "sgl_lnphi_exact_noalloc", "sgl_ref_roots", "sgl_ref_sub_bytes", "sgl_ref_sub_build",
    "sgl_ref_cells_emit_xi", "sgl_ref_xi_state_bytes", "sgl_ref_xi_init", "sgl_ref_xi_sort_bytes",
    "sgl_ref_xi_sort", "sgl_ref_xi_rows_bytes", "sgl_ref_xi_rows", "sgl_ref_xi_axis",
    "sgl_ref_xi_grid_bytes", "sgl_ref_xi_grid", "sgl_ref_xi_info",
    "sgl_ref_vp_runs_bytes", "sgl_ref_vp_runs", "sgl_ref_vp_state_bytes", "sgl_ref_vp_setup",
    "sgl_ref_vp_variance", "sgl_ref_vp_rates", "sgl_ref_vp_lnr_bytes", "sgl_ref_vp_lnr",
    "sgl_ref_vp_qxi", "sgl_ref_vp_spline_bytes", "sgl_ref_vp_spline", "sgl_ref_window",
    "sgl_ref_lss_bytes", "sgl_ref_lss_var", "sgl_ref_lss_psi", "sgl_ref_lss_axis",
)
NOALLOC_SCRATCH_DOUBLES = 1 << 22   # capacity of sgl_lnphi_exact_noalloc's own array
REF_MAX_SHELLS, REF_MASSES, REF_FINE_BINS, REF_ROOT_STRIDE = 208, 99, 961 * 300, 128
REF_STATUS_ROOTS, REF_STATUS_EDGES, REF_STATUS_BAD = 16, 32, 2
# sgl_ref_xi_info record: [n_runs, n_parts, n_z, nP, nQ, variance, largest s_0,
# shear variance of the m = 1 groups]
XI_INFO_RUNS, XI_INFO_PARTS, XI_INFO_NZ, XI_INFO_NP, XI_INFO_NQ, XI_INFO_VAR = 0, 1, 2, 3, 4, 5
XI_INFO_SHEAR = 7
# sgl_ref_vp_setup / sgl_ref_vp_rates record (int64): cells, groups, shells,
# visited groups, visited cells, status bits
VP_META_CELLS, VP_META_GROUPS, VP_META_SHELLS, VP_META_VGROUPS, VP_META_VCELLS = 0, 1, 2, 3, 4
VP_META_STATUS = 5
VP_STATUS_BINS, VP_STATUS_MODE, VP_STATUS_RATE, VP_STATUS_TRANSITION = 1, 2, 4, 8
EPI_PLAIN, EPI_BIAS, EPI_BIAS_SILU, EPI_DSILU = 0, 1, 2, 3
TUNE_BIN_THREADS, TUNE_LNPHI_BLOCK, TUNE_GEMM_BM, TUNE_BIN_SCHEDULE = 0, 1, 2, 3
PSI2_STRIDE = 8  # SGL_PSI2_STRIDE: batched two-point sub-grid stride


class CShape(ctypes.Structure):
    _fields_ = [("n_z_max", ctypes.c_int32), ("n_m", ctypes.c_int32), ("n_b", ctypes.c_int32),
                ("n_k", ctypes.c_int32), ("n_out", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("m_lo_h", ctypes.c_double), ("m_hi_h", ctypes.c_double),
                ("u_lo", ctypes.c_double), ("u_hi", ctypes.c_double),
                ("k_lo", ctypes.c_double), ("k_hi", ctypes.c_double),
                ("out_lo", ctypes.c_double), ("out_hi", ctypes.c_double)]


class SglError(RuntimeError):
    pass


_LIB = None


def load(path: str = LIB_PATH):
    """Load the library and declare every exported signature."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise SglError(f"{path} is missing: build it with `python -m sgl_analytic_amd.build` "
                       "(there is no CPU fallback)")
    L = ctypes.CDLL(path)
    vp, i32, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_size_t
    sp = ctypes.POINTER(CShape)
    L.sgl_version.restype = ctypes.c_char_p
    L.sgl_version.argtypes = []
    for name in ("sgl_tables_bytes", "sgl_hist_bytes", "sgl_phi_bytes"):
        f = getattr(L, name)
        f.restype = sz
        f.argtypes = [sp, i32] + ([i32] if name == "sgl_phi_bytes" else [])
    L.sgl_vp_bytes.restype = sz
    L.sgl_vp_bytes.argtypes = [sp, i32]
    L.sgl_filon_bytes.restype = sz
    L.sgl_filon_bytes.argtypes = [sp, i32]
    L.sgl_get_tuning.restype = i32
    L.sgl_get_tuning.argtypes = [i32]
    L.sgl_psi2_nodes.restype = i32
    L.sgl_psi2_nodes.argtypes = [sp]
    f64, i64 = ctypes.c_double, ctypes.c_int64
    L.sgl_ref_shell_count.restype = i32
    L.sgl_ref_shell_count.argtypes = [f64]
    L.sgl_ref_tasks.restype = i32
    L.sgl_ref_tasks.argtypes = [i32]
    L.sgl_ref_tables_bytes.restype = sz
    L.sgl_ref_tables_bytes.argtypes = [i32]
    L.sgl_ref_roots_bytes.restype = sz
    L.sgl_ref_roots_bytes.argtypes = []
    L.sgl_ref_sub_bytes.restype = sz
    L.sgl_ref_sub_bytes.argtypes = []
    L.sgl_ref_bin_bytes.restype = sz
    L.sgl_ref_bin_bytes.argtypes = [i64]
    L.sgl_ref_phi_bytes.restype = sz
    L.sgl_ref_phi_bytes.argtypes = [i32, i32, i32]
    L.sgl_ref_xi_state_bytes.restype = sz
    L.sgl_ref_xi_state_bytes.argtypes = []
    L.sgl_ref_xi_sort_bytes.restype = sz
    L.sgl_ref_xi_sort_bytes.argtypes = [i64, i64]
    L.sgl_ref_xi_rows_bytes.restype = sz
    L.sgl_ref_xi_rows_bytes.argtypes = [i64, i64]
    L.sgl_ref_xi_grid_bytes.restype = sz
    L.sgl_ref_xi_grid_bytes.argtypes = [i32, i32, i32]
    L.sgl_ref_p3_bytes.restype = sz
    L.sgl_ref_p3_bytes.argtypes = [i32, i32, i32, i32]
    L.sgl_ref_vp_runs_bytes.restype = sz
    L.sgl_ref_vp_runs_bytes.argtypes = [i64]
    L.sgl_ref_vp_state_bytes.restype = sz
    L.sgl_ref_vp_state_bytes.argtypes = [i64]
    L.sgl_ref_vp_lnr_bytes.restype = sz
    L.sgl_ref_vp_lnr_bytes.argtypes = [i32, i32, i32, i32]
    L.sgl_ref_vp_spline_bytes.restype = sz
    L.sgl_ref_vp_spline_bytes.argtypes = [i32, i32, i32]
    L.sgl_ref_lss_bytes.restype = sz
    L.sgl_ref_lss_bytes.argtypes = []
    sigs = {
        "sgl_tables_build": [vp, i32, sp, vp, vp, vp],
        "sgl_lnphi_exact": [vp, i32, sp, vp, vp, vp, vp],
        "sgl_lnphi_exact_noalloc": [vp, i32, sp, vp, vp, vp],
        "sgl_lnphi_exact_vp": [vp, i32, sp, vp, vp, vp, vp, vp, vp],
        "sgl_bin_cells": [vp, i32, sp, vp, vp],
        "sgl_bin_pass": [vp, i32, sp, vp, vp],
        "sgl_bin_reduce": [vp, i32, sp, vp, vp],
        "sgl_lnphi_binned": [vp, vp, i32, sp, vp, vp, vp],
        "sgl_quadform": [vp, vp, i32, sp, i32, vp, vp],
        "sgl_phi_pack": [vp, vp, vp, i32, sp, i32, vp, vp],
        "sgl_filon_build": [vp, i32, sp, i32, vp, vp],
        "sgl_invert_gemm": [vp, vp, i32, sp, i32, vp, vp],
        "sgl_invert_fp64": [vp, vp, i32, sp, vp, vp],
        "sgl_diagnostics": [vp, i32, i32, sp, vp, vp],
        "sgl_invert_fp64_at": [vp, vp, i32, sp, vp, i32, vp, vp],
        "sgl_mlp_gemm": [vp, ctypes.c_long, ctypes.c_long, vp, ctypes.c_long, ctypes.c_long,
                         i32, i32, i32, i32, vp, vp, ctypes.c_long, vp, vp, vp, ctypes.c_long,
                         ctypes.c_float, vp],
        "sgl_mlp_colsum": [vp, i32, i32, vp, vp],
        "sgl_to_bf16": [vp, sz, vp, vp],
        "sgl_lnphi_interp": [vp, ctypes.c_long, i32, i32, sp, vp, vp],
        "sgl_set_tuning": [i32, i32],
        "sgl_psi2_interp": [vp, i32, sp, vp, vp],
        "sgl_trunc_nfw_kg": [vp, i32, ctypes.c_double, ctypes.c_double, i32, vp, vp, vp],
        "sgl_ref_tables_build": [vp, i32, vp, vp, vp],
        "sgl_ref_roots": [vp, i32, vp, vp, vp],
        "sgl_ref_sub_build": [vp, i32, vp, vp, vp, vp],
        "sgl_ref_cells_count": [vp, i32, vp, vp, vp, vp, vp],
        "sgl_ref_cells_emit": [vp, i32, vp, vp, vp, vp, vp, vp, vp, vp, vp],
        "sgl_ref_cells_emit_xi": [vp, i32, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp],
        "sgl_ref_xi_init": [vp, i32, vp, vp],
        "sgl_ref_xi_info": [vp, vp, vp],
        "sgl_ref_xi_sort": [vp, i32, i64, vp, i64, vp, vp, vp, vp],
        "sgl_ref_xi_rows": [vp, i32, i64, vp, vp, vp, vp, i64, vp, vp, i64, i64, vp, vp, vp],
        "sgl_ref_xi_axis": [vp, i32, vp, vp, i32, i32, vp, vp],
        "sgl_ref_xi_grid": [vp, i32, vp, vp, i64, i32, i32, i32, vp, f64, f64, i32, i32, vp, vp],
        "sgl_ref_bin": [i64, vp, vp, vp, vp, vp, vp, vp, vp, vp],
        "sgl_ref_lam_axis": [vp, vp, vp, i32, i32, vp, vp],
        "sgl_ref_phi": [vp, vp, i32, vp, i32, vp, i32, vp, vp, vp, vp],
        "sgl_ref_p3": [vp, i32, i32, f64, f64, f64, f64, i32, f64, i32, vp, vp, vp],
        "sgl_ref_arcs": [vp, f64, f64, i32, f64, i32, f64, vp, i32, vp, vp, vp],
        "sgl_ref_moments": [vp, vp, i32, vp, vp],
        "sgl_ref_gemm_f64": [i32, i32, i32, vp, i64, vp, i64, vp, i64, vp],
        "sgl_ref_vp_runs": [vp, i32, i64, vp, vp, vp, vp, vp, vp, vp, vp],
        "sgl_ref_vp_setup": [vp, i32, i64, vp, i64, vp, vp, vp, vp, i64, i64, vp, vp, vp],
        "sgl_ref_vp_variance": [vp, i64, i32, f64, vp, vp],
        "sgl_ref_vp_rates": [vp, i64, i32, f64, vp, vp],
        "sgl_ref_vp_lnr": [vp, i64, i32, vp, i32, vp, i32, vp, i32, vp, vp, vp],
        "sgl_ref_vp_qxi": [vp, i32, i32, vp, f64, vp],
        "sgl_ref_vp_spline": [vp, i32, i32, i32, i32, vp, vp, vp],
        "sgl_ref_window": [vp, i32, i32, vp],
        "sgl_ref_lss_var": [vp, i32, vp, vp, vp],
        "sgl_ref_lss_psi": [f64, vp, i32, vp, i32, vp, vp],
        "sgl_ref_lss_axis": [f64, vp, i32, i32, vp, vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = args
    _LIB = L
    return L


def check(rc: int, what: str):
    if rc != SGL_OK:
        names = {SGL_ERR_INVALID: "invalid argument", SGL_ERR_HIP: "HIP error",
                 SGL_ERR_UNSUPPORTED: "unsupported"}
        raise SglError(f"{what} failed: {names.get(rc, rc)}")
