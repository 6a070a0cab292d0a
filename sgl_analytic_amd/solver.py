"""Device drivers of the magnification-PDF hot path (plain and batched).

Every step runs in libsgl_amd.so on the current HIP stream; PyTorch only
owns the device buffers and the stream.  Two paths:

* plain (``path="exact"``): FP64 everywhere -- every cell against every k
  node, FP64 two-point term, FP64 Filon inversion evaluated on the fly.
* batched (``path="binned"``): cells histogrammed on a jump lattice, exponent
  with FP64 phase reduction / FP32 sin-cos / FP64 accumulation, inversion as
  one [n_out x 2 n_q] . [2 n_q x B] contraction on MFMA (FP32 or split
  BF16x3), optionally replayed as a captured HIP graph.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, replace

import numpy as np
import torch

from . import _lib
from ._lib import (COSMO_STRIDE, FLAG_CLUSTERING, FLAG_EDGE_RVIR, FLAG_ELL, FLAG_FILAMENTS,
                   FLAG_WINDOW, MSTAR_N, PREC_BF16X3, PREC_FP32, PREC_FP64, REC_GROWTH,
                   REC_RADIATION)

PARAM_NAMES = ("h", "Om", "sigma8", "Ob", "ns", "zeq", "Ok", "w0", "wa", "z_s")


@dataclass(frozen=True)
class Shape:
    """Grids of one solve.  Defaults: 64 slices x 256 masses x 512 impact
    parameters, 8192 log-k nodes in [1e-2, 1e6], 2048 ln(mu) nodes on [-0.5, 3].
    ``filaments`` adds the cylinder-filament population (120 u bins x 13
    orientations per (slice, mass)).  ``edge`` is the halo truncation: "r200"
    (x_t = c, the reference default) or "rvir" (x_t = r_vir / r_s).  ``ell``
    adds the halo ellipticity (12 orientation nodes per halo cell)."""
    n_z_max: int = 64
    n_m: int = 256
    n_b: int = 512
    n_k: int = 8192
    n_out: int = 2048
    clustering: bool = True
    window: bool = True
    filaments: bool = False
    edge: str = "r200"
    ell: bool = False
    m_lo_h: float = 1e8
    m_hi_h: float = 1e16
    u_lo: float = 1e-4
    u_hi: float = 10.0
    k_lo: float = 1e-2
    k_hi: float = 1e6
    out_lo: float = -0.5
    out_hi: float = 3.0

    def c(self) -> _lib.CShape:
        if self.edge not in ("r200", "rvir"):
            raise ValueError(f"edge must be 'r200' or 'rvir', not {self.edge!r}")
        flags = ((FLAG_CLUSTERING if self.clustering else 0) | (FLAG_WINDOW if self.window else 0)
                 | (FLAG_FILAMENTS if self.filaments else 0)
                 | (FLAG_EDGE_RVIR if self.edge == "rvir" else 0)
                 | (FLAG_ELL if self.ell else 0))
        return _lib.CShape(self.n_z_max, self.n_m, self.n_b, self.n_k, self.n_out, flags,
                           self.m_lo_h, self.m_hi_h, self.u_lo, self.u_hi, self.k_lo, self.k_hi,
                           self.out_lo, self.out_hi)

    def out_grid(self) -> np.ndarray:
        return np.linspace(self.out_lo, self.out_hi, self.n_out)

    def k_nodes(self) -> np.ndarray:
        return np.concatenate([[0.0], np.logspace(math.log10(self.k_lo), math.log10(self.k_hi),
                                                  self.n_k)])


def slice_count(z_s) -> np.ndarray:
    """Slices per instance for batched solves: ceil(24 (1 + ln(1 + z_s)))."""
    return np.ceil(24.0 * (1.0 + np.log1p(np.asarray(z_s, float)))).astype(np.int64)


def cosmo_records(params: np.ndarray, n_z, radiation="engine", growth_mode="ode") -> np.ndarray:
    """[B, 10] parameters (PARAM_NAMES order) + slice counts -> [B, 16] records.

    radiation: "engine" (Omega_r = Omega_m / (1 + z_eq)) or None; growth_mode:
    "ode" or "legacy" (LambdaCDM only; resolve "auto" first).  Both apply to
    every record."""
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    if params.shape[1] != 10:
        raise ValueError("params must have 10 columns: " + ", ".join(PARAM_NAMES))
    if radiation not in ("engine", None):
        raise ValueError("radiation must be 'engine' or None")
    if growth_mode not in ("ode", "legacy"):
        raise ValueError("growth_mode must be 'ode' or 'legacy' (resolve 'auto' first)")
    B = params.shape[0]
    if growth_mode == "legacy" and not np.all((params[:, 7] == -1.0) & (params[:, 8] == 0.0)):
        raise ValueError("legacy growth is only valid for LambdaCDM")
    rec = np.zeros((B, COSMO_STRIDE))
    rec[:, :10] = params
    rec[:, 10] = np.broadcast_to(np.asarray(n_z, dtype=np.float64), (B,))
    rec[:, REC_RADIATION] = 1.0 if radiation is None else 0.0
    rec[:, REC_GROWTH] = 1.0 if growth_mode == "legacy" else 0.0
    return rec


def vp_groups(shape: Shape) -> int:
    """Closure groups per slice: the halo mass rows, then the filament rows."""
    return shape.n_m * (2 if shape.filaments else 1)


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Solver:
    """Preallocated buffers for B instances of one shape.

    ``run(records)`` -> (p [B, n_out] tensor, diag [B, 8] float64 tensor).
    diag = (grid mass, mean, var, skew, sigma_DL/D_L, <mu>, P(mu > 2), negative mass).
    """

    def __init__(self, shape: Shape, B: int, path: str = "binned", precision: int = PREC_FP32,
                 device: str = "cuda:0", closure: str = "gauss"):
        if path not in ("exact", "binned"):
            raise ValueError("path must be 'exact' or 'binned'")
        if closure not in ("gauss", "vp"):
            raise ValueError("closure must be 'gauss' or 'vp'")
        if closure == "vp" and (path != "exact" or not shape.clustering):
            raise ValueError("the vp closure runs on the exact path with clustering")
        if path == "exact":
            precision = PREC_FP64
        elif precision not in (PREC_FP32, PREC_BF16X3):
            raise ValueError("batched precision must be PREC_FP32 or PREC_BF16X3")
        if not torch.cuda.is_available():
            raise _lib.SglError("no HIP device visible: the solver has no CPU path")
        self.L = _lib.load()
        self.shape, self.B, self.path, self.precision = shape, int(B), path, precision
        self.closure = closure
        self.dev = torch.device(device)
        self.cs = shape.c()
        cs = ctypes.byref(self.cs)
        L, B = self.L, self.B

        def buf(nbytes, dtype=torch.uint8):
            if nbytes == 0:
                raise _lib.SglError("shape or batch rejected by the library")
            return torch.zeros(int(nbytes), dtype=torch.uint8, device=self.dev)

        nk, nz = shape.n_k, shape.n_z_max
        self.records = torch.zeros((B, COSMO_STRIDE), dtype=torch.float64, device=self.dev)
        self.tables = buf(L.sgl_tables_bytes(cs, B))
        self.status = torch.zeros(1, dtype=torch.int32, device=self.dev)
        self.lam = torch.zeros((B, nk, 2), dtype=torch.float64, device=self.dev)
        # slice amplitudes: FP64 on every k node on the exact path; FP32 on the
        # two-point sub-grid (every PSI2_STRIDE-th node) on the batched path,
        # with Psi_2 computed there (psi2c) and interpolated onto every node
        if path == "exact":
            self.amp = torch.zeros((B, nz, nk, 2), dtype=torch.float64, device=self.dev)
        else:
            self.n_kc = int(L.sgl_psi2_nodes(cs))
            self.amp = torch.zeros((B, nz, self.n_kc, 2), dtype=torch.float32, device=self.dev)
            self.psi2c = torch.zeros((B, self.n_kc, 2), dtype=torch.float64, device=self.dev)
        self.psi2 = torch.zeros((B, nk, 2), dtype=torch.float64, device=self.dev)
        self.diag = torch.zeros((B, 8), dtype=torch.float64, device=self.dev)
        if path == "exact":
            self.work = torch.zeros((B, nz, nk, 2), dtype=torch.float64, device=self.dev)
            if closure == "vp":
                self.vpw = buf(L.sgl_vp_bytes(cs, B))
            self.phi = buf(L.sgl_phi_bytes(cs, B, PREC_FP64))
            self.p = torch.zeros((B, shape.n_out), dtype=torch.float64, device=self.dev)
        else:
            self.hist = buf(L.sgl_hist_bytes(cs, B))
            self.phi = buf(L.sgl_phi_bytes(cs, B, precision))
            self.E = buf(L.sgl_filon_bytes(cs, precision))
            self.p = torch.zeros((B, shape.n_out), dtype=torch.float32, device=self.dev)
            self._filon_ready = False  # E needs the k grid written by sgl_tables_build
        self.graph = None

    # -- one solve on the current stream ------------------------------------
    def _launch(self):
        L, B, cs, st = self.L, self.B, ctypes.byref(self.cs), _stream()
        ck = _lib.check
        ck(L.sgl_tables_build(_ptr(self.records), B, cs, _ptr(self.tables), _ptr(self.status), st),
           "sgl_tables_build")
        psi = None
        if self.closure == "vp":
            ck(L.sgl_lnphi_exact_vp(_ptr(self.tables), B, cs, _ptr(self.lam), _ptr(self.psi2),
                                    _ptr(self.work), _ptr(self.vpw), _ptr(self.status), st),
               "sgl_lnphi_exact_vp")
            psi = _ptr(self.psi2)
        elif self.path == "exact":
            ck(L.sgl_lnphi_exact(_ptr(self.tables), B, cs, _ptr(self.lam), _ptr(self.amp),
                                 _ptr(self.work), st), "sgl_lnphi_exact")
        else:
            ck(L.sgl_bin_cells(_ptr(self.tables), B, cs, _ptr(self.hist), st), "sgl_bin_cells")
            ck(L.sgl_lnphi_binned(_ptr(self.tables), _ptr(self.hist), B, cs, _ptr(self.lam),
                                  _ptr(self.amp), st), "sgl_lnphi_binned")
        if self.shape.clustering and self.closure == "gauss":
            if self.path == "exact":
                ck(L.sgl_quadform(_ptr(self.tables), _ptr(self.amp), B, cs, PREC_FP64,
                                  _ptr(self.psi2), st), "sgl_quadform")
            else:
                ck(L.sgl_quadform(_ptr(self.tables), _ptr(self.amp), B, cs, self.precision,
                                  _ptr(self.psi2c), st), "sgl_quadform")
                ck(L.sgl_psi2_interp(_ptr(self.psi2c), B, cs, _ptr(self.psi2), st),
                   "sgl_psi2_interp")
            psi = _ptr(self.psi2)
        ck(L.sgl_phi_pack(_ptr(self.lam), psi, _ptr(self.tables), B, cs, self.precision,
                          _ptr(self.phi), st), "sgl_phi_pack")
        if self.path == "exact":
            ck(L.sgl_invert_fp64(_ptr(self.phi), _ptr(self.tables), B, cs, _ptr(self.p), st),
               "sgl_invert_fp64")
            ck(L.sgl_diagnostics(_ptr(self.p), 1, B, cs, _ptr(self.diag), st), "sgl_diagnostics")
        else:
            if not self._filon_ready:
                ck(L.sgl_filon_build(_ptr(self.tables), B, cs, self.precision, _ptr(self.E), st),
                   "sgl_filon_build")
                self._filon_ready = True
            ck(L.sgl_invert_gemm(_ptr(self.E), _ptr(self.phi), B, cs, self.precision,
                                 _ptr(self.p), st), "sgl_invert_gemm")
            ck(L.sgl_diagnostics(_ptr(self.p), 0, B, cs, _ptr(self.diag), st), "sgl_diagnostics")

    def vp_calibration(self):
        """(eta [B], target [B]) of the last vp solve (host copies)."""
        e = self.vpw[:16 * self.B].view(torch.float64).cpu().numpy()
        return e[:self.B].copy(), e[self.B:].copy()

    def set_records(self, records):
        r = torch.as_tensor(np.asarray(records, dtype=np.float64))
        if tuple(r.shape) != (self.B, COSMO_STRIDE):
            raise ValueError(f"records must be [{self.B}, {COSMO_STRIDE}]")
        nz = r[:, 10]
        if bool((nz < 1).any()) or bool((nz > self.shape.n_z_max).any()):
            raise ValueError("slice counts must lie in [1, n_z_max]")
        self.records.copy_(r.to(self.dev))

    def run(self, records=None):
        if records is not None:
            self.set_records(records)
        if self.graph is not None:
            self.graph.replay()
        else:
            self._launch()
        return self.p, self.diag

    def table_views(self):
        """Host copies of the table sub-arrays (same offsets as the library's
        TablesLayout: each sub-array starts on a 32-double boundary)."""
        sh, B = self.shape, self.B
        sizes = [("m_nodes", sh.n_m), ("uedge", sh.n_b + 1), ("knodes", sh.n_k), ("status", 1),
                 ("cosmo", B * 16), ("derived", B * 16), ("slices", B * sh.n_z_max * 8),
                 ("mass", B * sh.n_m * 4), ("rows", B * sh.n_z_max * sh.n_m * 8),
                 ("frows", B * sh.n_z_max * sh.n_m * 4 if sh.filaments else 0),
                 ("cov", B * sh.n_z_max * sh.n_z_max),
                 ("sigtab", B * MSTAR_N if sh.ell else 0)]
        T = self.tables.view(torch.float64).cpu().numpy()
        out, o = {}, 0
        for name, n in sizes:
            out[name] = T[o:o + n]
            o += (n + 31) & ~31
        out["derived"] = out["derived"].reshape(B, 16)
        out["slices"] = out["slices"].reshape(B, sh.n_z_max, 8)
        out["mass"] = out["mass"].reshape(B, sh.n_m, 4)
        out["rows"] = out["rows"].reshape(B, sh.n_z_max, sh.n_m, 8)
        if sh.filaments:
            out["frows"] = out["frows"].reshape(B, sh.n_z_max, sh.n_m, 4)
        out["cov"] = out["cov"].reshape(B, sh.n_z_max, sh.n_z_max)
        return out

    def capture(self):
        """Capture one solve as a HIP graph (replayed by run())."""
        if self.path != "binned":
            raise ValueError("graph capture is for the batched path")
        torch.cuda.synchronize()
        self._launch()             # builds E outside the graph
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self._launch()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = g
        return g
