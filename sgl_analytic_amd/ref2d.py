"""The reference's own 2-D (k_kappa, k_gamma) chain on the GPU.

This is the chain ``pipeline.run`` executes (sgl_analytic/pipeline.py:114-274)
for the ingredient sets built here (halos, ellipticity, filaments), with the
reference's defaults: edge="r200", kap_floor=1e-8, zint="midpoint",
phi_tol=1e-8, tail box, no window, n_xi_wide=500.

Host side (this module): the scalar grid rules of ``invert``
(sgl_full.py:159-270, ``sgl.vector_grid`` sgl.py:1229-1258) and the k-reach
search (``_reach``, sgl_full.py:38-47), which read back a handful of
device scalars.  Device side (libsgl_amd.so, include/sgl_amd.h ``sgl_ref_*``):
the tables on the reference's shell and mass grids, the adaptive cells with
their critical-curve refinement, the binning into the fine (kappa, gamma)
lattice, Lambda on the reach probes and on the full tensor grid (FP64 MFMA),
the Hankel + Fourier inversion (FP64 MFMA), the constant-mu arc integral and
the normalisation.  There is no CPU fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from ._lib import FLAG_ELL, FLAG_FILAMENTS, REF_FINE_BINS, check

# invert / vector_grid constants (sgl.py:1176-1186, sgl_full.py:35-37, pipeline.py:117)
SIG_REACH, PPP_K, PPP_G, BOX_REACH = 12.0, 18.0, 96.0, 6.0
PTS_PER_SIG_KAP, PTS_PER_SIG_GAM, XI_REACH = 72.0, 36.0, 24.0
TAIL_KAP_MAX, TAIL_GAM_MAX, TAIL_PPP_G = 0.99, 0.8, 12.0
PHI_TOL = 1e-8
N_XI_WIDE = 500
N_XI_WIDE_VP = 4000              # pipeline.py:236 (positive closure)
XI_COARSE = 101                  # Lam xi_coarse
OUT_GRID_MAX = 3.4e7             # sgl_full.py:60 (geometric output axis above it)
E_BLOCK_BYTES = 4 << 30          # Lambda operand rows per call (k_kappa row blocks)
INGREDIENTS_2D = ("ell", "fil", "sub", "bias")  # ingredient sets this chain computes


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _records(cosmo: dict, zs: float, n_z: int, zint: str = "midpoint",
             zint_n: int = 1) -> np.ndarray:
    rec = np.zeros(_lib.COSMO_STRIDE)
    rec[:10] = [cosmo["h"], cosmo["Om"], cosmo["s8"], cosmo["Ob"], cosmo["ns"], cosmo["zeq"],
                cosmo.get("Ok", 0.0), cosmo.get("w0", -1.0), cosmo.get("wa", 0.0), zs]
    rec[10] = n_z
    rec[_lib.REC_RADIATION] = 1.0 if cosmo.get("radiation", "engine") is None else 0.0
    rec[_lib.REC_GROWTH] = 1.0 if cosmo.get("growth_mode", "ode") == "legacy" else 0.0
    rec[_lib.REC_ZINT] = {"midpoint": 0.0, "engine": 1.0, "gauss": 2.0}[zint]
    rec[_lib.REC_ZINT_N] = float(zint_n) if zint == "gauss" else 0.0
    return rec


@dataclass
class XiTerm:
    """The xi_lin two-point term on the device (sgl_ref_xi_*): the state (k_perp
    rule, first moments, rotated per-shell tables, spline), the per-run rows
    that the coarse grid reads, and their sizes."""
    tables: torch.Tensor
    flags: int
    state: torch.Tensor
    rows: torch.Tensor
    n_runs: int
    n_parts: int
    n_z: int
    nP: int
    nQ: int
    var: float


@dataclass
class VpTerm:
    """The positive closure on the device (sgl_ref_vp_*): the per-PDF state
    (cells, groups, rates, transitions), its sizes, the calibrated eta and the
    two clustering targets."""
    state: torch.Tensor
    n_cells: int            # (shell, bin) runs: the state's allocation
    cells: int              # the shells' cells (bins with nonzero sums)
    n_groups: int
    n_z: int
    eta: float
    target_kappa: float
    target_gamma: float
    n_bins: int


@dataclass
class Population:
    """Occupied fine bins on the device ([w][kappa centroid][gamma centroid],
    REF_FINE_BINS each) and the builder's bookkeeping (sgl_full.py:1535-1808,
    2742-2755); xi: the clustering term when built with bias; lss_var: the
    variance of the Gaussian term of clustering="limber" (0: none)."""
    pop: torch.Tensor
    npop_dev: torch.Tensor
    n_bins: int
    sigma: float
    kbar: float
    n_points: int
    N_expl: float
    N_fil: float
    timings: dict = field(default_factory=dict)
    xi: XiTerm = None
    vp: VpTerm = None
    lss_var: float = 0.0


class _Stages:
    """Wall time of each stage (synchronising the device) when a dict is given."""

    def __init__(self, out):
        self.out = out
        self.t = None
        if out is not None:
            import time
            self._clock = time.perf_counter
            torch.cuda.synchronize()
            self.t = self._clock()

    def mark(self, name):
        if self.out is None:
            return
        torch.cuda.synchronize()
        t = self._clock()
        self.out[name] = self.out.get(name, 0.0) + (t - self.t)
        self.t = t


def _xi_info(L, state: torch.Tensor) -> np.ndarray:
    out = torch.empty(8, dtype=torch.float64, device=state.device)
    check(L.sgl_ref_xi_info(_ptr(state), _ptr(out), _stream()), "sgl_ref_xi_info")
    return out.cpu().numpy()


def build(cosmo: dict, zs: float, ell: bool = False, fil: bool = False,
          edge: str = "r200", timings: dict = None, sub: bool = False,
          keep_sub: bool = False, bias: bool = False, vp: bool = False,
          lss: bool = False, zint: str = "midpoint", zint_n: int = 2) -> Population:
    """build_population(..., kap_floor=1e-8, trunc=edge, zint="midpoint")
    + Lam's flattening (sgl_full.py:1535, 2742-2755), on the device.  edge
    "rvir" truncates every halo at its virial radius (sgl_full.py:1876-1880).
    sub: subhalos on their hosts (the exact decorated-host sector with
    clumps truncated at r200, the pipeline's default for edge "r200";
    pipeline.py:171-183); keep_sub keeps the subhalo tables on the result.
    bias: the xi_lin clustering term (clustering="xilin", the pipeline's
    default for the "bias" ingredient, pipeline.py:135-207): every point's
    lens distance deposited per shell and coarse bin, the rows compressed,
    and sigma_kappa with the clustering variance (Lam.sigma).
    vp: the positive closure on that population (xi_closure="vp",
    pos_closure.install, pos_closure.py:708-733): the per-shell cells, their
    loadings and groups, the line-of-sight chain, eta calibrated on the
    clustering variance (brentq in ln eta, as pos_closure.py:575-591) and the
    rates in interval form; needs bias.
    lss: clustering="limber" with the bias ingredient (pipeline.py:195-202):
    no xi_lin deposit; the Gaussian term of limber_lss_var
    (sgl_full.py:2541-2572) in Lambda and in sigma_kappa; excludes bias.
    zint: the lens shells of z_shells (sgl_full.py:480-518), "midpoint" (the
    pipeline default), "engine" (the engine grid's intervals below z_s,
    one node at the upper edge) or "gauss" (the midpoint shells at zint_n
    Gauss-Legendre nodes each, every node its own slice); engine and gauss
    not with sub or vp."""
    L = _lib.load()
    dev = torch.device("cuda")
    if edge not in ("r200", "rvir"):
        raise ValueError("edge must be 'r200' or 'rvir'")
    if sub and edge != "r200":
        raise NotImplementedError("subhalos are built with the r200 edge only (clump_edge r200)")
    if vp and not bias:
        raise ValueError("the positive closure needs the xi_lin term (bias)")
    if lss and bias:
        raise ValueError("the limber term replaces the xi_lin term (bias)")
    flags = ((FLAG_ELL if ell else 0) | (FLAG_FILAMENTS if fil else 0)
             | (_lib.FLAG_EDGE_RVIR if edge == "rvir" else 0)
             | (_lib.FLAG_SUBHALOS if sub else 0)
             | (_lib.FLAG_CLUSTERING if bias else 0))
    if zint not in ("midpoint", "engine", "gauss"):
        raise ValueError("zint must be 'midpoint', 'engine' or 'gauss'")
    if zint != "midpoint" and (sub or vp):
        raise NotImplementedError(f"{zint} shells are built without subhalos and the vp closure")
    n_mid = int(L.sgl_ref_shell_count(float(zs)))
    n_z = {"midpoint": n_mid, "engine": n_mid - 2, "gauss": n_mid * int(zint_n)}[zint]
    if zint == "gauss" and n_z > _lib.REF_MAX_SHELLS:
        raise NotImplementedError(f"zint='gauss' with {zint_n} nodes on {n_mid} shells exceeds the "
                                  f"{_lib.REF_MAX_SHELLS} lens nodes of the tables")
    if n_z < 1 or n_z > _lib.REF_MAX_SHELLS:
        raise ValueError("z_s outside the reference grid's shells")
    st = _stream()
    clock = _Stages(timings)
    rec = torch.tensor(_records(cosmo, float(zs), n_z, zint, int(zint_n)), dtype=torch.float64,
                       device=dev)
    T = torch.empty(L.sgl_ref_tables_bytes(flags) // 8, dtype=torch.float64, device=dev)
    status = torch.zeros(1, dtype=torch.int32, device=dev)
    check(L.sgl_ref_tables_build(_ptr(rec), flags, _ptr(T), _ptr(status), st), "sgl_ref_tables_build")
    lss_var = 0.0
    if lss:
        lws = torch.empty(L.sgl_ref_lss_bytes() // 8, dtype=torch.float64, device=dev)
        lv = torch.empty(1, dtype=torch.float64, device=dev)
        check(L.sgl_ref_lss_var(_ptr(T), flags, _ptr(lws), _ptr(lv), st), "sgl_ref_lss_var")
        lss_var = float(lv.item())
        del lws
    clock.mark("tables")
    roots = torch.empty(L.sgl_ref_roots_bytes() // 8, dtype=torch.float64, device=dev)
    check(L.sgl_ref_roots(_ptr(T), flags, _ptr(roots), _ptr(status), st), "sgl_ref_roots")
    subw = None
    if sub:
        subw = torch.empty(L.sgl_ref_sub_bytes() // 8, dtype=torch.float64, device=dev)
        check(L.sgl_ref_sub_build(_ptr(T), flags, _ptr(roots), _ptr(subw), _ptr(status), st),
              "sgl_ref_sub_build")
        clock.mark("subhalos")
    sp = _ptr(subw) if subw is not None else None
    ntask = int(L.sgl_ref_tasks(flags))
    counts = torch.empty((2 if bias else 1) * (ntask + 1), dtype=torch.int64, device=dev)
    check(L.sgl_ref_cells_count(_ptr(T), flags, _ptr(roots), sp, _ptr(counts), _ptr(status), st),
          "sgl_ref_cells_count")
    tot = counts[[ntask, 2 * ntask + 1] if bias else [ntask]].cpu().numpy()
    n = int(tot[0])
    clock.mark("cells_count")
    if n < 1:
        raise _lib.SglError(f"reference population is empty (status {int(status.item())})")
    K = torch.empty(n, dtype=torch.float64, device=dev)
    G = torch.empty_like(K)
    W = torch.empty_like(K)
    key = torch.empty(n, dtype=torch.int32, device=dev)
    if bias:
        n_cells = int(tot[1])
        xs = torch.empty(L.sgl_ref_xi_state_bytes() // 8, dtype=torch.float64, device=dev)
        check(L.sgl_ref_xi_init(_ptr(T), flags, _ptr(xs), st), "sgl_ref_xi_init")
        pts = torch.empty(2 * n, dtype=torch.int32, device=dev)
        cells = torch.empty(2 * n_cells, dtype=torch.float64, device=dev)
        check(L.sgl_ref_cells_emit_xi(_ptr(T), flags, _ptr(roots), sp, _ptr(counts), n, _ptr(K),
                                      _ptr(G), _ptr(W), _ptr(key), _ptr(xs), _ptr(pts),
                                      _ptr(cells), _ptr(status), st), "sgl_ref_cells_emit_xi")
    else:
        check(L.sgl_ref_cells_emit(_ptr(T), flags, _ptr(roots), sp, _ptr(counts), _ptr(K),
                                   _ptr(G), _ptr(W), _ptr(key), _ptr(status), st),
              "sgl_ref_cells_emit")
    clock.mark("cells_emit")
    vruns = None
    if vp:
        # the cells keyed by (shell, fine bin): the per-shell pieces
        vruns = torch.empty(L.sgl_ref_vp_runs_bytes(n), dtype=torch.uint8, device=dev)
        ncell_d = torch.zeros(1, dtype=torch.int64, device=dev)
        check(L.sgl_ref_vp_runs(_ptr(T), flags, n, _ptr(counts), _ptr(key), _ptr(K), _ptr(G),
                                _ptr(W), _ptr(vruns), _ptr(ncell_d), st), "sgl_ref_vp_runs")
        clock.mark("vp_runs")
    ws = torch.empty(L.sgl_ref_bin_bytes(n), dtype=torch.uint8, device=dev)
    pop = torch.empty(3 * REF_FINE_BINS, dtype=torch.float64, device=dev)
    npop = torch.zeros(1, dtype=torch.int32, device=dev)
    scal = torch.zeros(2, dtype=torch.float64, device=dev)
    check(L.sgl_ref_bin(n, _ptr(key), _ptr(K), _ptr(G), _ptr(W), _ptr(ws), _ptr(pop), _ptr(npop),
                        _ptr(scal), st), "sgl_ref_bin")
    clock.mark("bin")
    s = int(status.item())
    if s:
        raise _lib.SglError(f"reference population status {s} (16: critical-curve roots beyond the "
                            "slots, 32: edges beyond the per-row capacity, 2: bad record)")
    del ws
    xi = None
    if bias:
        sws = torch.empty(L.sgl_ref_xi_sort_bytes(n, n_cells), dtype=torch.uint8, device=dev)
        check(L.sgl_ref_xi_sort(_ptr(T), flags, n, _ptr(pts), n_cells, _ptr(cells), _ptr(sws),
                                _ptr(xs), st), "sgl_ref_xi_sort")
        info = _xi_info(L, xs)
        n_runs, n_parts = int(info[_lib.XI_INFO_RUNS]), int(info[_lib.XI_INFO_PARTS])
        rws = torch.empty(L.sgl_ref_xi_rows_bytes(n_runs, n_parts), dtype=torch.uint8, device=dev)
        check(L.sgl_ref_xi_rows(_ptr(T), flags, n, _ptr(K), _ptr(G), _ptr(W), _ptr(pts), n_cells,
                                _ptr(cells), _ptr(sws), n_runs, n_parts, _ptr(rws), _ptr(xs), st),
              "sgl_ref_xi_rows")
        info = _xi_info(L, xs)
        del sws, pts, cells
        xi = XiTerm(tables=T, flags=flags, state=xs, rows=rws, n_runs=n_runs, n_parts=n_parts,
                    n_z=int(info[_lib.XI_INFO_NZ]), nP=int(info[_lib.XI_INFO_NP]),
                    nQ=int(info[_lib.XI_INFO_NQ]), var=float(info[_lib.XI_INFO_VAR]))
        clock.mark("xi_rows")
    vpt = None
    if vp:
        vpt = _vp_setup(L, T, flags, n, vruns, int(ncell_d.item()), pop, npop, xi, n_z,
                        float(info[_lib.XI_INFO_SHEAR]))
        del vruns
        clock.mark("vp_setup")
    del K, G, W, key
    sc = scal.cpu().numpy()
    # bookkeeping of the info record (sgl_full.py:2039, 1992): expected counts
    lay_rows = _row_table(T, flags, n_z)
    var = sc[0] + (xi.var if xi is not None else 0.0) + lss_var
    out = Population(pop=pop, npop_dev=npop, n_bins=int(npop.item()), sigma=math.sqrt(var),
                     kbar=float(sc[1]), n_points=n, N_expl=lay_rows[0],
                     N_fil=lay_rows[1] if fil else 0.0, xi=xi, vp=vpt, lss_var=lss_var)
    if vpt is not None:
        vpt.n_bins = out.n_bins
    if keep_sub:
        out.sub_tables = subw
    return out


def _vp_setup(L, T, flags, n_points, vruns, n_cells, pop, npop, xi, n_z, shear_var) -> VpTerm:
    """The positive closure's set-up on the device, then eta: the root of
    variance(eta) = the clustering variance in ln eta, bracketed from
    [1/1.5, 1.5] widened by 1.5 per side inside [1e-3, 1e4] and refined by
    brentq to 2e-12 (pos_closure.Closure.calibrate, pos_closure.py:575-591);
    each variance is one device evaluation (rates, per-shell sums, chain)."""
    from scipy.optimize import brentq
    dev = T.device
    st = _stream()
    state = torch.empty(L.sgl_ref_vp_state_bytes(n_cells), dtype=torch.uint8, device=dev)
    meta = torch.zeros(16, dtype=torch.int64, device=dev)
    check(L.sgl_ref_vp_setup(_ptr(T), flags, n_points, _ptr(vruns), n_cells, _ptr(pop),
                             _ptr(npop), _ptr(xi.state), _ptr(xi.rows), xi.n_runs, xi.n_parts,
                             _ptr(state), _ptr(meta), st), "sgl_ref_vp_setup")
    m = meta.cpu().numpy()
    bad = int(m[_lib.VP_META_STATUS])
    if bad & _lib.VP_STATUS_BINS:
        raise _lib.SglError("positive closure: the per-shell bins disagree with the pooled lattice")
    if bad & _lib.VP_STATUS_MODE:
        raise NotImplementedError("positive closure: a shell without a kappa clustering mode")
    if bad & _lib.VP_STATUS_TRANSITION:
        raise _lib.SglError("positive closure: transition balancing did not converge")
    target = xi.var
    vd = torch.empty(1, dtype=torch.float64, device=dev)

    def variance(eta):
        check(L.sgl_ref_vp_variance(_ptr(state), n_cells, n_z, float(eta), _ptr(vd), st),
              "sgl_ref_vp_variance")
        return float(vd.item())

    def fun(le):
        return variance(math.exp(le)) - target

    lo, hi = 1e-3, 1e4
    a, b = -math.log(1.5), math.log(1.5)
    fa, fb = fun(a), fun(b)
    while fa * fb > 0:
        if fa > 0 and a > math.log(lo):
            b, fb = a, fa
            a -= math.log(1.5)
            fa = fun(a)
        elif fb < 0 and b < math.log(hi):
            a, fa = b, fb
            b += math.log(1.5)
            fb = fun(b)
        else:
            raise ValueError(f"positive closure: kappa clustering target {target:.3e} outside "
                             f"[{lo}, {hi}]")
    eta = float(math.exp(brentq(fun, a, b, xtol=2e-12)))
    check(L.sgl_ref_vp_rates(_ptr(state), n_cells, n_z, eta, _ptr(meta), st), "sgl_ref_vp_rates")
    m = meta.cpu().numpy()
    if int(m[_lib.VP_META_STATUS]) & _lib.VP_STATUS_RATE:
        raise ValueError("vp closure: a relative rate variance beyond the 48-state cap")
    return VpTerm(state=state, n_cells=n_cells, cells=int(m[_lib.VP_META_CELLS]),
                  n_groups=int(m[_lib.VP_META_GROUPS]), n_z=n_z,
                  eta=eta, target_kappa=target, target_gamma=shear_var, n_bins=0)


def vp_qxi(pop: Population, kv: np.ndarray, qv: np.ndarray) -> torch.Tensor:
    """The positive closure's ln R - shear_var k_gamma^2 / 4 on the tensor grid
    kv x qv, Im ln R unwrapped (pos_closure.py:690-705): [nk][nq][2]."""
    L = _lib.load()
    v = pop.vp
    dev = pop.pop.device
    kd = torch.tensor(np.asarray(kv, float), dtype=torch.float64, device=dev)
    qd = torch.tensor(np.asarray(qv, float), dtype=torch.float64, device=dev)
    nk, nq = kd.numel(), qd.numel()
    ws = torch.empty(L.sgl_ref_vp_lnr_bytes(v.n_z, v.n_bins, nk, nq), dtype=torch.uint8,
                     device=dev)
    out = torch.empty((nk, nq, 2), dtype=torch.float64, device=dev)
    st = _stream()
    check(L.sgl_ref_vp_lnr(_ptr(v.state), v.n_cells, v.n_z, _ptr(pop.pop), v.n_bins, _ptr(kd), nk,
                           _ptr(qd), nq, _ptr(ws), _ptr(out), st), "sgl_ref_vp_lnr")
    del ws
    check(L.sgl_ref_vp_qxi(_ptr(out), nk, nq, _ptr(qd), v.target_gamma, st), "sgl_ref_vp_qxi")
    return out


def vp_psi(pop: Population, k_k: np.ndarray, k_g: np.ndarray, n: int, nk: int) -> torch.Tensor:
    """qxi on the closure's coarse grid (k_kappa = k_k_max s_k^2 on nk nodes,
    k_gamma = k_g_max s^2 on n nodes) and its not-a-knot bicubic interpolant on
    the full grid (Lam._Q_interpolated, sgl_full.py:2953-3022): [nk_k][nk_g][2]."""
    L = _lib.load()
    dev = pop.pop.device
    s = np.linspace(0.0, 1.0, n)
    s_k = s if nk == n else np.linspace(0.0, 1.0, nk)
    coarse = vp_qxi(pop, float(k_k[-1]) * s_k ** 2, float(k_g[-1]) * s ** 2)
    ws = torch.empty(L.sgl_ref_vp_spline_bytes(nk, n, k_k.size), dtype=torch.uint8, device=dev)
    psi = torch.empty((k_k.size, k_g.size, 2), dtype=torch.float64, device=dev)
    check(L.sgl_ref_vp_spline(_ptr(coarse), nk, n, k_k.size, k_g.size, _ptr(ws), _ptr(psi),
                              _stream()), "sgl_ref_vp_spline")
    return psi


def _row_table(T: torch.Tensor, flags: int, n_z: int):
    """N_expl = sum over rows of pi w_row (x_m r_s)^2 and N_fil = sum of
    pi w_fil r_sF^2, read from the device rows (sgl_full.py:2039, 1992)."""
    # tables layout: grids (m_nodes 99, uedge 2, knodes 2, status 1), cosmo,
    # derived, slices, mass, rows, frows (each rounded to 32 doubles)
    r32 = lambda v: (v + 31) // 32 * 32
    S, NM = _lib.REF_MAX_SHELLS, _lib.REF_MASSES
    o = r32(NM) + r32(2) + r32(2) + r32(1) + r32(16) + r32(16) + r32(S * 8) + r32(NM * 4)
    rows = T[o:o + S * NM * 8].view(S, NM, 8)[:n_z].cpu().numpy()
    ks, rs, xt, w = rows[..., 0], rows[..., 1], rows[..., 2], rows[..., 4]
    mut = np.log(1 + xt) - xt / (1 + xt)
    g = ks * 4 * mut / xt ** 2
    xm = np.where(g > 1e-8, xt * np.sqrt(np.maximum(g, 0) / 1e-8), xt)
    n_expl = float(np.sum(math.pi * w * (xm * rs) ** 2))
    n_fil = 0.0
    if flags & FLAG_FILAMENTS:
        of = o + r32(S * NM * 8)
        fr = T[of:of + S * NM * 4].view(S, NM, 4)[:n_z].cpu().numpy()
        n_fil = float(np.sum(math.pi * fr[..., 2] * fr[..., 1] ** 2))
    return n_expl, n_fil


def lam_axis(pop: Population, k: np.ndarray, axis: int) -> np.ndarray:
    """Lambda(k, 0) (axis 0) or Lambda(0, k) (axis 1) at the probe points."""
    L = _lib.load()
    kd = torch.tensor(np.asarray(k, float), dtype=torch.float64, device="cuda")
    out = torch.empty(2 * kd.numel(), dtype=torch.float64, device="cuda")
    check(L.sgl_ref_lam_axis(_ptr(pop.pop), _ptr(pop.npop_dev), _ptr(kd), kd.numel(), axis,
                             _ptr(out), _stream()), "sgl_ref_lam_axis")
    if pop.vp is not None:
        zero = np.zeros(1)
        q = vp_qxi(pop, k, zero) if axis == 0 else vp_qxi(pop, zero, k)
        o = out.cpu().numpy()
        qq = q.reshape(-1, 2).cpu().numpy()
        return (o[0::2] + qq[:, 0]) + 1j * (o[1::2] + qq[:, 1])
    if pop.xi is not None:
        x = pop.xi
        check(L.sgl_ref_xi_axis(_ptr(x.tables), x.flags, _ptr(x.state), _ptr(kd), kd.numel(), axis,
                                _ptr(out), _stream()), "sgl_ref_xi_axis")
    if pop.lss_var:
        check(L.sgl_ref_lss_axis(pop.lss_var, _ptr(kd), kd.numel(), axis, _ptr(out), _stream()),
              "sgl_ref_lss_axis")
    o = out.cpu().numpy()
    return o[0::2] + 1j * o[1::2]


def vector_grid(sigma: float) -> dict:
    """sgl.vector_grid (sgl.py:1229-1258)."""
    sig_xi = 2.0 * sigma
    xi_max = XI_REACH * sig_xi
    xi_min = -BOX_REACH * sig_xi
    kap_max = 1.0 - math.exp(-0.5 * xi_max)
    kap_min = -BOX_REACH * sigma
    gam_max = BOX_REACH * sigma
    k_max = SIG_REACH / sigma
    kap_ext = max(abs(kap_min), kap_max)
    dk_k = 2.0 * math.pi / (PPP_K * kap_ext)
    dk_g = 2.0 * math.pi / (PPP_G * gam_max)
    return dict(nk_k=int(math.ceil(k_max / dk_k)) + 1, nk_g=int(math.ceil(k_max / dk_g)) + 1,
                k_k_max=k_max, k_g_max=k_max,
                n_kap=int(math.ceil((kap_max - kap_min) / (sigma / PTS_PER_SIG_KAP))) + 1,
                n_gam=int(math.ceil(gam_max / (sigma / PTS_PER_SIG_GAM))) + 1,
                kap_min=kap_min, kap_max=kap_max, gam_max=gam_max, xi_min=xi_min,
                xi_max=xi_max, sigma=sigma)


def reach(pop: Population, k0: float, axis: int, tol: float = PHI_TOL):
    """_reach (sgl_full.py:38-47): the first of k0 * 1.25^n, n = 0..19, with
    |Phi| <= tol on the axis, else the smallest |Phi|."""
    mults = 1.25 ** np.arange(0, int(np.ceil(np.log(64.0) / np.log(1.25))) + 1)
    ks = k0 * mults
    phi = np.abs(np.exp(lam_axis(pop, ks, axis)))
    ok = np.nonzero(phi <= tol)[0]
    i = int(ok[0]) if len(ok) else int(np.argmin(phi))
    return float(ks[i]), float(phi[i])


def grid_plan(pop: Population, tol: float = PHI_TOL) -> dict:
    """The k and (kappa, gamma) grids of invert with the tail box
    (sgl_full.py:190-262)."""
    sigma = pop.sigma
    grid = vector_grid(sigma)
    g0 = grid["gam_max"]
    kap_max = max(grid["kap_max"], TAIL_KAP_MAX)
    gam_max = max(g0, TAIL_GAM_MAX)
    dk0 = (2 * math.pi / (PPP_K * max(abs(grid["kap_min"]), grid["kap_max"])),
           2 * math.pi / (PPP_G * grid["gam_max"]))
    dk0 = (dk0[0], min(dk0[1], 2 * math.pi / (TAIL_PPP_G * gam_max)))
    if kap_max != grid["kap_max"] or gam_max != grid["gam_max"]:
        grid["kap_max"], grid["gam_max"] = float(kap_max), float(gam_max)
        grid["n_kap"] = int(math.ceil((grid["kap_max"] - grid["kap_min"])
                                      / (sigma / PTS_PER_SIG_KAP))) + 1
        grid["n_gam"] = int(math.ceil(grid["gam_max"] / (sigma / PTS_PER_SIG_GAM))) + 1
    grid["k_k_max"], phi_k = reach(pop, grid["k_k_max"], 0, tol)
    grid["k_g_max"], phi_g = reach(pop, grid["k_g_max"], 1, tol)
    grid["nk_k"] = int(math.ceil(grid["k_k_max"] / dk0[0])) + 1
    grid["nk_g"] = int(math.ceil(grid["k_g_max"] / dk0[1])) + 1
    grid.update(phi_at_k_cut=phi_k, phi_at_g_cut=phi_g, window=0.0)
    if pop.vp is not None:
        # the closure's coarse grid follows the box, per axis (sgl_full.py:226-238)
        n = int(min(401, max(XI_COARSE, 1 + math.ceil(12.65 * math.sqrt(grid["k_g_max"] * sigma)))))
        nk = int(min(4001, max(n, 1 + math.ceil(66.7 * math.sqrt(grid["k_k_max"] * sigma)))))
        grid.update(xi_coarse=float(n), xi_coarse_k=float(nk))
        if max(phi_k, phi_g) > tol:
            grid["window"] = 1.0
    if grid["n_kap"] * grid["n_gam"] > OUT_GRID_MAX:
        raise NotImplementedError("the geometric output axis of the narrowest low-z_s grids "
                                  "(n_kap * n_gam > 3.4e7, sgl_full.py:316-324) is not built")
    return grid


def phi_grid(pop: Population, k_k: np.ndarray, k_g: np.ndarray,
             grid: dict = None) -> torch.Tensor:
    """Phi = exp(Lambda + clustering term) on the tensor grid as [2 nk_k][nk_g]
    (rows re, im); with the positive closure the grid plan gives its coarse
    grid and whether the positive-definite window applies."""
    L = _lib.load()
    dev = torch.device("cuda")
    nk, ng = k_k.size, k_g.size
    phi = torch.empty((2 * nk, ng), dtype=torch.float64, device=dev)
    kg = torch.tensor(k_g, dtype=torch.float64, device=dev)
    lda = (pop.n_bins + 63) // 64 * 64
    blk = max(1, min(nk, 65535, E_BLOCK_BYTES // (16 * lda)))
    if grid is not None and pop.vp is not None:
        psi = vp_psi(pop, k_k, k_g, int(grid["xi_coarse"]), int(grid["xi_coarse_k"]))
    elif pop.xi is not None:
        psi = xi_psi(pop, k_k, k_g)
    elif pop.lss_var:
        psi = torch.empty((nk, ng, 2), dtype=torch.float64, device=dev)
        kkd = torch.tensor(k_k, dtype=torch.float64, device=dev)
        check(L.sgl_ref_lss_psi(pop.lss_var, _ptr(kkd), nk, _ptr(kg), ng, _ptr(psi), _stream()),
              "sgl_ref_lss_psi")
    else:
        psi = None
    ws = torch.empty(L.sgl_ref_phi_bytes(blk, ng, pop.n_bins) // 8, dtype=torch.float64, device=dev)
    for lo in range(0, nk, blk):
        hi = min(nk, lo + blk)
        kk = torch.tensor(k_k[lo:hi], dtype=torch.float64, device=dev)
        check(L.sgl_ref_phi(_ptr(pop.pop), _ptr(pop.npop_dev), pop.n_bins, _ptr(kk), hi - lo,
                            _ptr(kg), ng, _ptr(psi[lo]) if psi is not None else None, _ptr(ws),
                            _ptr(phi[2 * lo:]), _stream()),
              "sgl_ref_phi")
    if grid is not None and grid.get("window", 0.0):
        check(L.sgl_ref_window(_ptr(phi), nk, ng, _stream()), "sgl_ref_window")
    return phi


def xi_psi(pop: Population, k_k: np.ndarray, k_g: np.ndarray) -> torch.Tensor:
    """The xi_lin term Psi_2 on the tensor grid k_k x k_g (linspaces from 0),
    [nk][ng][2]: the 101 x 101 coarse grid and its bicubic spline
    (Lam._Q_interpolated with fn=_Qxi, sgl_full.py:2953-3022, 3114-3128)."""
    L = _lib.load()
    x = pop.xi
    nk, ng = k_k.size, k_g.size
    gws = torch.empty(L.sgl_ref_xi_grid_bytes(x.n_z, x.nP, x.nQ), dtype=torch.uint8,
                      device=x.state.device)
    psi = torch.empty((nk, ng, 2), dtype=torch.float64, device=x.state.device)
    check(L.sgl_ref_xi_grid(_ptr(x.tables), x.flags, _ptr(x.state), _ptr(x.rows), x.n_runs, x.n_z,
                            x.nP, x.nQ, _ptr(gws), float(k_k[-1]), float(k_g[-1]), nk, ng,
                            _ptr(psi), _stream()), "sgl_ref_xi_grid")
    del gws
    return psi


@dataclass
class Result:
    xi: np.ndarray
    P: np.ndarray
    sigma: float
    mean: float
    var: float
    skew: float
    sigmaDL: float
    norm: float
    grid: dict
    pop: Population = None


def invert(pop: Population, xi_out: np.ndarray, tol: float = PHI_TOL,
           timings: dict = None) -> Result:
    """sgl_full.invert (sgl_full.py:159-375), Gaussian closure, no window."""
    L = _lib.load()
    dev = torch.device("cuda")
    clock = _Stages(timings)
    g = grid_plan(pop, tol)
    k_k = np.linspace(0.0, g["k_k_max"], g["nk_k"])
    k_g = np.linspace(0.0, g["k_g_max"], g["nk_g"])
    clock.mark("reach")
    phi = phi_grid(pop, k_k, k_g, g)
    clock.mark("phi")
    nk, ng, nkap, ngam = g["nk_k"], g["nk_g"], g["n_kap"], g["n_gam"]
    ws = torch.empty(L.sgl_ref_p3_bytes(nk, ng, nkap, ngam) // 8, dtype=torch.float64, device=dev)
    p3 = torch.empty((nkap, ngam), dtype=torch.float64, device=dev)
    st = _stream()
    check(L.sgl_ref_p3(_ptr(phi), nk, ng, g["k_k_max"], g["k_g_max"], g["kap_min"], g["kap_max"],
                       nkap, g["gam_max"], ngam, _ptr(ws), _ptr(p3), st), "sgl_ref_p3")
    clock.mark("p3")
    del phi, ws
    g_def = min(BOX_REACH * pop.sigma, g["gam_max"])
    n_wide = N_XI_WIDE_VP if pop.vp is not None else N_XI_WIDE
    xw = np.linspace(g["xi_min"], g["xi_max"], n_wide)
    xwd = torch.tensor(xw, dtype=torch.float64, device=dev)
    pw = torch.empty(n_wide, dtype=torch.float64, device=dev)
    check(L.sgl_ref_arcs(_ptr(p3), g["kap_min"], g["kap_max"], nkap, g["gam_max"], ngam, g_def,
                         _ptr(xwd), n_wide, None, _ptr(pw), st), "sgl_ref_arcs")
    mom = torch.empty(5, dtype=torch.float64, device=dev)
    check(L.sgl_ref_moments(_ptr(xwd), _ptr(pw), n_wide, _ptr(mom), st), "sgl_ref_moments")
    xo = np.ascontiguousarray(xi_out, dtype=np.float64)
    xod = torch.tensor(xo, dtype=torch.float64, device=dev)
    P = torch.empty(xo.size, dtype=torch.float64, device=dev)
    check(L.sgl_ref_arcs(_ptr(p3), g["kap_min"], g["kap_max"], nkap, g["gam_max"], ngam, g_def,
                         _ptr(xod), xo.size, _ptr(mom), _ptr(P), st), "sgl_ref_arcs")
    m = mom.cpu().numpy()
    clock.mark("arcs")
    return Result(xo, P.cpu().numpy(), pop.sigma, float(m[1]), float(m[2]), float(m[3]),
                  float(m[4]), float(m[0]), g, pop)


def run(zs: float, cosmo: dict, ell: bool = False, fil: bool = False, xi_out=None,
        edge: str = "r200", timings: dict = None, sub: bool = False,
        bias: bool = False, vp: bool = False, lss: bool = False,
        zint: str = "midpoint", zint_n: int = 2) -> Result:
    """One PDF on the reference's chain; xi_out None: the CLI body and tail
    grids linspace(-1, 1, 1601) and linspace(1.005, 9, 1600) (pipeline.py:211-215);
    vp: the positive clustering closure (needs bias); lss: the limber term
    in place of bias."""
    pop = build(cosmo, zs, ell=ell, fil=fil, edge=edge, timings=timings, sub=sub, bias=bias,
                vp=vp, lss=lss, zint=zint, zint_n=zint_n)
    if xi_out is None:
        xi_out = np.concatenate([np.linspace(-1.0, 1.0, 1601), np.linspace(1.005, 9.0, 1600)])
    return invert(pop, xi_out, timings=timings)
