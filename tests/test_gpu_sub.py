"""GPU parity of the subhalo ingredient on the reference-shape path.

* the per-host tables (clump law, projected clump density, clump cells, the
  largest host radius within reach) against the oracle's SubhaloModel.host /
  clump_cells at z_s = 1;
* the coarse point tables (K offset, removed mass, weight, shear terms, the
  host-mass table) against the oracle's host_cells_moments restatement;
* the halos + subhalos population at z_s = 1 against the oracle's bins;
* the explorer's "+ subhalos" ladder row (5 significant figures) on the device;
* the halos + subhalos PDF against the oracle's on the same grid.
"""
import json
import os

from sgl_analytic_amd import _lib  # noqa: E402

import numpy as np
import pytest
import torch

from oracle import ref2d as oref
from oracle import sub as osub
from oracle.cosmo import Cosmology, PLANCK

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LADDER = json.load(open(os.path.join(HERE, "golden", "reference_scalars.json")))["explorer_ladder_zs1"]

# device layout of the subhalo workspace (sgl_analytic_amd/csrc/ref_sub.h)
NPSI, ND, NR, NC, NQ, NMT = 28, 90, 256, 128, 15, 18
HR = dict(valid=0, f=1, eta=2, Mvir=3, xsub=4, Dmax=5, rs=6, M=7, xref=8, nq=9, ms=10, kstar=11,
          nmt=12, xt=13, r200=14)
HR_DN = 32
HR_MC, HR_D, HR_RSC, HR_KSC, HR_CC = (HR_DN + NPSI * k for k in range(1, 6))
HR_N2 = HR_CC + NPSI
HR_DC = HR_N2 + NR
HR_DA = HR_DC + ND
HR_CK = HR_DA + ND
HR_CG = HR_CK + NPSI * ND
HOST_STRIDE = (HR_CG + NPSI * ND + 31) // 32 * 32
CM_K = NC
CM_S, CM_W, CM_A1, CM_A2 = (CM_K + NC * NQ * k for k in range(1, 5))
CM_LMT = CM_A2 + NC * NQ
CM_LRS, CM_LKS, CM_EPT = CM_LMT + NMT, CM_LMT + 2 * NMT, CM_LMT + 3 * NMT
COARSE_STRIDE = (CM_EPT + NMT + 31) // 32 * 32
ROWS = _lib.REF_MAX_SHELLS * _lib.REF_MASSES   # the table capacity, not the shell count
TASK_DOUBLES = (ROWS * 12 * 4 // 2 + 31) // 32 * 32   # per-task counts after the host records


@pytest.fixture(scope="module")
def dev_sub():
    from sgl_analytic_amd import ref2d
    pop = ref2d.build(PLANCK, 1.0, sub=True, keep_sub=True)
    w = pop.sub_tables.cpu().numpy()
    host = w[:ROWS * HOST_STRIDE].reshape(ROWS, HOST_STRIDE)
    c0 = ROWS * HOST_STRIDE + TASK_DOUBLES
    coarse = w[c0:c0 + ROWS * COARSE_STRIDE].reshape(ROWS, COARSE_STRIDE)
    return pop, host, coarse


@pytest.fixture(scope="module")
def oracle_model():
    cos = Cosmology()
    zs = 1.0
    zl = np.array([s[2] for s in oref.z_shells(zs)])
    return cos, osub.SubhaloModel(cos, zs, zl, oref.engine_grids()[1])


def _rel(a, b, floor=0.0):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))) if a.size else 0.0


HOSTS = [(0, 40), (20, 60), (40, 80), (40, 98), (60, 30), (66, 90)]


def test_host_count(dev_sub, oracle_model):
    _, host, _ = dev_sub
    cos, model = oracle_model
    n_dev = int(np.sum(host[:, HR["valid"]] != 0))
    n_or = sum(model.host(i, j) is not None for i in range(67) for j in range(1, 100))
    print(f"hosts with subhalos: device {n_dev}, oracle {n_or}", flush=True)
    assert n_dev == n_or


@pytest.mark.parametrize("jz,jM", HOSTS)
def test_host_tables(dev_sub, oracle_model, jz, jM):
    _, host, _ = dev_sub
    cos, model = oracle_model
    h = model.host(jz, jM)
    H = host[jz * 99 + jM - 1]
    assert (H[HR["valid"]] != 0) == (h is not None)
    if h is None:
        return
    checks = {
        "f": (H[HR["f"]], h["f"]), "eta": (H[HR["eta"]], h["eta"]),
        "Mvir": (H[HR["Mvir"]], h["Mvir"]), "dN": (H[HR_DN:HR_DN + NPSI], h["dN"]),
        "m": (H[HR_MC:HR_MC + NPSI], h["m"]), "D": (H[HR_D:HR_D + NPSI], h["D"]),
        "ksc": (H[HR_KSC:HR_KSC + NPSI], h["ksc"]), "n2": (H[HR_N2:HR_N2 + NR], h["n2"]),
    }
    _, dc, dA = model.dgrid(h)
    checks["dc"] = (H[HR_DC:HR_DC + ND], dc)
    checks["dA"] = (H[HR_DA:HR_DA + ND], dA)
    cc = model.clump_cells(h)
    ck = H[HR_CK:HR_CK + NPSI * ND].reshape(NPSI, ND)[cc["i"], cc["d"]]
    cg = H[HR_CG:HR_CG + NPSI * ND].reshape(NPSI, ND)[cc["i"], cc["d"]]
    checks["ck"] = (ck, cc["k"])
    checks["cg"] = (cg, cc["g"])
    for k, (a, b) in checks.items():
        r = _rel(a, b, 1e-300)
        print(f"host ({jz}, {jM}) {k}: max rel {r:.2e}", flush=True)
        assert r < 1e-8, k


@pytest.mark.parametrize("jz,jM", HOSTS)
def test_coarse_tables(dev_sub, oracle_model, jz, jM):
    _, host, coarse = dev_sub
    cos, model = oracle_model
    h = model.host(jz, jM)
    if h is None:
        pytest.skip("no subhalos on this host")
    H = host[jz * 99 + jM - 1]
    Cm = coarse[jz * 99 + jM - 1]
    r = [q for q in oref.rows(cos, 1.0, shells={jz}) if q["jM"] == jM][0]
    grid = oref.host_grid(r["ks"], r["xt"], r["xm"], r["Nbar"])
    x_sub = (r["xt"] * h["rs"] + h["D"].max()) / h["rs"] * 1.001
    xref = np.concatenate([grid["xc"], grid["xo"]])
    xref = xref[xref <= x_sub]
    # the host rows agree with the oracle's to ~1e-10 (tables), so do their grids
    assert abs(H[HR["xref"]] / xref.max() - 1) < 1e-9
    cm = model.coarse(h, xref)
    nq = cm["nq"]
    assert int(H[HR["nq"]]) == nq
    print(f"host ({jz}, {jM}): kstar dev {H[HR['kstar']]:.6e} oracle {cm['kstar']:.6e}, "
          f"ms {H[HR['ms']]:.6e} {cm['ms']:.6e}", flush=True)
    assert abs(H[HR["kstar"]] - cm["kstar"]) <= 1e-12 * abs(cm["kstar"])
    assert abs(H[HR["ms"]] / cm["ms"] - 1) < 1e-8
    assert np.max(np.abs(Cm[:NC] - cm["lxg"])) < 1e-9
    for name, off in (("Kp_c", CM_K), ("Sp_c", CM_S), ("wp_c", CM_W), ("A1_c", CM_A1),
                      ("A2_c", CM_A2)):
        d = Cm[off:off + NC * NQ].reshape(NC, NQ)[:, :nq]
        o = cm[name]
        scale = np.max(np.abs(o))
        err = np.max(np.abs(d - o)) / scale
        print(f"host ({jz}, {jM}) {name}: max |d - o| / max|o| = {err:.2e}", flush=True)
        assert err < 1e-7, name
    nmt = int(H[HR["nmt"]])
    assert nmt == cm["Mt"].size
    assert _rel(Cm[CM_LMT:CM_LMT + nmt], cm["lMt"]) < 1e-12
    assert _rel(Cm[CM_LRS:CM_LRS + nmt], cm["log_rs"], 1.0) < 1e-9
    assert _rel(Cm[CM_LKS:CM_LKS + nmt], cm["log_ks"], 1.0) < 1e-9


def _bin_index(k, g):
    """Fine-lattice bin of a bin centroid (a centroid lies inside its bin)."""
    ek, eg = oref.fine_edges()
    return (np.searchsorted(ek, k, side="right") - 1) * (eg.size - 1) + (
        np.searchsorted(eg, g, side="right") - 1)


def test_population_halo_sub_zs1(dev_sub):
    """Same point count; bins compared by lattice index: the bins present on
    one side only carry a negligible weight (points of ~1e-65 weight whose
    kappa rounds across a bin edge), the bins that carry the weight agree to 1e-6."""
    from sgl_analytic_amd._lib import REF_FINE_BINS
    dpop = dev_sub[0]
    opop = oref.build(Cosmology(), 1.0, sub=True)
    print(f"halo+sub: device bins {dpop.n_bins} points {dpop.n_points}; oracle {opop.w.size} "
          f"{opop.n_points}", flush=True)
    assert dpop.n_points == opop.n_points
    p = dpop.pop.cpu().numpy()
    n = dpop.n_bins
    w, k, g = p[:n], p[REF_FINE_BINS:REF_FINE_BINS + n], p[2 * REF_FINE_BINS:2 * REF_FINE_BINS + n]
    bd = _bin_index(k, g)
    bo = np.asarray(opop.bins)
    assert np.all(np.diff(bd) > 0)
    common, id_, io = np.intersect1d(bd, bo, return_indices=True)
    only_d = np.setdiff1d(np.arange(n), id_)
    only_o = np.setdiff1d(np.arange(bo.size), io)
    wtot = opop.w.sum()
    print(f"bins only on device {only_d.size} (weight {w[only_d].sum() / wtot:.2e}), only in the "
          f"oracle {only_o.size} (weight {opop.w[only_o].sum() / wtot:.2e})", flush=True)
    assert only_d.size + only_o.size <= 5e-3 * bo.size
    assert w[only_d].sum() < 1e-15 * wtot and opop.w[only_o].sum() < 1e-15 * wtot
    # a decorated point's kappa comes out of Gauss rules whose nodes move by up
    # to ~1e-7 when the moments agree to ~1e-12 (near-degenerate cubics), so
    # points of tiny weight lying that close to a bin edge fall on the other
    # side: bins above 1e-9 of the weight agree bin by bin, and the weight
    # that changes bins is below 1e-9 of the total
    big = opop.w[io] > 1e-9 * wtot
    dw = np.abs(w[id_] - opop.w[io]).sum() + w[only_d].sum() + opop.w[only_o].sum()
    ew = np.max(np.abs(w[id_][big] / opop.w[io][big] - 1))
    ek = np.max(np.abs(k[id_][big] - opop.k[io][big]) / (np.abs(opop.k[io][big]) + 1e-12))
    eg = np.max(np.abs(g[id_][big] / opop.g[io][big] - 1))
    print(f"bins above 1e-9 of the weight: {big.sum()} of {io.size}; w {ew:.2e} kappa {ek:.2e} "
          f"gamma {eg:.2e}; total |dw| / w {dw / wtot:.2e}", flush=True)
    assert ew < 1e-6 and ek < 1e-6 and eg < 1e-6
    assert dw < 1e-9 * wtot
    assert abs(dpop.sigma / opop.sigma() - 1) < 1e-7


def test_ladder_row_subhalos_on_device():
    from sgl_analytic_amd import ref2d
    r = ref2d.run(1.0, PLANCK, ell=True, fil=True, sub=True)
    row = oref.ladder_row(r)
    tgt = LADDER["+ subhalos"]
    print("+ subhalos", row, flush=True)
    for key in ("sigma_kappa", "clipped_var", "clipped_skew", "sigmaDL", "p_tail1"):
        digits = 5 - 1 - int(np.floor(np.log10(abs(tgt[key]))))
        assert abs(row[key] - tgt[key]) <= 0.5 * 10.0 ** (-digits) * 1.0001, key


def test_halo_sub_pdf_matches_oracle():
    from sgl_analytic_amd import ref2d
    xo = np.concatenate([oref.XI_BODY, oref.XI_TAIL])
    d = ref2d.run(1.0, PLANCK, xi_out=xo, sub=True)
    o = oref.run(1.0, xi_out=xo, sub=True)
    l1 = np.trapezoid(np.abs(d.P - o.P), xo)
    print(f"halo+sub z_s=1: bins {d.pop.n_bins} vs {o.pop.w.size}, grid "
          f"{d.grid['nk_k']}x{d.grid['nk_g']} vs {o.grid['nk_k']}x{o.grid['nk_g']}, L1 {l1:.3e}",
          flush=True)
    assert l1 < 1e-6
