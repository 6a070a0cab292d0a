"""GPU parity of the reference-shape 2-D chain (sgl_ref_* entry points).

* sgl_ref_p3 (Hankel + Fourier FP64 MFMA contractions and the
  Euler-Maclaurin epilogue) against the oracle's P_vector_joint restatement
  on the same Gaussian Phi, on grids that are not multiples of the tiles;
* sgl_ref_phi (Lambda contraction with the exp epilogue) and
  sgl_ref_lam_axis against the oracle's Lambda on a random bin table;
* the device population at z_s = 1 (halos) against the oracle's bins;
* generate-level runs at z_s = 1 against the explorer ladder rows (5
  significant figures) and against the oracle's PDF on the same grid.
"""
import ctypes
import json
import os

import numpy as np
import pytest
import torch

from oracle import ref2d as oref
from oracle.cosmo import Cosmology, PLANCK

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LADDER = json.load(open(os.path.join(HERE, "golden", "reference_scalars.json")))["explorer_ladder_zs1"]


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def lib():
    from sgl_analytic_amd import _lib
    assert torch.cuda.is_available()
    return _lib.load()


def _pop_tensor(w, k, g):
    from sgl_analytic_amd._lib import REF_FINE_BINS
    pop = torch.zeros(3 * REF_FINE_BINS, dtype=torch.float64, device="cuda")
    n = w.size
    pop[:n] = torch.tensor(w)
    pop[REF_FINE_BINS:REF_FINE_BINS + n] = torch.tensor(k)
    pop[2 * REF_FINE_BINS:2 * REF_FINE_BINS + n] = torch.tensor(g)
    return pop, torch.tensor([n], dtype=torch.int32, device="cuda")


def test_gemm_f64_layout(lib):
    """The FP64 MFMA contraction against numpy on shapes that are not tile
    multiples; on a mismatch, report where each entry of the first 16 x 16
    block landed (the lane layout of the MFMA result)."""
    rng = np.random.default_rng(5)
    for M, N, K in ((64, 64, 16), (131, 77, 45), (600, 257, 1000)):
        A = rng.normal(size=(M, K))
        B = rng.normal(size=(K, N))
        Ad, Bd = torch.tensor(A, device="cuda"), torch.tensor(B, device="cuda")
        Cd = torch.zeros((M, N), dtype=torch.float64, device="cuda")
        assert lib.sgl_ref_gemm_f64(M, N, K, _p(Ad), K, _p(Bd), N, _p(Cd), N, None) == 0
        torch.cuda.synchronize()
        C = Cd.cpu().numpy()
        ref = A @ B
        err = np.max(np.abs(C - ref)) / np.max(np.abs(ref))
        if err > 1e-13 and M == 64:
            where = {}
            for i in range(16):
                for j in range(16):
                    hit = np.argwhere(np.abs(ref - C[i, j]) < 1e-9 * np.max(np.abs(ref)))
                    where[(i, j)] = [tuple(h) for h in hit[:2]]
            print("device (i, j) holds true entries:", {k: v for k, v in list(where.items())[:40]},
                  flush=True)
        assert err < 1e-13, (M, N, K, err)


def test_p3_gaussian_matches_oracle(lib):
    sk, sg = 0.03, 0.025
    nk, ng, nkap, ngam = 301, 203, 157, 119
    k_k = np.linspace(0.0, 9.0 / sk, nk)
    k_g = np.linspace(0.0, 9.0 / sg, ng)
    kap = np.linspace(-0.2, 0.99, nkap)
    gam = np.linspace(0.0, 0.8, ngam)
    Phi = np.exp(-0.5 * sk ** 2 * k_k[:, None] ** 2 - 0.5 * sg ** 2 * k_g[None, :] ** 2) * (
        1 + 0.1j * np.sin(k_k[:, None] * 0.01))
    ref = oref.p3_joint(k_k, k_g, Phi, kap, gam)
    pd = torch.empty((2 * nk, ng), dtype=torch.float64, device="cuda")
    pd[0::2] = torch.tensor(Phi.real)
    pd[1::2] = torch.tensor(Phi.imag)
    ws = torch.empty(lib.sgl_ref_p3_bytes(nk, ng, nkap, ngam) // 8, dtype=torch.float64,
                     device="cuda")
    out = torch.empty((nkap, ngam), dtype=torch.float64, device="cuda")
    rc = lib.sgl_ref_p3(_p(pd), nk, ng, k_k[-1], k_g[-1], kap[0], kap[-1], nkap, gam[-1], ngam,
                        _p(ws), _p(out), None)
    assert rc == 0
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.max(np.abs(o - ref)) < 1e-12 * np.max(np.abs(ref))


def test_phi_and_axis_match_oracle(lib):
    rng = np.random.default_rng(3)
    n = 777
    w = rng.uniform(0.1, 40.0, n)
    k = rng.normal(0, 0.03, n) * rng.choice([1e-5, 1e-2, 1.0, 5.0], n)
    g = np.abs(rng.normal(0, 0.05, n)) * rng.choice([1e-6, 1.0], n) + 1e-9
    pop = oref.Population(w=w, k=k, g=g)
    pdev, npop = _pop_tensor(w, k, g)
    kk = np.linspace(0.0, 300.0, 67)
    kg = np.linspace(0.0, 250.0, 45)
    ref = np.exp(oref.lam(pop, kk, kg))
    kkd = torch.tensor(kk, device="cuda")
    kgd = torch.tensor(kg, device="cuda")
    ws = torch.empty(lib.sgl_ref_phi_bytes(kk.size, kg.size, n) // 8, dtype=torch.float64,
                     device="cuda")
    phi = torch.empty((2 * kk.size, kg.size), dtype=torch.float64, device="cuda")
    assert lib.sgl_ref_phi(_p(pdev), _p(npop), n, _p(kkd), kk.size, _p(kgd), kg.size, None, _p(ws),
                           _p(phi), None) == 0
    torch.cuda.synchronize()
    ph = phi.cpu().numpy()
    got = ph[0::2] + 1j * ph[1::2]
    assert np.max(np.abs(got - ref)) < 1e-12 * max(1.0, np.max(np.abs(ref)))
    for axis in (0, 1):
        probe = np.array([0.0, 0.37, 12.0, 401.0, 5000.0])
        pd = torch.tensor(probe, device="cuda")
        out = torch.empty(2 * probe.size, dtype=torch.float64, device="cuda")
        assert lib.sgl_ref_lam_axis(_p(pdev), _p(npop), _p(pd), probe.size, axis, _p(out), None) == 0
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        got = o[0::2] + 1j * o[1::2]
        want = oref.lam(pop, probe, [0.0])[:, 0] if axis == 0 else oref.lam(pop, [0.0], probe)[0]
        assert np.max(np.abs(got - want)) < 1e-11 * w.sum()


@pytest.fixture(scope="module")
def halo_pops():
    from sgl_analytic_amd import ref2d
    dpop = ref2d.build(PLANCK, 1.0)
    opop = oref.build(Cosmology(), 1.0)
    return dpop, opop


def test_population_halo_zs1(halo_pops):
    from sgl_analytic_amd._lib import REF_FINE_BINS
    dpop, opop = halo_pops
    print(f"device bins {dpop.n_bins} points {dpop.n_points}; oracle {opop.w.size} {opop.n_points}",
          flush=True)
    assert dpop.n_points == opop.n_points
    assert dpop.n_bins == opop.w.size
    p = dpop.pop.cpu().numpy()
    n = dpop.n_bins
    w, k, g = p[:n], p[REF_FINE_BINS:REF_FINE_BINS + n], p[2 * REF_FINE_BINS:2 * REF_FINE_BINS + n]
    assert np.max(np.abs(w / opop.w - 1)) < 1e-6
    assert np.max(np.abs(k - opop.k) / (np.abs(opop.k) + 1e-12)) < 1e-6
    assert np.max(np.abs(g / opop.g - 1)) < 1e-6
    assert abs(dpop.sigma / opop.sigma() - 1) < 1e-7
    assert abs(dpop.N_expl / sum(r["Nbar"] for r in oref.rows(Cosmology(), 1.0)) - 1) < 1e-6


@pytest.mark.parametrize("model,kw", [("halos", {}), ("+ ellipticity", dict(ell=True)),
                                      ("+ filaments", dict(ell=True, fil=True))])
def test_ladder_rows_on_device(model, kw):
    from sgl_analytic_amd import ref2d
    r = ref2d.run(1.0, PLANCK, **kw)
    row = oref.ladder_row(r)
    tgt = LADDER[model]
    print(model, row, flush=True)
    for key in ("sigma_kappa", "clipped_var", "clipped_skew", "sigmaDL", "p_tail1"):
        digits = 5 - 1 - int(np.floor(np.log10(abs(tgt[key]))))
        assert abs(row[key] - tgt[key]) <= 0.5 * 10.0 ** (-digits) * 1.0001, (model, key)


def test_halo_pdf_matches_oracle(halo_pops):
    from sgl_analytic_amd import ref2d
    dpop, opop = halo_pops
    xo = np.concatenate([oref.XI_BODY, oref.XI_TAIL])
    d = ref2d.invert(dpop, xo)
    o = oref.invert(lambda a, b: oref.lam(opop, a, b), opop.sigma(), xo)
    assert d.grid["nk_k"] == o.grid["nk_k"] and d.grid["n_kap"] == o.grid["n_kap"]
    l1 = np.trapezoid(np.abs(d.P - o.P), xo)
    print(f"halo z_s=1 device vs oracle: L1 {l1:.3e}, max {np.max(np.abs(d.P - o.P)):.3e}", flush=True)
    assert l1 < 1e-6


@pytest.mark.parametrize("zs,edge", [(0.5, "r200"), (1.0, "rvir"), (3.0, "r200"), (10.0, "r200")])
def test_halo_pdf_other_shapes(zs, edge):
    """Device vs oracle on the reference grids at other source redshifts
    (their own shell counts and k grids) and with the r_vir truncation."""
    from sgl_analytic_amd import ref2d
    xo = np.concatenate([oref.XI_BODY, oref.XI_TAIL])
    d = ref2d.run(zs, PLANCK, xi_out=xo, edge=edge)
    o = oref.run(zs, xi_out=xo, edge=edge)
    l1 = np.trapezoid(np.abs(d.P - o.P), xo)
    print(f"halo z_s={zs} edge={edge}: bins {d.pop.n_bins} vs {o.pop.w.size}, grid "
          f"{d.grid['nk_k']}x{d.grid['nk_g']} vs {o.grid['nk_k']}x{o.grid['nk_g']}, "
          f"L1 {l1:.3e}", flush=True)
    assert d.pop.n_bins == o.pop.w.size
    assert l1 < 1e-6
