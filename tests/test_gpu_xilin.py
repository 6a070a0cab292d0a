"""GPU parity of the full default chain (halos + ell + fil + sub + bias, the
xi_lin Limber two-point term with the Gaussian closure) against the
reference's own stored PDFs.

The reference's regression test (tests/test_reference.py:29-45) runs
run(zs, "full", xi_out=linspace(-1, 1, 1601), xi_tail=linspace(1.005, 9,
1600)) and requires max|dP| <= 1e-9 of the peak on the body and on the tail
against tests/reference/pdf_zs{0.5,1,3,10}.json.  The same call goes here
through the drop-in api.run on the device, with the same gate.

At z_s = 10 the tail box's gamma grid sits on a knife edge: n_gam =
ceil(6 sigma / (sigma / 36)) + 1 rounds 216 + 1 ulp up (217 stored), so
sigma must come out bit-close to the reference's; the device's sums run in a
fixed order (stable sorts, per-run sums in emission order), so it does.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_scalars.json")))
PDFS = np.load(os.path.join(HERE, "golden", "reference_pdfs.npz"))
BODY = np.linspace(-1.0, 1.0, 1601)
TAIL = np.linspace(1.005, 9.0, 1600)
TOL = 1e-9   # tests/test_reference.py:28


def _rel5(v, target):
    digits = 5 - 1 - int(np.floor(np.log10(abs(target))))
    return abs(v - target) <= 0.5 * 10.0 ** (-digits) * 1.0001


@pytest.mark.parametrize("zs,key", [(1.0, "zs1"), (0.5, "zs0p5"), (3.0, "zs3"), (10.0, "zs10")])
def test_full_default_matches_stored_reference(zs, key):
    from sgl_analytic_amd import api
    print(f"full default chain z_s={zs} on the device vs pdf_{key}", flush=True)
    p, info = api.run(zs, "full", cosmo={}, xi_out=np.concatenate([BODY, TAIL]))
    P0, T0 = PDFS[key + "_P_s"], PDFS[key + "_P_s_tail"]
    P1, T1 = p[:BODY.size], p[BODY.size:]
    d = np.max(np.abs(P1 - P0)) / P0.max()
    dt = np.max(np.abs(T1 - T0)) / P0.max()
    fx = GOLD["fixtures"]["pdf_zs%g" % zs]
    print(f"z_s={zs}: max|dP|/peak body {d:.2e}, tail {dt:.2e}; sigma_kappa "
          f"{info['sigma_kappa']:.10f} (ref {fx['sigma_kappa']:.10f}); grid "
          f"{[info['grid'][k] for k in ('nk_k', 'nk_g', 'n_kap', 'n_gam')]} (ref "
          f"{[fx['grid'][k] for k in ('nk_k', 'nk_g', 'n_kap', 'n_gam')]}); "
          f"{info['seconds']:.3f} s", flush=True)
    assert abs(info["sigma_kappa"] / fx["sigma_kappa"] - 1) < 1e-9
    for k in ("nk_k", "nk_g", "n_kap", "n_gam"):
        assert info["grid"][k] == fx["grid"][k], k
    assert d <= TOL and dt <= TOL
    for k in ("mean", "var", "skew", "sigmaDL"):
        assert abs(info[k] / fx[k] - 1) < 1e-8, k
    assert abs(info["N_expl"] / fx["N_expl"] - 1) < 1e-9
    assert abs(info["N_fil"] / fx["N_fil"] - 1) < 1e-9


def test_clustering_ladder_row():
    """The explorer's '+ clustering' row (halos + ell + fil + sub + bias at
    z_s = 1) at 5 significant figures."""
    from oracle import ref2d as oref
    from oracle.cosmo import PLANCK
    from sgl_analytic_amd import ref2d
    r = ref2d.run(1.0, PLANCK, ell=True, fil=True, sub=True, bias=True)
    row = oref.ladder_row(r)
    tgt = GOLD["explorer_ladder_zs1"]["+ clustering"]
    print("+ clustering", row, flush=True)
    for k in ("sigma_kappa", "clipped_var", "clipped_skew", "sigmaDL", "p_tail1"):
        assert _rel5(row[k], tgt[k]), (k, row[k], tgt[k])


def test_xi_grid_matches_axis_probes():
    """Psi_2 two ways on the device: the reach-probe path (sgl_ref_xi_axis,
    _Qxi_gauss evaluated directly) against the coarse-grid path
    (sgl_ref_xi_grid: batched FP64 MFMA contractions, then the not-a-knot
    bicubic spline) at the coarse nodes k = k_max (i / 100)^2, where the
    spline interpolates: a linspace of 10001 points holds every such node.
    And Psi_2(k, 0) -> -k^2 var / 2 at small k, var the clustering variance
    (Lam.sigma, sgl_full.py:3437-3446)."""
    from oracle.cosmo import PLANCK
    from sgl_analytic_amd import ref2d
    dpop = ref2d.build(PLANCK, 1.0, bias=True)
    hpop = ref2d.build(PLANCK, 1.0)
    kmax = 8.0 / dpop.sigma
    i = np.arange(101)
    nodes = kmax * (i / 100.0) ** 2
    for axis in (0, 1):
        psi_ax = ref2d.lam_axis(dpop, nodes, axis) - ref2d.lam_axis(hpop, nodes, axis)
        full = np.linspace(0.0, kmax, 10001)
        two = np.array([0.0, kmax])
        grid = ref2d.xi_psi(dpop, full, two) if axis == 0 else ref2d.xi_psi(dpop, two, full)
        g = grid.cpu().numpy()
        g = g[..., 0] + 1j * g[..., 1]
        on_nodes = g[i ** 2, 0] if axis == 0 else g[0, i ** 2]
        err = np.max(np.abs(on_nodes - psi_ax)) / np.max(np.abs(psi_ax))
        print(f"axis {axis}: max |grid - probe| / max |probe| = {err:.2e}", flush=True)
        assert err < 1e-9
    k = 1e-3 / dpop.sigma
    v = (ref2d.lam_axis(dpop, np.array([k]), 0) - ref2d.lam_axis(hpop, np.array([k]), 0))[0]
    assert abs(v.real / (-0.5 * k * k * dpop.xi.var) - 1) < 1e-4
