"""zint="engine" on the reference chain: the engine grid's shells below z_s,
one node at the upper edge with the backward difference as its width
(z_shells and the builder's dz=None rule, sgl_full.py:480-518, 1858-1859),
against the oracle on the CLI grids: halos at z_s = 1 and 3, the xi_lin
term ("+bias") and the limber term at z_s = 1, and the run() record.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import ref2d as oref  # noqa: E402
from oracle.cosmo import PLANCK  # noqa: E402

pytestmark = pytest.mark.gpu

XO = np.concatenate([oref.XI_BODY, oref.XI_TAIL])


@pytest.mark.parametrize("zs", [1.0, 3.0])
def test_engine_halo_pdf_matches_oracle(zs):
    from sgl_analytic_amd import ref2d
    d = ref2d.run(zs, PLANCK, xi_out=XO, zint="engine")
    o = oref.run(zs, xi_out=XO, zint="engine")
    m = ref2d.run(zs, PLANCK, xi_out=XO)
    l1 = np.trapezoid(np.abs(d.P - o.P), XO)
    print(f"engine halos z_s={zs}: bins {d.pop.n_bins} vs {o.pop.w.size}, sigma {d.sigma:.10f} vs "
          f"{o.sigma:.10f} (midpoint {m.sigma:.10f}), L1 {l1:.3e}", flush=True)
    assert d.pop.n_bins == o.pop.w.size and d.pop.n_points == o.pop.n_points
    assert abs(d.sigma / o.sigma - 1) < 1e-9
    assert l1 < 1e-8
    assert d.sigma < m.sigma        # the engine rule drops [0, z_0] and [z_last, z_s]


def test_engine_xilin_pdf_matches_oracle():
    from sgl_analytic_amd import api
    P, info = api.run(1.0, "+bias", cosmo={}, xi_out=XO, zint="engine")
    o = oref.run(1.0, xi_out=XO, bias=True, zint="engine")
    peak = np.max(o.P)
    err = np.max(np.abs(P - o.P)) / peak
    print(f"engine xi_lin z_s=1: sigma {info['sigma_kappa']:.10f} vs {o.sigma:.10f}, grid "
          f"{info['grid']['nk_k']}x{info['grid']['nk_g']} vs {o.grid['nk_k']}x{o.grid['nk_g']}, "
          f"max |dP| / peak {err:.3e}", flush=True)
    assert info["zint"] == "engine" and info["clustering"] == "xilin"
    assert abs(info["sigma_kappa"] / o.sigma - 1) < 1e-9
    assert err < 1e-9


def test_engine_limber_pdf_matches_oracle():
    from sgl_analytic_amd import api
    P, info = api.run(1.0, "+bias", cosmo={}, xi_out=XO, clustering="limber", zint="engine")
    o = oref.run(1.0, xi_out=XO, lss=True, zint="engine")
    l1 = np.trapezoid(np.abs(P - o.P), XO)
    print(f"engine limber z_s=1: L1 {l1:.3e}", flush=True)
    assert info["zint"] == "engine" and info["lss_var"] > 0
    assert l1 < 1e-8


@pytest.mark.parametrize("zs,n", [(1.0, 2), (0.5, 3)])
def test_gauss_halo_pdf_matches_oracle(zs, n):
    """zint="gauss": the midpoint shells at n Gauss-Legendre nodes each (the
    device's own Newton nodes against numpy's leggauss in the oracle); at
    z_s = 1 with 2 nodes 134 lens nodes, past the 104 midpoint shells."""
    from sgl_analytic_amd import api
    P, info = api.run(zs, "halo", cosmo={}, xi_out=XO, zint="gauss", zint_n=n)
    o = oref.run(zs, xi_out=XO, zint=f"gauss{n}")
    l1 = np.trapezoid(np.abs(P - o.P), XO)
    print(f"gauss{n} halos z_s={zs}: sigma {info['sigma_kappa']:.10f} vs {o.sigma:.10f}, "
          f"L1 {l1:.3e}", flush=True)
    assert info["zint"] == "gauss" and info["zint_n"] == n
    assert abs(info["sigma_kappa"] / o.sigma - 1) < 1e-9
    assert l1 < 1e-8


def test_gauss_xilin_pdf_matches_oracle():
    from sgl_analytic_amd import api
    P, info = api.run(1.0, "+bias", cosmo={}, xi_out=XO, zint="gauss", zint_n=2)
    o = oref.run(1.0, xi_out=XO, bias=True, zint="gauss2")
    err = np.max(np.abs(P - o.P)) / np.max(o.P)
    print(f"gauss2 xi_lin z_s=1: sigma {info['sigma_kappa']:.10f} vs {o.sigma:.10f}, "
          f"max |dP| / peak {err:.3e}", flush=True)
    assert abs(info["sigma_kappa"] / o.sigma - 1) < 1e-9
    assert err < 1e-9
