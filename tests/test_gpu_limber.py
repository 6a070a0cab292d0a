"""clustering="limber" on the reference chain (pipeline.py:126-134, 195-202)
against the oracle:

* sgl_ref_lss_var against oracle.ref2d.limber_lss_var (sgl_full.py:2541-2572)
  on flat ODE growth, legacy (Carroll-Press-Turner) growth and a curved w0
  cosmology;
* sgl_ref_lss_psi / sgl_ref_lss_axis against the oracle's lss_term
  (sgl_full.py:3104-3105);
* api.run(z_s, "+bias", clustering="limber") against the oracle's PDF on the
  CLI grids, and its run() record (clustering, lss_var, sigma_kappa);
* clustering="linear" without the bias ingredient is the default chain.
"""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

from oracle import ref2d as oref  # noqa: E402
from oracle.cosmo import Cosmology  # noqa: E402

pytestmark = pytest.mark.gpu

CASES = [(1.0, dict(growth_mode="ode")), (0.5, dict(growth_mode="legacy")),
         (2.0, dict(Ok=0.05, w0=-0.9, growth_mode="ode")), (3.0, dict(growth_mode="ode"))]


def _p(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


@pytest.fixture(scope="module")
def lib():
    from sgl_analytic_amd import _lib
    return _lib.load()


@pytest.mark.parametrize("zs,kw", CASES)
def test_lss_var_matches_oracle(lib, zs, kw):
    from sgl_analytic_amd import api, ref2d
    cp = api.check_cosmology(kw)
    rec = torch.tensor(ref2d._records(cp, zs, int(lib.sgl_ref_shell_count(zs))),
                       dtype=torch.float64, device="cuda")
    T = torch.empty(lib.sgl_ref_tables_bytes(0) // 8, dtype=torch.float64, device="cuda")
    status = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert lib.sgl_ref_tables_build(_p(rec), 0, _p(T), _p(status), None) == 0
    ws = torch.empty(lib.sgl_ref_lss_bytes() // 8, dtype=torch.float64, device="cuda")
    var = torch.empty(1, dtype=torch.float64, device="cuda")
    assert lib.sgl_ref_lss_var(_p(T), 0, _p(ws), _p(var), None) == 0
    torch.cuda.synchronize()
    assert int(status.item()) == 0
    got = float(var.item())
    want = oref.limber_lss_var(Cosmology(**kw), zs)
    print(f"z_s={zs} {kw}: device {got:.15e} oracle {want:.15e} rel {got / want - 1:.2e}",
          flush=True)
    assert abs(got / want - 1) < 1e-9


def test_lss_term_on_grid_and_axes(lib):
    v = 2.3e-4
    kk = np.linspace(0.0, 170.0, 37)
    kg = np.linspace(0.0, 95.0, 23)
    kkd = torch.tensor(kk, device="cuda")
    kgd = torch.tensor(kg, device="cuda")
    psi = torch.empty((kk.size, kg.size, 2), dtype=torch.float64, device="cuda")
    assert lib.sgl_ref_lss_psi(v, _p(kkd), kk.size, _p(kgd), kg.size, _p(psi), None) == 0
    torch.cuda.synchronize()
    p = psi.cpu().numpy()
    want = oref.lss_term(v, kk, kg)
    assert np.max(np.abs(p[..., 0] - want)) <= 1e-15 * np.max(np.abs(want))
    assert np.all(p[..., 1] == 0.0)
    for axis, k in ((0, kk), (1, kg)):
        kd = torch.tensor(k, device="cuda")
        out = torch.zeros(2 * k.size, dtype=torch.float64, device="cuda")
        assert lib.sgl_ref_lss_axis(v, _p(kd), k.size, axis, _p(out), None) == 0
        torch.cuda.synchronize()
        o = out.cpu().numpy()
        w = want[:, 0] if axis == 0 else want[0, :]
        assert np.max(np.abs(o[0::2] - w)) <= 1e-15 * np.max(np.abs(w))
        assert np.all(o[1::2] == 0.0)


@pytest.mark.parametrize("zs", [1.0, 3.0])
def test_limber_pdf_matches_oracle(zs):
    from sgl_analytic_amd import api
    xo = np.concatenate([oref.XI_BODY, oref.XI_TAIL])
    P, info = api.run(zs, "+bias", cosmo={}, xi_out=xo, clustering="limber")
    o = oref.run(zs, xi_out=xo, lss=True)
    l1 = np.trapezoid(np.abs(P - o.P), xo)
    print(f"limber z_s={zs}: lss_var {info['lss_var']:.6e} vs {o.pop.lss_var:.6e}, "
          f"sigma {info['sigma_kappa']:.10f} vs {o.sigma:.10f}, grid "
          f"{info['grid']['nk_k']}x{info['grid']['nk_g']} vs {o.grid['nk_k']}x{o.grid['nk_g']}, "
          f"L1 {l1:.3e}", flush=True)
    assert info["clustering"] == "limber" and info["xi_kperp"] is None
    assert abs(info["lss_var"] / o.pop.lss_var - 1) < 1e-9
    assert abs(info["sigma_kappa"] / o.sigma - 1) < 1e-9
    assert info["grid"]["nk_k"] == o.grid["nk_k"] and info["grid"]["n_kap"] == o.grid["n_kap"]
    assert l1 < 1e-8


def test_linear_without_bias_is_the_default_chain():
    from sgl_analytic_amd import api
    xo = np.linspace(-0.5, 2.0, 301)
    P0, i0 = api.run(1.0, "halo", cosmo={}, xi_out=xo)
    for mode in ("linear", "limber"):
        P1, i1 = api.run(1.0, "halo", cosmo={}, xi_out=xo, clustering=mode)
        assert np.array_equal(P0, P1), mode
        assert i1["clustering"] == mode and i1["lss_var"] == 0.0
        assert i1["sigma_kappa"] == i0["sigma_kappa"]
