"""Truncated-NFW profile on the device (needs a GPU): the exact path's FP64
closed forms against the oracle's C restatement, and the batched path's FP32
form (no cancellation at x = 1, small-x series of the truncation term)
against the FP64 closed forms over the whole lattice range, x near 1 and the
truncation edge included."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

CS = (1.5, 2.0, 3.5, 5.0, 8.0, 14.0, 25.0)


def _dev(x, ks, xt, prec):
    from sgl_analytic_amd import _lib
    from sgl_analytic_amd.solver import _ptr, _stream
    L = _lib.load()
    xd = torch.as_tensor(x, dtype=torch.float64).cuda().contiguous()
    k = torch.empty_like(xd)
    g = torch.empty_like(xd)
    _lib.check(L.sgl_trunc_nfw_kg(_ptr(xd), xd.numel(), ks, xt, prec, _ptr(k), _ptr(g), _stream()),
               "sgl_trunc_nfw_kg")
    torch.cuda.synchronize()
    return k.cpu().numpy(), g.cpu().numpy()


def _grid(c):
    # lattice range (u in [1e-4, 10] r_vir, r_vir / r_s >= 2: x >= 2e-4), dense
    # around x = 1 and up to the truncation edge, plus points outside
    x = np.concatenate([np.geomspace(2e-4, 1.5 * c, 6000), 1.0 + np.linspace(-0.2, 0.2, 4001),
                        c * (1.0 - np.geomspace(1e-6, 0.1, 200))])
    return np.unique(x.astype(np.float32).astype(np.float64))


def test_fp64_profile_matches_oracle():
    from oracle import chain
    L = chain.lib()
    for c in CS:
        print(f"fp64 profile c={c}", flush=True)
        x = _grid(c)
        kd, gd = _dev(x, 0.37, c, 0)
        ko, go = np.empty_like(x), np.empty_like(x)
        L.oracle_trunc_kg(len(x), chain._p(x), 0.37, c, chain._p(ko), chain._p(go))
        # the same FP64 closed forms evaluated in a different order: they
        # differ by round-off, which the forms amplify where their terms
        # cancel (near x = 1 as 1 / |1 - x^2|^(3/2); in the truncation term of
        # gamma at small x as ln(2/x) / x^2, up to ~3e-9 at x = 2e-3 just
        # above the small-a series) and where kappa -> 0 at the truncation edge
        ek = np.abs(kd - ko) / (1e-9 * np.abs(ko) + 1e-15)
        eg = np.abs(gd - go) / (1e-8 * np.abs(go))
        print(f"  worst kappa at x={x[ek.argmax()]:.8g} ({ek.max():.2f}), gamma at "
              f"x={x[eg.argmax()]:.8g} ({eg.max():.2f})", flush=True)
        assert ek.max() <= 1.0 and eg.max() <= 1.0, c


def test_fp32_profile_vs_fp64():
    worst_k = worst_g = 0.0
    for c in CS:
        x = _grid(c)
        k64, g64 = _dev(x, 1.0, c, 0)
        k32, g32 = _dev(x, 1.0, c, 1)
        inside = x < c
        ek = np.abs(k32 - k64)[inside] / np.abs(k64[inside])
        eg = np.abs(g32 - g64) / np.abs(g64)
        print(f"fp32 profile c={c}: max rel err kappa {ek.max():.2e} gamma {eg.max():.2e}",
              flush=True)
        assert np.all(k32[~inside] == 0.0)
        worst_k, worst_g = max(worst_k, ek.max()), max(worst_g, eg.max())
    # hardware rcp / rsq / sqrt / log are ~1 ulp: a few 1e-6 at worst
    assert worst_k < 2e-5 and worst_g < 3e-5, (worst_k, worst_g)


def test_rejects_bad_arguments():
    from sgl_analytic_amd import _lib
    L = _lib.load()
    x = torch.ones(4, dtype=torch.float64, device="cuda")
    from sgl_analytic_amd.solver import _ptr
    assert L.sgl_trunc_nfw_kg(_ptr(x), 4, 1.0, -1.0, 0, _ptr(x), _ptr(x), None) == _lib.SGL_ERR_INVALID
    assert L.sgl_trunc_nfw_kg(_ptr(x), 0, 1.0, 3.0, 0, _ptr(x), _ptr(x), None) == _lib.SGL_ERR_INVALID
    assert L.sgl_trunc_nfw_kg(_ptr(x), 4, 1.0, 3.0, 2, _ptr(x), _ptr(x), None) == _lib.SGL_ERR_UNSUPPORTED
