"""Projected NFW pins: the oracle's closed forms vs independent quadrature
of the 3-D NFW density (mpmath), untruncated and truncated at x_t."""
import math

import mpmath as mp
import numpy as np
import pytest

from oracle import nfw
from oracle.chain import lib

mp.mp.dps = 20


def F_los(x, xt=None):
    """Sigma / (2 rho_s r_s) = int_0^zmax dz / (y (1+y)^2), y = sqrt(x^2+z^2)."""
    f = lambda z: 1 / (mp.sqrt(x * x + z * z) * (1 + mp.sqrt(x * x + z * z)) ** 2)
    if xt is None:
        return mp.quad(f, [0, 1, 10, mp.inf])
    zmax = mp.sqrt(xt * xt - x * x)
    return mp.quad(f, [0, zmax / 2, zmax])


def mu3d(y):
    return mp.log(1 + y) - y / (1 + y)


@pytest.mark.parametrize("x", [1e-5, 1e-3, 0.2, 0.9, 0.99, 1.0 + 1e-7, 1.3, 3.0, 40.0])
def test_F_against_line_of_sight(x):
    F, h = nfw.nfw_Fh(x)
    ref = F_los(mp.mpf(x))
    assert abs(F / float(ref) - 1) < 1e-13


@pytest.mark.parametrize("x", [1e-3, 0.95, 1.05, 20.0])
def test_h_is_enclosed_projected_mass(x):
    """h(x) = int_0^x F(x') x' dx' (so kbar = 4 ks h / x^2)."""
    F, h = nfw.nfw_Fh(x)
    ref = mp.quad(lambda u: F_los(u) * u, [0, min(x, 1) / 2, min(x, 1), x] if x > 1 else [0, x / 2, x])
    assert abs(h / float(ref) - 1) < 1e-10


@pytest.mark.parametrize("xt,a", [(4.85, 5e-4), (4.85, 0.5), (2.0, 0.999), (9.0, 0.1)])
def test_truncated_kappa_gamma(xt, a):
    x = a * xt
    k, g = nfw.nfw_trunc_kg_scalar(x, 1.0, xt)
    kref = 2 * F_los(mp.mpf(x), mp.mpf(xt))
    # enclosed projected mass of the truncated sphere within radius x
    menc = mp.quad(lambda u: 2 * F_los(u, mp.mpf(xt)) * u, [0, x / 2, x])
    kbar = 2 * menc / x ** 2
    assert abs(k - float(kref)) < 1e-10 * max(1.0, abs(float(kref)))
    assert abs((g + k) / float(kbar) - 1) < 1e-9


def test_outside_truncation_is_point_mass():
    xt = 4.0
    for x in (4.0, 5.0, 30.0):
        k, g = nfw.nfw_trunc_kg_scalar(x, 0.3, xt)
        assert k == 0.0
        assert abs(g / (0.3 * 4 * float(mu3d(xt)) / x ** 2) - 1) < 1e-14


def test_c_restatement_matches_python():
    """oracle/chain.c and oracle/nfw.py are two restatements of the same
    formulas; they must agree to rounding."""
    import ctypes
    L = lib()
    x = np.concatenate([np.geomspace(1e-6, 20, 400), [1 - 1e-6, 1 + 1e-6, 1.0]])
    k = np.zeros_like(x)
    g = np.zeros_like(x)
    P = ctypes.POINTER(ctypes.c_double)
    L.oracle_trunc_kg(x.size, x.ctypes.data_as(P), 0.2, 6.0, k.ctypes.data_as(P), g.ctypes.data_as(P))
    k2, g2 = nfw.nfw_trunc_kg(x, 0.2, 6.0)
    np.testing.assert_allclose(k, k2, rtol=1e-13, atol=1e-300)
    np.testing.assert_allclose(g, g2, rtol=1e-13)
