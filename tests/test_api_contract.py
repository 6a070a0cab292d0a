"""Public input contracts of the drop-in API without device solves (mirrors
the reference's tests/test_api.py with the solver mocked out)."""
import warnings
from unittest.mock import patch

import numpy as np
import pytest

from sgl_analytic_amd import api, generate_pdf, generate_pdf_lnmu


def fake_info(x):
    return dict(kernel_grid=dict(xi_min=-0.5, xi_max=3.0))


def test_invalid_configuration():
    for cfg in ("", "garbage", "+typo", "-typo"):
        with pytest.raises(ValueError):
            api.on(cfg)
    assert api.on("halo") == set()
    assert api.on("+bias") == {"bias"}
    assert api.on("-sub") == {"ell", "fil", "bias"}


def test_invalid_grid_fails_before_solver():
    with patch("sgl_analytic_amd.api.run") as solver:
        for grid in ([0, 0], [1, 0], [0, np.nan], [[0, 1]], [0]):
            with pytest.raises(ValueError):
                generate_pdf_lnmu(1, lnmu=grid)
        with pytest.raises(ValueError):
            generate_pdf(1, mu=[0, 1])
        with pytest.raises(ValueError):
            generate_pdf_lnmu(1, n_lnmu=1)
        solver.assert_not_called()


def test_cosmology_validation():
    for cp in (dict(h=0), dict(Om=np.nan), dict(Ob=0.4), dict(zeq=-1), dict(sigma8=.8, s8=.8),
               dict(Ok=np.nan), dict(Ok=True), dict(w0=0.0), dict(wa=np.inf), dict(wa=1.2),
               dict(w0=-.5, wa=.45), dict(growth_mode="exact"), dict(Ok=-1.5),
               dict(w0=-.9, growth_mode="legacy"), dict(radiation="yes"), dict(foo=1)):
        with pytest.raises(ValueError):
            api.check_cosmology(cp)
    cp = api.check_cosmology(dict(Ok=-.05, w0=-1.1, wa=.2))
    assert (cp["Ok"], cp["w0"], cp["wa"]) == (-.05, -1.1, .2)


def test_paths_not_computed_here_raise():
    with patch("sgl_analytic_amd.api._solver") as solver:
        with pytest.raises(ValueError):
            api.run(1.0, "full", cosmo={}, xi_out=np.array([0.0, 1.0]), closure="lognormal")
        with pytest.raises(NotImplementedError):    # subhalos pair with the r200 edge
            api.run(1.0, "+sub", cosmo={}, xi_out=np.array([0.0, 1.0]), edge="rvir")
        with pytest.raises(ValueError):     # vp needs the bias ingredient, as in the reference
            api.run(1.0, "-bias", cosmo={}, xi_out=np.array([0.0, 1.0]), closure="vp")
        with pytest.raises(NotImplementedError):    # the threshold edge is not built
            api.run(1.0, "-bias", cosmo={}, xi_out=np.array([0.0, 1.0]), edge="kthr")
        with pytest.raises(ValueError):
            api.run(1.0, "full", cosmo={}, xi_out=np.array([0.0, 1.0]), edge="r500")
        solver.assert_not_called()


def test_run_checks_arguments_as_the_reference_does():
    """pipeline.run's argument errors (pipeline.py:139-164, its own
    test_invalid_redshift_and_workers_fail_before_build) raise the same
    ValueError before any device work; the modes that are not built raise
    NotImplementedError, also before any device work."""
    from sgl_analytic_amd import ref2d
    with patch.object(ref2d, "build") as build:
        for zs in (0, -1, np.nan, np.inf):
            with pytest.raises(ValueError):
                api.run(zs, "full")
        for nproc in (0, -1, 1.5, True):
            with pytest.raises(ValueError):
                api.run(1, "full", nproc=nproc)
        for kw in (dict(edge="r500"), dict(closure="lognormal"),
                   dict(closure="vp", clustering="linear"), dict(edge="kthr"),
                   dict(clustering="halofit"), dict(zint="trapezoid"),
                   dict(zint="gauss", zint_n=0), dict(closure="vp", clustering="limber")):
            with pytest.raises(ValueError):
                api.run(1, "full", **kw)
        with pytest.raises(ValueError):     # the vp closure needs the bias ingredient
            api.run(1, "-bias", closure="vp")
        for cfg, kw in (("full", dict(clustering="linear")), ("full", dict(clustering="closure")),
                        ("-bias", dict(clustering="closure")), ("-bias", dict(edge="kthr")),
                        ("full", dict(zint="engine")), ("+sub", dict(zint="engine")),
                        ("+bias", dict(zint="engine", closure="vp")),
                        ("full", dict(zint="gauss", zint_n=3)),
                        ("+bias", dict(zint="gauss", closure="vp"))):
            with pytest.raises(NotImplementedError):
                api.run(1, cfg, **kw)
        build.assert_not_called()


def test_vp_closure_is_an_exact_path_option():
    """The positive closure runs on the plain (exact) path and needs the slice
    covariance; the solver rejects it elsewhere before touching a device."""
    from sgl_analytic_amd.solver import Shape, Solver
    with pytest.raises(ValueError):
        Solver(Shape(), 1, path="binned", closure="vp")
    with pytest.raises(ValueError):
        Solver(Shape(clustering=False), 1, path="exact", closure="vp")
    with pytest.raises(ValueError):
        Solver(Shape(), 1, path="exact", closure="lognormal")


def test_growth_and_radiation_options():
    """growth_mode "auto" resolves as sgl.Cosmology does (legacy for LambdaCDM,
    the ODE otherwise, with the pipeline's warning), radiation "none" is None,
    and both reach the device record slots."""
    from sgl_analytic_amd import _lib
    from sgl_analytic_amd.solver import cosmo_records
    assert api.check_cosmology(dict(growth_mode="auto"))["growth_mode"] == "legacy"
    with pytest.warns(UserWarning):
        assert api.check_cosmology(dict(growth_mode="auto", w0=-0.9))["growth_mode"] == "ode"
    assert api.check_cosmology(dict(radiation="none"))["radiation"] is None
    p = np.array([[0.674, 0.315, 0.811, 0.0493, 0.965, 3402.0, 0.0, -1.0, 0.0, 1.0]])
    r = cosmo_records(p, 8)
    assert r[0, _lib.REC_RADIATION] == 0.0 and r[0, _lib.REC_GROWTH] == 0.0
    r = cosmo_records(p, 8, radiation=None, growth_mode="legacy")
    assert r[0, _lib.REC_RADIATION] == 1.0 and r[0, _lib.REC_GROWTH] == 1.0
    q = p.copy()
    q[0, 7] = -0.9
    with pytest.raises(ValueError):
        cosmo_records(q, 8, growth_mode="legacy")
    with pytest.raises(ValueError):
        cosmo_records(p, 8, growth_mode="auto")


def test_invalid_redshift_and_workers_fail_before_build():
    with patch("sgl_analytic_amd.api._solver") as solver:
        for zs in (0, -1, np.nan, np.inf):
            with pytest.raises(ValueError):
                api.run(zs, "full", xi_out=np.array([0.0, 1.0]))
        for nproc in (0, -1, 1.5, True):
            with pytest.raises(ValueError):
                generate_pdf_lnmu(1, nproc=nproc)
        solver.assert_not_called()


def test_parameter_forwarding_and_jacobian():
    def fake_run(zs, cfg, **kw):
        assert zs == 2
        assert kw["cosmo"] == dict(h=.7, Om=.3, sigma8=.9, Ob=.05, ns=.97, zeq=3450, Ok=0.0,
                                   w0=-1.0, wa=0.0, growth_mode="ode", radiation="engine")
        assert kw["closure"] == "gauss"
        x = kw["xi_out"]
        return np.ones_like(x), fake_info(x)
    mu = np.array([.5, 1, 2])
    with patch("sgl_analytic_amd.api.run", side_effect=fake_run):
        x, p, info = generate_pdf(2, .7, .3, .9, .05, .97, 3450, mu=mu, return_info=True)
    np.testing.assert_array_equal(x, mu)
    np.testing.assert_array_equal(p, 1 / mu)
    assert info["moment_interval"] == [-.5, 3.0]


def test_negative_residual_is_visible():
    with patch("sgl_analytic_amd.api.run", return_value=(np.array([-.1, 1.0]), fake_info(None))):
        with pytest.warns(RuntimeWarning):
            _, p, info = generate_pdf_lnmu(1, lnmu=[0, 1], return_info=True)
    assert p[0] == -.1
    assert abs(info["diagnostics"]["negative_mass"] - .05) < 1e-15


def test_nonfinite_output_raises():
    with patch("sgl_analytic_amd.api.run", return_value=(np.array([np.nan, 1.0]), fake_info(None))):
        with pytest.raises(RuntimeError):
            generate_pdf_lnmu(1, lnmu=[0, 1])


def test_solver_refuses_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from sgl_analytic_amd import _lib
    from sgl_analytic_amd.solver import Shape, Solver
    with pytest.raises(_lib.SglError):
        Solver(Shape(n_z_max=4, n_m=4, n_b=4, n_k=16, n_out=8), 1, path="exact")


def test_to_image_plane_closed_form():
    """A Gaussian source-plane PDF N(m, s^2) in ln mu maps to N(m + s^2, s^2)
    in the image plane (mu N(m, s^2) = e^{m + s^2/2} N(m + s^2, s^2))."""
    from sgl_analytic_amd import to_image_plane
    x = np.linspace(-3.0, 3.0, 6001)
    m, s = -0.02, 0.15
    g = lambda mm: np.exp(-0.5 * ((x - mm) / s) ** 2) / (s * np.sqrt(2 * np.pi))
    pi = to_image_plane(x, g(m))
    assert abs(np.trapezoid(pi, x) - 1.0) < 1e-12
    assert np.max(np.abs(pi - g(m + s * s))) < 1e-9 * np.max(g(m))


def test_clipped_moments_closed_form():
    """Uniform density on [-1, 1] clipped at 0.5: mean 0, variance 1/12,
    skewness 0; a shifted Gaussian well inside the clip keeps its moments."""
    from sgl_analytic_amd import clipped_moments
    x = np.linspace(-1.0, 1.0, 200001)
    c = clipped_moments(x, np.full_like(x, 0.5), clip=0.5)
    # the clip takes whole grid points (step 1e-5): mean and variance to a step
    assert abs(c["mean"]) < 1e-5 and abs(c["var"] - 1.0 / 12.0) < 1e-5 and abs(c["skew"]) < 1e-4
    s = 0.05
    g = np.exp(-0.5 * ((x - 0.1) / s) ** 2)
    c = clipped_moments(x, g, clip=1.0)
    assert abs(c["mean"] - 0.1) < 1e-9 and abs(c["var"] - s * s) < 1e-9 and abs(c["skew"]) < 1e-6
