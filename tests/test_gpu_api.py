"""The drop-in API on the device (needs a GPU)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_generate_pdf_lnmu_runs_the_reference_chain():
    import sgl_analytic_amd as sa
    from sgl_analytic_amd import ref2d
    print("api generate_pdf_lnmu z_s=1 default grid, ell + fil", flush=True)
    x, p, info = sa.generate_pdf_lnmu(1.0, config={"ell", "fil"}, return_info=True)
    assert x.shape == (3201,) and p.shape == (3201,)
    assert np.all(np.isfinite(p))
    assert info["backend"].startswith("MI355X") and info["N_fil"] > 0
    # the user grid shares its nodes with the body / tail grids of the explorer
    # run: the same arc integrals, so the same values
    r = ref2d.run(1.0, sa.api.check_cosmology({}), ell=True, fil=True)
    on_body = np.isin(np.round(r.xi, 10), np.round(x, 10))
    xb = r.xi[on_body]
    np.testing.assert_allclose(np.interp(xb, x, p), r.P[on_body], rtol=1e-12, atol=1e-14)
    assert on_body.sum() > 321
    mass = np.trapezoid(p, x)
    assert 0.99 < mass < 1.01
    # the positive closure needs the bias ingredient, as in the reference
    with pytest.raises(ValueError):
        sa.generate_pdf_lnmu(1.0, lnmu=x[:10], closure="vp", config="-bias")


def test_generate_pdf_jacobian_on_device():
    import sgl_analytic_amd as sa
    print("api generate_pdf", flush=True)
    mu, pm = sa.generate_pdf(1.0, mu=np.geomspace(0.8, 3.0, 64), config="+ell")
    x, pl = sa.generate_pdf_lnmu(1.0, lnmu=np.log(mu), config="+ell")
    np.testing.assert_allclose(pm, pl / mu, rtol=1e-14)


def test_batch_api_matches_plain_within_tolerance():
    import sgl_analytic_amd as sa
    from sgl_analytic_amd.solver import Shape, Solver, cosmo_records, slice_count
    print("api batch B=3", flush=True)
    params = np.array([[0.674, 0.315, 0.811, 0.0493, 0.965, 3402.0, 0.0, -1.0, 0.0, 0.5],
                       [0.7, 0.3, 0.85, 0.045, 0.97, 3402.0, 0.03, -0.9, 0.2, 2.0],
                       [0.65, 0.35, 0.75, 0.05, 0.94, 3402.0, -0.04, -1.1, -0.3, 8.0]])
    x, pb, db = sa.generate_pdf_lnmu_batch(params, graph=False)
    n_z = slice_count(params[:, 9])
    pl = Solver(Shape(n_z_max=int(n_z.max())), 1, path="exact")
    for i in range(3):
        p1, d1 = pl.run(cosmo_records(params[i:i + 1], n_z[i]))
        l1 = np.trapezoid(np.abs(pb[i] - p1[0].cpu().numpy()), x)
        print("instance", i, "L1", l1, flush=True)
        assert l1 < 1e-3
