"""Halos truncated at r_vir (the reference's edge="rvir") against the FP64
oracle on identical inputs (needs a GPU): rows, the plain FP64 PDF and the
batched PDF."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import chain  # noqa: E402
from oracle.cosmo import Cosmology  # noqa: E402

PLANCK = dict(h=0.674, Om=0.315, s8=0.811, Ob=0.0493, ns=0.965, zeq=3402.0)
SMALL = dict(n_z=8, n_m=32, n_b=64, n_k=1024, n_out=256)


def _shapes():
    from sgl_analytic_amd.solver import Shape
    osh = chain.Shape(edge="rvir", **SMALL)
    dsh = Shape(n_z_max=8, n_m=32, n_b=64, n_k=1024, n_out=256, edge="rvir")
    return osh, dsh


def _record(zs, n_z):
    from sgl_analytic_amd.solver import cosmo_records
    c = PLANCK
    return cosmo_records(np.array([[c["h"], c["Om"], c["s8"], c["Ob"], c["ns"], c["zeq"], 0.0,
                                    -1.0, 0.0, zs]]), n_z)


@pytest.mark.parametrize("zs", [0.5, 2.0])
def test_rvir_rows_and_plain_pdf(zs):
    from sgl_analytic_amd.solver import Solver
    osh, dsh = _shapes()
    print(f"edge rvir plain zs={zs}", flush=True)
    s = Solver(dsh, 1, path="exact")
    p, diag = s.run(_record(zs, dsh.n_z_max))
    torch.cuda.synchronize()
    T = chain.rows(Cosmology(**PLANCK), zs, osh)
    Rd = s.table_views()["rows"][0, :, :, :6]
    worst = (np.abs(Rd - T["rows"]) / np.abs(T["rows"])).reshape(-1, 6).max(axis=0)
    print("rows max rel err", worst)
    assert np.all(worst < 2e-6), worst
    # x_t = r_vir / r_s on the device too
    assert np.allclose(Rd[:, :, 2], Rd[:, :, 3] / Rd[:, :, 1], rtol=1e-12)
    r = chain.pdf(zs, osh)
    l1 = float(np.trapezoid(np.abs(p[0].cpu().numpy() - r.p), r.x))
    d = diag[0].cpu().numpy()
    print("plain L1", l1, "mass", d[0], r.info["moments"]["norm"])
    assert l1 < 1e-6
    assert abs(d[5] - r.info["moments"]["mu_mean"]) < 1e-6


@pytest.mark.parametrize("zs", [0.5, 2.0])
def test_rvir_batched_pdf(zs):
    from sgl_analytic_amd._lib import PREC_FP32
    from sgl_analytic_amd.solver import Solver
    osh, dsh = _shapes()
    print(f"edge rvir batched zs={zs}", flush=True)
    s = Solver(dsh, 1, path="binned", precision=PREC_FP32)
    p, diag = s.run(_record(zs, dsh.n_z_max))
    torch.cuda.synchronize()
    r = chain.pdf(zs, osh)
    l1 = float(np.trapezoid(np.abs(p[0].cpu().numpy().astype(float) - r.p), r.x))
    print("batched L1", l1)
    assert l1 < 1e-3


def test_api_edge_rvir_differs_from_r200():
    import sgl_analytic_amd as sa
    print("api edge rvir", flush=True)
    from sgl_analytic_amd import api
    x = np.linspace(-0.2, 0.6, 801)
    p200, _ = api.run_kernel_shape(1.0, "halo", cosmo={}, xi_out=x, n_z=16)
    pv, info = api.run_kernel_shape(1.0, "halo", cosmo={}, xi_out=x, n_z=16, edge="rvir")
    assert info["edge"] == "rvir"
    # more projected mass per halo (r_vir > r200 at z ~ 0.5): a wider PDF
    v = lambda p: np.trapezoid(x * x * p, x) - np.trapezoid(x * p, x) ** 2
    print("var r200", v(p200), "rvir", v(pv))
    assert v(pv) > v(p200)
    assert sa.generate_pdf_lnmu is not None
