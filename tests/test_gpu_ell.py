"""Halo ellipticity (the reference's "ell" ingredient) against the FP64 oracle
on identical inputs (needs a GPU): the rows' eps(M, z), the plain FP64 PDF and
the batched PDF, and through the public API."""
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
    osh = chain.Shape(ell=True, **SMALL)
    dsh = Shape(n_z_max=8, n_m=32, n_b=64, n_k=1024, n_out=256, ell=True)
    return osh, dsh


def _record(zs, n_z):
    from sgl_analytic_amd.solver import cosmo_records
    c = PLANCK
    return cosmo_records(np.array([[c["h"], c["Om"], c["s8"], c["Ob"], c["ns"], c["zeq"], 0.0,
                                    -1.0, 0.0, zs]]), n_z)


@pytest.mark.parametrize("zs", [0.5, 1.5])
def test_ell_rows_and_plain_pdf(zs):
    from sgl_analytic_amd.solver import Solver
    osh, dsh = _shapes()
    print(f"ell plain zs={zs}", flush=True)
    s = Solver(dsh, 1, path="exact")
    p, diag = s.run(_record(zs, dsh.n_z_max))
    torch.cuda.synchronize()
    T = chain.rows(Cosmology(**PLANCK), zs, osh)
    tv = s.table_views()
    Rd = tv["rows"][0]
    worst = (np.abs(Rd[:, :, :6] - T["rows"]) / np.abs(T["rows"])).reshape(-1, 6).max(axis=0)
    print("rows max rel err", worst)
    assert np.all(worst < 2e-6), worst
    eps_err = np.max(np.abs(Rd[:, :, 6] - T["eps"]))
    print("eps max abs err", eps_err, "eps range", T["eps"].min(), T["eps"].max())
    assert eps_err < 1e-6 and T["eps"].max() > 0.1
    r = chain.pdf(zs, osh)
    l1 = float(np.trapezoid(np.abs(p[0].cpu().numpy() - r.p), r.x))
    print("plain L1", l1)
    assert l1 < 1e-6


@pytest.mark.parametrize("zs", [0.5, 1.5])
def test_ell_batched_pdf(zs):
    from sgl_analytic_amd._lib import PREC_FP32
    from sgl_analytic_amd.solver import Solver
    osh, dsh = _shapes()
    print(f"ell batched zs={zs}", flush=True)
    s = Solver(dsh, 1, path="binned", precision=PREC_FP32)
    p, diag = s.run(_record(zs, dsh.n_z_max))
    torch.cuda.synchronize()
    r = chain.pdf(zs, osh)
    l1 = float(np.trapezoid(np.abs(p[0].cpu().numpy().astype(float) - r.p), r.x))
    print("batched L1", l1)
    assert l1 < 1e-3


def test_api_ell_differs_from_halo():
    from sgl_analytic_amd import api
    print("api +ell", flush=True)
    x = np.linspace(-0.2, 0.6, 801)
    ph, _ = api.run_kernel_shape(1.0, "halo", cosmo={}, xi_out=x, n_z=16)
    pe, info = api.run_kernel_shape(1.0, "+ell", cosmo={}, xi_out=x, n_z=16)
    assert "ell" in info["ingredients"]
    d = float(np.max(np.abs(pe - ph))) / float(np.max(ph))
    print("relative change with ellipticity", d)
    assert d > 1e-4
