"""The reference's cosmology options beyond the pipeline defaults (needs a
GPU): radiation=None and growth_mode="legacy" (Heath-integral D in the mass
function, Carroll-Press-Turner growth in the bias), each against the FP64
oracle on identical inputs -- slices, rows, the plain FP64 PDF and the batched
PDF -- and through the public API."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import chain  # noqa: E402
from oracle.cosmo import Cosmology  # noqa: E402

PLANCK = dict(h=0.674, Om=0.315, s8=0.811, Ob=0.0493, ns=0.965, zeq=3402.0)
SMALL = dict(n_z=8, n_m=32, n_b=64, n_k=1024, n_out=256)
OPTIONS = [(None, "ode"), ("engine", "legacy"), (None, "legacy")]


def _record(zs, n_z, radiation, growth, Ok=0.0):
    from sgl_analytic_amd.solver import cosmo_records
    c = PLANCK
    return cosmo_records(np.array([[c["h"], c["Om"], c["s8"], c["Ob"], c["ns"], c["zeq"], Ok,
                                    -1.0, 0.0, zs]]), n_z, radiation=radiation,
                         growth_mode=growth)


@pytest.mark.parametrize("radiation,growth", OPTIONS)
def test_options_rows_and_plain_pdf(radiation, growth):
    from sgl_analytic_amd.solver import Shape, Solver
    zs, Ok = 1.5, 0.03
    print(f"cosmology options radiation={radiation} growth={growth} zs={zs} Ok={Ok}", flush=True)
    osh = chain.Shape(**SMALL)
    dsh = Shape(n_z_max=8, n_m=32, n_b=64, n_k=1024, n_out=256)
    s = Solver(dsh, 1, path="exact")
    p, diag = s.run(_record(zs, dsh.n_z_max, radiation, growth, Ok))
    torch.cuda.synchronize()
    cos = Cosmology(**PLANCK, Ok=Ok, radiation=radiation, growth_mode=growth)
    T = chain.rows(cos, zs, osh)
    tv = s.table_views()
    Rd = tv["rows"][0, :, :, :6]
    worst = (np.abs(Rd - T["rows"]) / np.abs(T["rows"])).reshape(-1, 6).max(axis=0)
    print("rows max rel err", worst)
    assert np.all(worst < 2e-6), worst
    sl = tv["slices"][0]
    np.testing.assert_allclose(sl[:, 0], T["z"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(sl[:, 4], cos.D(T["z"]), rtol=1e-8)     # D(z)
    np.testing.assert_allclose(sl[:, 7], cos.Dg(T["z"]), rtol=1e-8)    # Dg(z)
    r = chain.pdf(zs, osh, cosmo=dict(**PLANCK, Ok=Ok, radiation=radiation, growth_mode=growth))
    l1 = float(np.trapezoid(np.abs(p[0].cpu().numpy() - r.p), r.x))
    print("plain L1", l1)
    assert l1 < 1e-6


@pytest.mark.parametrize("radiation,growth", OPTIONS)
def test_options_batched_pdf(radiation, growth):
    from sgl_analytic_amd._lib import PREC_FP32
    from sgl_analytic_amd.solver import Shape, Solver
    zs = 1.0
    print(f"cosmology options batched radiation={radiation} growth={growth}", flush=True)
    osh = chain.Shape(**SMALL)
    dsh = Shape(n_z_max=8, n_m=32, n_b=64, n_k=1024, n_out=256)
    s = Solver(dsh, 1, path="binned", precision=PREC_FP32)
    p, _ = s.run(_record(zs, dsh.n_z_max, radiation, growth))
    torch.cuda.synchronize()
    r = chain.pdf(zs, osh, cosmo=dict(**PLANCK, radiation=radiation, growth_mode=growth))
    l1 = float(np.trapezoid(np.abs(p[0].cpu().numpy().astype(float) - r.p), r.x))
    print("batched L1", l1)
    assert l1 < 1e-3


def test_api_growth_and_radiation():
    """The options change the PDF through the public API; "auto" is "legacy"
    for LambdaCDM."""
    from sgl_analytic_amd import api
    print("api growth/radiation options", flush=True)
    x = np.linspace(-0.2, 0.6, 801)
    p0, i0 = api.run_kernel_shape(1.0, "halo", cosmo={}, xi_out=x, n_z=16)
    pa, ia = api.run_kernel_shape(1.0, "halo", cosmo=dict(growth_mode="auto"), xi_out=x, n_z=16)
    pl, il = api.run_kernel_shape(1.0, "halo", cosmo=dict(growth_mode="legacy"), xi_out=x, n_z=16)
    pn, i_n = api.run_kernel_shape(1.0, "halo", cosmo=dict(radiation=None), xi_out=x, n_z=16)
    assert ia["growth_mode"] == il["growth_mode"] == "legacy" and i_n["radiation"] is None
    np.testing.assert_array_equal(pa, pl)
    d_leg = float(np.max(np.abs(pl - p0))) / float(np.max(p0))
    d_rad = float(np.max(np.abs(pn - p0))) / float(np.max(p0))
    print("relative change legacy", d_leg, "no radiation", d_rad)
    assert d_leg > 1e-6 and d_rad > 1e-6
