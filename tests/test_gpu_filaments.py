"""Filament population on the device vs the FP64 CPU oracle (needs a GPU)."""
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
    d = SMALL
    osh = chain.Shape(n_z=d["n_z"], n_m=d["n_m"], n_b=d["n_b"], n_k=d["n_k"], n_out=d["n_out"],
                      filaments=True)
    dsh = Shape(n_z_max=d["n_z"], n_m=d["n_m"], n_b=d["n_b"], n_k=d["n_k"], n_out=d["n_out"],
                filaments=True)
    return osh, dsh


def _record(zs, n_z):
    from sgl_analytic_amd.solver import cosmo_records
    c = PLANCK
    return cosmo_records(np.array([[c["h"], c["Om"], c["s8"], c["Ob"], c["ns"], c["zeq"], 0.0,
                                    -1.0, 0.0, zs]]), n_z)


@pytest.fixture(scope="module")
def plain():
    from sgl_analytic_amd.solver import Solver
    osh, dsh = _shapes()
    return osh, dsh, Solver(dsh, 1, path="exact")


@pytest.mark.parametrize("zs", [0.5, 2.0])
def test_filament_rows_match_oracle(plain, zs):
    osh, dsh, s = plain
    print(f"filament rows zs={zs}", flush=True)
    s.run(_record(zs, dsh.n_z_max))
    torch.cuda.synchronize()
    Fd = s.table_views()["frows"][0]
    F = chain.fil_rows(Cosmology(**PLANCK), zs, osh)
    rel = (np.abs(Fd - F) / np.abs(F)).reshape(-1, 4).max(axis=0)
    print("max rel err (k0base, r_s, w_row, beta)", rel)
    assert np.all(rel < 2e-6), rel


@pytest.mark.parametrize("zs", [0.5, 2.0])
def test_lnphi_exact_with_filaments(plain, zs):
    """Exponent and slice amplitudes of halos + filaments on the device's own
    rows: FP64 to 1e-10 of the oracle's."""
    osh, dsh, s = plain
    print(f"lnphi exact with filaments zs={zs}", flush=True)
    s.run(_record(zs, dsh.n_z_max))
    torch.cuda.synchronize()
    tv = s.table_views()
    R = tv["rows"][0, :, :, :6].copy()
    F = tv["frows"][0].copy()
    kq = osh.k_nodes()
    xi, ws, _, _ = chain.cells(R, osh)
    lam, amp = chain.lnphi(xi, ws, R[:, :, 5], kq[1:])
    fxi, fws, _, _ = chain.cells_fil(F)
    lam_f, amp_f = chain.lnphi(fxi, fws, F[:, :, 3], kq[1:])
    lam, amp = lam + lam_f, amp + amp_f
    lam_d = s.lam[0].cpu().numpy()
    lam_d = lam_d[:, 0] + 1j * lam_d[:, 1]
    amp_d = s.amp[0].cpu().numpy()
    amp_d = amp_d[..., 0] + 1j * amp_d[..., 1]
    err = np.max(np.abs(lam_d - lam) / np.maximum(np.abs(lam), 1e-3))
    aerr = np.max(np.abs(amp_d - amp)) / np.max(np.abs(amp))
    print("lam rel err", err, "amp rel err", aerr, "filament share of Lambda(k_max)",
          abs(lam_f[-1]) / abs(lam[-1]))
    assert err < 1e-10 and aerr < 1e-10


def test_plain_pdf_with_filaments_vs_oracle(plain):
    osh, dsh, s = plain
    zs = 1.0
    print(f"plain pdf with filaments zs={zs}", flush=True)
    p, diag = s.run(_record(zs, dsh.n_z_max))
    torch.cuda.synchronize()
    r = chain.pdf(zs, osh)
    l1 = float(np.trapezoid(np.abs(p[0].cpu().numpy() - r.p), r.x))
    d = diag[0].cpu().numpy()
    print("L1", l1, "mass", d[0], r.info["moments"]["norm"])
    assert l1 < 1e-6
    assert abs(d[5] - r.info["moments"]["mu_mean"]) < 1e-6


@pytest.mark.parametrize("prec", [1, 2])
def test_batched_pdf_with_filaments_vs_oracle(prec):
    from sgl_analytic_amd.solver import Solver
    osh, dsh = _shapes()
    zs = 1.0
    print(f"batched pdf with filaments zs={zs} prec={prec}", flush=True)
    s = Solver(dsh, 1, path="binned", precision=prec)
    p, diag = s.run(_record(zs, dsh.n_z_max))
    torch.cuda.synchronize()
    r = chain.pdf(zs, osh)
    l1 = float(np.trapezoid(np.abs(p[0].cpu().numpy().astype(np.float64) - r.p), r.x))
    d = diag[0].cpu().numpy()
    print("L1", l1, "<mu>", d[5], r.info["moments"]["mu_mean"])
    assert l1 < 1e-3
    assert abs(d[5] - r.info["moments"]["mu_mean"]) < 1e-3


def test_api_fil_ingredient():
    """generate_pdf_lnmu with '+fil' runs the filament population; 'halo' does not."""
    import sgl_analytic_amd as sa
    print("api +fil", flush=True)
    x = np.linspace(-0.5, 1.5, 401)
    _, p_h, i_h = sa.generate_pdf_lnmu(1.0, lnmu=x, config="halo", return_info=True)
    _, p_f, i_f = sa.generate_pdf_lnmu(1.0, lnmu=x, config="+fil", return_info=True)
    assert "fil" in i_f["ingredients"] and "fil" not in i_h["ingredients"]
    assert float(np.trapezoid(np.abs(p_f - p_h), x)) > 1e-4
