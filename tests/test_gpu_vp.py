"""Positive ("vp") clustering closure on the plain path against the FP64
oracle on identical inputs (needs a GPU): the calibrated eta and target, the
closure term ln R(k), and the PDF; halo rows alone, with the filament rows as
further groups, a batch of two instances with different slice counts, and
the public API."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from oracle import chain  # noqa: E402

PLANCK = dict(h=0.674, Om=0.315, s8=0.811, Ob=0.0493, ns=0.965, zeq=3402.0)
SMALL = dict(n_m=32, n_b=64, n_k=1024, n_out=256)


def _records(zs_list, n_z_list):
    from sgl_analytic_amd.solver import cosmo_records
    c = PLANCK
    P = np.array([[c["h"], c["Om"], c["s8"], c["Ob"], c["ns"], c["zeq"], 0.0, -1.0, 0.0, zs]
                  for zs in zs_list])
    return cosmo_records(P, np.array(n_z_list))


def _check(s, b, zs, osh, tag):
    r = chain.pdf(zs, osh, closure="vp", keep=True)
    eta, target = s.vp_calibration()
    p = s.p[b].cpu().numpy()
    R_or = np.exp(r.info["psi2"])
    nk1 = osh.n_k + 1
    ph = s.phi.view(torch.float64).reshape(s.B, nk1, 2)[b].cpu().numpy() @ np.array([1.0, 1j])
    dphi = float(np.max(np.abs(ph - r.info["phi"])))
    l1 = float(np.trapezoid(np.abs(p - r.p), r.x))
    print(f"{tag}: eta {eta[b]:.15g} / {r.info['eta']:.15g}, target {target[b]:.6g} / "
          f"{r.info['vp_target']:.6g}, max |dphi| {dphi:.3g}, L1 {l1:.3g}", flush=True)
    # the device tables (rows, covariance) agree with the oracle's to ~1e-8
    # relative, which bounds the agreement of everything downstream
    assert abs(target[b] / r.info["vp_target"] - 1) < 1e-6
    assert abs(eta[b] / r.info["eta"] - 1) < 1e-6
    assert dphi < 1e-8                          # phi = window exp(lam + ln R)
    assert np.max(np.abs(R_or - 1)) > 1e-3     # the closure term is not trivially 1
    assert l1 < 1e-9


@pytest.mark.parametrize("zs", [0.5, 1.5])
def test_vp_plain_pdf(zs):
    from sgl_analytic_amd.solver import Shape, Solver
    print(f"vp plain zs={zs}", flush=True)
    s = Solver(Shape(n_z_max=8, **SMALL), 1, path="exact", closure="vp")
    s.run(_records([zs], [8]))
    torch.cuda.synchronize()
    assert int(s.status.cpu()[0]) == 0
    _check(s, 0, zs, chain.Shape(n_z=8, **SMALL), f"vp zs={zs}")


def test_vp_filament_groups():
    from sgl_analytic_amd.solver import Shape, Solver
    print("vp filaments zs=0.5", flush=True)
    s = Solver(Shape(n_z_max=8, filaments=True, **SMALL), 1, path="exact", closure="vp")
    s.run(_records([0.5], [8]))
    torch.cuda.synchronize()
    assert int(s.status.cpu()[0]) == 0
    _check(s, 0, 0.5, chain.Shape(n_z=8, filaments=True, **SMALL), "vp fil")


def test_vp_batch_of_two_slice_counts():
    from sgl_analytic_amd.solver import Shape, Solver
    print("vp batch n_z = 8, 5", flush=True)
    s = Solver(Shape(n_z_max=8, **SMALL), 2, path="exact", closure="vp")
    s.run(_records([1.0, 0.3], [8, 5]))
    torch.cuda.synchronize()
    assert int(s.status.cpu()[0]) == 0
    _check(s, 0, 1.0, chain.Shape(n_z=8, **SMALL), "vp b0")
    _check(s, 1, 0.3, chain.Shape(n_z=5, **SMALL), "vp b1")


def test_vp_through_the_api():
    from sgl_analytic_amd import api
    print("vp api zs=0.3", flush=True)
    x = np.linspace(-0.5, 2.0, 801)
    pv, info = api.run_kernel_shape(0.3, "full", cosmo={}, xi_out=x, closure="vp", n_z=16)
    pg, _ = api.run_kernel_shape(0.3, "full", cosmo={}, xi_out=x, closure="gauss", n_z=16)
    pc = info["pos_closure"]
    print("api vp", pc, float(np.trapezoid(pv, x)), float(np.trapezoid(np.abs(pv - pg), x)),
          flush=True)
    assert pc["kind"] == "vp" and pc["nstate"] == 48 and pc["shells"] == 16
    assert np.isfinite(pc["eta"]) and pc["eta"] > 0 and pc["target_kappa"] > 0
    assert abs(np.trapezoid(pv, x) - 1) < 1e-3
    assert 1e-5 < float(np.trapezoid(np.abs(pv - pg), x)) < 0.1


def test_vp_full_shape_properties():
    """The plain-path workload with the vp closure (z_s = 0.3: 64 slices x 256
    masses x 512 impact parameters, 8192 k nodes, 2048 ln mu nodes) checked
    through properties that hold at any size: |phi(k)| <= 1 (given the latent
    states the rates are non-negative, so phi is a mixture of compound-Poisson
    characteristic functions), ln R -> -1/2 k^2 target at the smallest k (the
    calibration's defining limit), and the grid mass and <mu> of the Gaussian
    closure on the same grids (the grid mass misses the no-jump point mass,
    which phi keeps as k -> infinity: e^{-sum w} times the closure's factor,
    E e^{-sum (lambda - 1) w} for vp and e^{Psi_2} for gauss, equal to second
    order in the clustering), and closeness to the Gaussian closure's PDF."""
    from sgl_analytic_amd.solver import Shape, Solver
    print("vp full shape zs=0.3", flush=True)
    sh = Shape()
    rec = _records([0.3], [64])
    sv = Solver(sh, 1, path="exact", closure="vp")
    p, d = sv.run(rec)
    torch.cuda.synchronize()
    assert int(sv.status.cpu()[0]) == 0
    eta, target = sv.vp_calibration()
    lnr = sv.psi2[0].cpu().numpy() @ np.array([1.0, 1j])
    lam = sv.lam[0].cpu().numpy() @ np.array([1.0, 1j])
    k = sh.k_nodes()[1:]
    phi_abs = np.abs(np.exp(lam + lnr))
    small = -0.5 * k[0] ** 2 * target[0]
    pv, dv = p[0].cpu().numpy().copy(), d[0].cpu().numpy().copy()
    sg = Solver(sh, 1, path="exact")
    pg, dg = sg.run(rec)
    torch.cuda.synchronize()
    x = sh.out_grid()
    l1 = float(np.trapezoid(np.abs(pv - pg[0].cpu().numpy()), x))
    print(f"eta {eta[0]:.6g} target {target[0]:.4g} max|phi| {phi_abs.max():.15g} "
          f"lnR(k0) {lnr[0]:.6g} vs {small:.6g}; mass {dv[0]:.9f} <mu> {dv[5]:.9f} "
          f"(gauss {dg[0, 0].item():.9f}, {dg[0, 5].item():.9f}); L1 vs gauss {l1:.3g}",
          flush=True)
    assert phi_abs.max() <= 1.0 + 1e-12
    assert abs(lnr[0].real / small - 1) < 1e-3 and abs(lnr[0].imag) < 1e-3 * abs(small)
    assert abs(dv[0] - dg[0, 0].item()) < 1e-3
    assert abs(dv[5] - dg[0, 5].item()) < 1e-4
    assert 1e-6 < l1 < 0.05


def test_vp_batch_api_uneven_slices():
    """generate_pdf_lnmu_batch(closure="vp"): three cosmologies with their own
    slice counts ceil(24 (1 + ln(1 + z_s))) in one plain-path launch, each
    against the oracle at its own slice count."""
    from sgl_analytic_amd import api
    from sgl_analytic_amd.solver import Shape, slice_count
    print("vp batch api", flush=True)
    zs = np.array([0.3, 1.0, 2.0])
    c = PLANCK
    P = np.array([[c["h"], c["Om"], c["s8"], c["Ob"], c["ns"], c["zeq"], 0.0, -1.0, 0.0, z]
                  for z in zs])
    n_z = slice_count(zs)
    x, p, d = api.generate_pdf_lnmu_batch(P, shape=Shape(n_z_max=int(n_z.max()), **SMALL),
                                          closure="vp")
    assert p.dtype == np.float64 and p.shape == (3, SMALL["n_out"])
    for b in range(3):
        r = chain.pdf(float(zs[b]), chain.Shape(n_z=int(n_z[b]), **SMALL), closure="vp")
        l1 = float(np.trapezoid(np.abs(p[b] - r.p), r.x))
        print(f"zs={zs[b]} n_z={n_z[b]} L1 {l1:.3g}", flush=True)
        assert l1 < 1e-9
    # chunked by the workspace cap (chunks of 2 + 1): the same numbers
    import ctypes
    from sgl_analytic_amd import _lib
    sh = Shape(n_z_max=int(n_z.max()), **SMALL)
    per = _lib.load().sgl_vp_bytes(ctypes.byref(sh.c()), 1)
    keep = api.VP_BATCH_BYTES
    try:
        api.VP_BATCH_BYTES = 2 * per
        _, p2, d2 = api.generate_pdf_lnmu_batch(P, shape=sh, closure="vp")
    finally:
        api.VP_BATCH_BYTES = keep
    assert np.array_equal(p2, p) and np.array_equal(d2, d)
    with pytest.raises(ValueError):
        api.generate_pdf_lnmu_batch(P, closure="lognormal")
