"""Launch-shape knobs and the autotuner (needs a GPU).

Every launch shape must give the same PDFs: the GEMM tile and the exponent
block keep the order of every sum (their stages are bit-identical on the
same inputs), the binning block only reorders FP64 histogram adds (whole
solves agree to round-off)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SMALL = dict(n_z_max=8, n_m=32, n_b=64, n_k=1024, n_out=256)


def _solver(B=4, prec=1):
    from sgl_analytic_amd.solver import Shape, Solver, cosmo_records
    s = Solver(Shape(**SMALL), B, path="binned", precision=prec)
    zs = np.linspace(0.5, 3.0, B)
    p = np.array([[0.674, 0.315, 0.811, 0.0493, 0.965, 3402.0, 0.0, -1.0, 0.0, z] for z in zs])
    s.set_records(cosmo_records(p, SMALL["n_z_max"]))
    return s


def test_knob_range_checks():
    from sgl_analytic_amd import _lib
    L = _lib.load()
    assert L.sgl_set_tuning(0, 3) == _lib.SGL_ERR_INVALID
    assert L.sgl_set_tuning(3, 0) == _lib.SGL_ERR_INVALID
    assert L.sgl_set_tuning(-1, 0) == _lib.SGL_ERR_INVALID
    assert L.sgl_get_tuning(7) == -1


@pytest.mark.parametrize("prec", [1, 2])
def test_every_launch_shape_same_pdf(prec):
    from sgl_analytic_amd import autotune
    s = _solver(prec=prec)
    base_choice = autotune.current()
    base = {"bin_threads": 512, "lnphi_block": 256, "gemm_bm": 64}
    try:
        autotune.apply(base)
        p0 = s.run()[0].clone()
        torch.cuda.synchronize()
        # whole solves: the binning block reorders FP64 LDS atomic adds, and
        # so does every repeat of the same shape, so solves agree to round-off
        for name, (_, sizes) in autotune.KNOBS.items():
            for size in sizes:
                print(f"knob {name}={size} prec={prec}", flush=True)
                autotune.apply({**base, name: size})
                p = s.run()[0]
                torch.cuda.synchronize()
                diff = float((p - p0).abs().max()) / float(p0.abs().max())
                assert diff < 1e-5, (name, size, diff)
        # the exponent block and the GEMM tile keep the order of every sum:
        # rerun only their stage on the same inputs, bit-identical outputs
        autotune.apply(base)
        s.run()
        torch.cuda.synchronize()
        lam0, amp0, pp0 = s.lam.clone(), s.amp.clone(), s.p.clone()
        for name, outs in (("lnphi_block", (s.lam, s.amp)), ("gemm_bm", (s.p,))):
            ref = (lam0, amp0) if name == "lnphi_block" else (pp0,)
            for size in autotune.KNOBS[name][1]:
                print(f"stage {name}={size} prec={prec}", flush=True)
                autotune.apply({**base, name: size})
                autotune._stage(s, name)()
                torch.cuda.synchronize()
                for o, r in zip(outs, ref):
                    assert torch.equal(o, r), (name, size)
    finally:
        autotune.apply(base_choice)


def test_tune_picks_a_candidate_and_caches():
    from sgl_analytic_amd import autotune
    s = _solver()
    base_choice = autotune.current()
    try:
        r = autotune.tune(s, reps=1)
        for name, (_, sizes) in autotune.KNOBS.items():
            assert r["choice"][name] in sizes
            assert set(r["ms"][name]) == set(sizes)
            assert all(v > 0 for v in r["ms"][name].values())
        assert autotune.current() == r["choice"]
        assert autotune.tune(s) is r
    finally:
        autotune.apply(base_choice)
