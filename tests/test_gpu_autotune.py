"""Launch-shape knobs and the autotuner (needs a GPU).

Every launch shape must give the same PDFs bit for bit: the GEMM tile and
the exponent block keep the order of every sum, and the binning block size
and schedule (one workgroup per slice, or persistent workgroups over a work
queue) only reorder the histogram's fixed-point integer adds."""
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
    # uneven slice counts, as in a batch of mixed source redshifts
    n_z = np.array([SMALL["n_z_max"], 5, 3, 7] * B)[:B]
    s.set_records(cosmo_records(p, n_z))
    return s


def test_knob_range_checks():
    from sgl_analytic_amd import _lib
    L = _lib.load()
    assert L.sgl_set_tuning(0, 3) == _lib.SGL_ERR_INVALID
    assert L.sgl_set_tuning(3, 2) == _lib.SGL_ERR_INVALID
    assert L.sgl_set_tuning(4, 0) == _lib.SGL_ERR_INVALID
    assert L.sgl_set_tuning(-1, 0) == _lib.SGL_ERR_INVALID
    assert L.sgl_get_tuning(7) == -1


@pytest.mark.parametrize("prec", [1, 2])
def test_every_launch_shape_same_pdf(prec):
    from sgl_analytic_amd import autotune
    s = _solver(prec=prec)
    base_choice = autotune.current()
    base = {"bin_threads": 512, "lnphi_block": 256, "gemm_bm": 64, "bin_schedule": 0}
    try:
        autotune.apply(base)
        p0 = s.run()[0].clone()
        torch.cuda.synchronize()
        # whole solves: every knob value, and the queue schedule at every
        # binning block size, gives the same bits
        cases = [{name: size} for name, (_, sizes) in autotune.KNOBS.items() for size in sizes]
        cases += [{"bin_schedule": 1, "bin_threads": t} for t in autotune.KNOBS["bin_threads"][1]]
        for case in cases:
            print(f"knobs {case} prec={prec}", flush=True)
            autotune.apply({**base, **case})
            p = s.run()[0]
            torch.cuda.synchronize()
            diff = float((p - p0).abs().max()) / float(p0.abs().max())
            assert torch.equal(p, p0), (case, diff)
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
