#!/usr/bin/env python
"""Benchmark: PDFs/s of the batched MI355X solver at <= 1e-3 L1 vs FP64, and of
the surrogate path.

Workload (one "step"): one batched solve of B = 4096 cosmologies drawn by a
seeded Latin hypercube over h[0.6,0.8], Om[0.2,0.4], sigma8[0.7,0.9],
Ob[0.04,0.06], ns[0.92,1.0], Ok[-0.05,0.05], w0[-1.2,-0.8], wa[-0.5,0.5],
zeq = 3402, z_s log-uniform in [0.3, 10]; ceil(24 (1 + ln(1 + z_s))) slices x
256 masses x 512 impact parameters per instance, 8192 log-k nodes in
[1e-2, 1e6], 2048 ln(mu) nodes on [-0.5, 3].  The step is one replay of the
captured HIP graph (tables -> binning -> exponent -> two-point term -> phi ->
MFMA inversion -> diagnostics); the cosmology records are resident in HBM.
Launch shapes are picked first by the autotuner on this batch.

Reported beside it:
- the plain single-cosmology FP64 GPU path on the 64 x 256 x 512 shape at
  z_s = 1 (PDFs/s and the ratio);
- the L1 / <mu> / mass / P(mu>2) errors of sampled batch members against the
  FP64 plain path;
- the surrogate path on 65536 fresh LHS cosmologies: throughput, acceptance,
  every accepted PDF checked against the batched solver, and the same batch
  without the networks (what the networks gain);
- the roofline of the dominant kernel (k_bin_cells) and the FP64 C oracle
  timed on the host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEED = 20261203
# MI355X dense peaks; FP32 vector without packed-FP32 instructions (this build
# disables them), which is the FP32 rate the binning kernel can reach.
PEAK = {"fp64": 78.6, "fp32": 78.65, "bf16": 2500.0, "hbm": 8000.0}
# Algorithmic work of the dominant kernel, k_bin_cells, per generated cell: the
# arithmetic of its formula on the common path, counted at source level (FMA =
# 2, transcendental = 1): column position and log (3), 1/x^2 (1), the FP32
# truncated-NFW profile -- P, Q, R, z with three square roots (14), G(z) (4),
# kappa (7), the mean convergence with its log (17), kappa and gamma (3) --
# the FP64 jump -ln detA and tilted weight (11), the bin from the FP32 log and
# its base exponent (7), the fixed-point place and the four rounded moments
# (13): 80.  The compiled kernel issues 126.0 VALU lane instructions per cell
# (profiles/batched_B4096_counters.txt: moves, compares, selects, conversions
# and both sides of divergent profile branches on top of these), at a
# saturated VALU issue rate.
CELL_FLOPS = 80
# HBM bytes of one k_bin_cells dispatch of the B = 4096 workload: FETCH_SIZE x 2.0
# + WRITE_SIZE x 1.0 KiB, the factors being what those counters read for a 1 GiB
# device copy (bench.py --copy-calibration) divided by 1 GiB.
BIN_TRAFFIC_B4096 = 2.10804e6 * 1024 * 2.0 + 6.03000e6 * 1024 * 1.0


def lhs_params(B, seed=SEED):
    from scipy.stats import qmc
    u = qmc.LatinHypercube(d=9, seed=seed).random(B)
    lo = np.array([0.6, 0.2, 0.7, 0.04, 0.92, -0.05, -1.2, -0.5, math.log(0.3)])
    hi = np.array([0.8, 0.4, 0.9, 0.06, 1.0, 0.05, -0.8, 0.5, math.log(10.0)])
    v = lo + u * (hi - lo)
    zs = np.exp(v[:, 8])
    return np.column_stack([v[:, 0], v[:, 1], v[:, 2], v[:, 3], v[:, 4], np.full(B, 3402.0),
                            v[:, 5], v[:, 6], v[:, 7], zs])


def trapz_l1(x, a, b):
    return float(np.trapezoid(np.abs(a - b), x))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--precision", choices=["fp32", "bf16x3"], default="fp32")
    ap.add_argument("--verify", type=int, default=8, help="batch members checked vs FP64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--precision-study", type=int, default=0, metavar="N",
                    help="instead of the bench line: FP32 vs split-BF16x3 inversion on the "
                         "same batch, N members each checked against the FP64 plain path")
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--sur-batch", type=int, default=65536, help="surrogate inference batch")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-sur-verify", action="store_true",
                    help="skip the solver check of the surrogate batch (for kernel profiles)")
    ap.add_argument("--sur-fp64-frac", type=float, default=0.01,
                    help="accepted surrogate PDFs with the largest network disagreement checked "
                         "against the plain FP64 path (fraction of the accepted)")
    ap.add_argument("--sur-fp64-random", type=int, default=64,
                    help="further accepted surrogate PDFs, drawn at random, checked vs FP64")
    ap.add_argument("--ref-only", action="store_true",
                    help="only the reference-chain leg (generate_pdf_lnmu's 2-D chain at z_s = 1)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only a 1 GiB device-to-device copy: the reference kernel for scaling "
                         "the FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the batch dimension "
              "shards over GPUs with no exchange, so the per-GPU figure is the N=1 line.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_calibration:
        a = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": a.numel() * 4}))
        return
    from sgl_analytic_amd import solver as S
    from sgl_analytic_amd._lib import PREC_BF16X3, PREC_FP32
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    if args.ref_only:
        print(json.dumps({"ref_chain": ref_chain_leg(log)}), flush=True)
        return

    prec = PREC_FP32 if args.precision == "fp32" else PREC_BF16X3
    B = args.batch
    if args.precision_study:
        precision_study(args, B)
        return
    params = lhs_params(B)
    n_z = S.slice_count(params[:, 9])
    shape = S.Shape(n_z_max=int(n_z.max()))
    rec = S.cosmo_records(params, n_z)

    # ---------------------------------------------------------- batched --
    log(f"[bench] batched B={B} n_z_max={shape.n_z_max} precision={args.precision}")
    sv = S.Solver(shape, B, path="binned", precision=prec)
    sv.set_records(rec)
    tuning = None
    if not args.no_tune:
        from sgl_analytic_amd import autotune
        tuning = autotune.tune(sv)
        log("[bench] tuned launch shapes", tuning["choice"],
            {k: {s_: round(v, 2) for s_, v in m.items()} for k, m in tuning["ms"].items()})
    stage = stage_times(sv, S)          # eager pass: per-stage event times (also builds E)
    log("[bench] stage ms", {k: round(v, 3) for k, v in stage.items()})
    t_bin = bin_pass_ms(sv, S)
    log(f"[bench] k_bin_cells {t_bin:.2f} ms per launch")
    if not args.no_graph:
        sv.capture()
    for _ in range(args.warmup):
        sv.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        sv.run()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = e0.elapsed_time(e1)
    ms_per_step = 1e3 * wall / args.steps
    value = B * args.steps / wall
    p_batch = sv.p.float().cpu().numpy().astype(np.float64)
    d_batch = sv.diag.cpu().numpy()
    log(f"[bench] batched {ms_per_step:.1f} ms/step, {value:.1f} PDFs/s (events {gpu_ms / args.steps:.1f} ms)")

    # ----------------------------------------------------- plain FP64 path --
    x = shape.out_grid()
    rng = np.random.default_rng(SEED)
    idx = np.sort(rng.choice(B, size=min(args.verify, B), replace=False))
    pl = S.Solver(shape, 1, path="exact")
    errs, t_members = [], []
    if len(idx):
        pl.run(rec[idx[0]:idx[0] + 1])          # warm-up (first launches)
        torch.cuda.synchronize()
    for i in idx:
        tm = time.perf_counter()
        p1, d1 = pl.run(rec[i:i + 1])
        torch.cuda.synchronize()
        t_members.append(time.perf_counter() - tm)
        pe = p1[0].cpu().numpy()
        de = d1[0].cpu().numpy()
        errs.append(dict(i=int(i), z_s=float(params[i, 9]), L1=trapz_l1(x, p_batch[i], pe),
                         maxabs=float(np.max(np.abs(p_batch[i] - pe))),
                         dmu=float(d_batch[i, 5] - de[5]), dmass=float(d_batch[i, 0] - de[0]),
                         dtail=float(d_batch[i, 6] - de[6])))
        log("[bench] verify", errs[-1])
    del pl
    # plain path timing on the single-cosmology kernel shape (64 slices, z_s = 1, Planck)
    plain_shape = S.Shape(n_z_max=64)
    pp = S.Solver(plain_shape, 1, path="exact")
    prec1 = S.cosmo_records(np.array([[0.674, 0.315, 0.811, 0.0493, 0.965, 3402.0, 0.0, -1.0,
                                       0.0, 1.0]]), 64)
    pp.run(prec1)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    nplain = 2
    for _ in range(nplain):
        pp.run(prec1)
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - t1) / nplain
    plain_rate = 1.0 / plain_s
    log(f"[bench] plain FP64 path {plain_s * 1e3:.1f} ms/PDF -> {plain_rate:.2f} PDFs/s")
    del pp

    ref = ref_chain_leg(log)

    # ------------------------------------------------------ surrogate --
    sur = None if args.no_surrogate else surrogate_leg(args, shape, sv, value, log, plain_rate)
    del sv

    # ------------------------------------------------------- roofline --
    roof = roofline(t_bin, shape, B, n_z)

    # ----------------------------------------------------- CPU baseline --
    cpu = None if args.no_cpu_baseline else cpu_baseline(log)

    worst = max(errs, key=lambda e: e["L1"]) if errs else None
    out = {
        "metric": "PDFs/sec at <=1e-3 L1 vs FP64: batched+surrogate vs plain GPU solver",
        "value": round(value, 2),
        "unit": "PDFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32" if prec == PREC_FP32 else "f64+bf16x3",
        "data": "synthetic (seeded Latin-hypercube cosmologies; no external data)",
        "config": {"workload": f"batched solve, B={B} LHS cosmologies, z_s in [0.3,10], "
                               f"ceil(24(1+ln(1+z_s))) slices x 256 masses x 512 b, 8192 k, "
                               f"2048 ln mu nodes; ingredients: truncated NFW halos (r200 "
                               f"edge) + the slice-covariance two-point term (Gaussian "
                               f"closure), no ellipticity, filaments or subhalos",
                   "ingredients": ["halos", "two-point (gauss)"], "batch": B, "seed": SEED,
                   "precision": args.precision, "graph": not args.no_graph},
        "plain_gpu_pdfs_per_s": round(plain_rate, 3),
        "batched_over_plain": round(value / plain_rate, 1),
        # the same ratio against the plain path on the batch's own members
        # (their own slice counts, mean below 64)
        "plain_gpu_pdfs_per_s_on_batch_members":
            round(len(t_members) / sum(t_members), 3) if t_members else None,
        "batched_over_plain_on_batch_members":
            round(value * sum(t_members) / len(t_members), 1) if t_members else None,
        "batch_mean_slices": round(float(np.mean(n_z)), 2),
        "verify": {"n": len(errs), "max_L1": worst["L1"] if worst else None,
                   "max_abs_dmu": max(abs(e["dmu"]) for e in errs) if errs else None,
                   "max_abs_dmass": max(abs(e["dmass"]) for e in errs) if errs else None,
                   "max_abs_dtail": max(abs(e["dtail"]) for e in errs) if errs else None,
                   "against": "plain FP64 GPU path (itself <= 1e-10 L1 vs the FP64 C oracle)"},
        "stage_ms": {k: round(v, 3) for k, v in stage.items()},
        "tuning": tuning["choice"] if tuning else None,
        "surrogate": sur,
        "ref_chain": ref,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def ref_chain_leg(log, zs=1.0, reps=3):
    """generate_pdf_lnmu's chain (the reference's 2-D chain) at z_s = 1, Planck,
    on the default ln mu grid linspace(-1, 9, 3201): seconds per PDF and the
    stage split for the ingredient ladder up to the default config="full"
    (halos + ellipticity + filaments + subhalos + the xi_lin term) and for
    "full" with the positive closure (closure="vp"), best of ``reps`` after
    one warm-up; then "full" and "full vp" once more on the reference's own
    body and tail grids against its stored pdf_zs1.json / pdf_zs1_vp.json
    (max |dP| over the peak, the reference's gate being 1e-9)."""
    import torch
    from sgl_analytic_amd import api, ref2d
    cp = api.check_cosmology({})
    x = np.linspace(-1.0, 9.0, 3201)
    out = {}
    full = dict(ell=True, fil=True, sub=True, bias=True)
    for name, kw in (("halo", {}), ("ell+fil", dict(ell=True, fil=True)),
                     ("ell+fil+sub", dict(ell=True, fil=True, sub=True)),
                     ("full", full), ("full vp", dict(full, vp=True))):
        ref2d.run(zs, cp, xi_out=x, **kw)
        torch.cuda.synchronize()
        best, stages, res = None, None, None
        for _ in range(reps):
            tm = {}
            t0 = time.perf_counter()
            res = ref2d.run(zs, cp, xi_out=x, timings=tm, **kw)
            dt = time.perf_counter() - t0
            if best is None or dt < best:
                best, stages = dt, tm
        g = res.grid
        out[name] = dict(seconds_per_pdf=round(best, 4), pdfs_per_s=round(1.0 / best, 3),
                         stage_s={k: round(v, 4) for k, v in stages.items()},
                         cells=res.pop.n_points, bins=res.pop.n_bins,
                         k_grid=[g["nk_k"], g["nk_g"]], out_grid=[g["n_kap"], g["n_gam"]],
                         sigma_kappa=res.sigma)
        log(f"[bench] reference chain z_s={zs} {name}: {best * 1e3:.1f} ms/PDF", out[name])
    gold = np.load(os.path.join(ROOT, "tests", "golden", "reference_pdfs.npz"))
    xb = np.concatenate([np.linspace(-1.0, 1.0, 1601), np.linspace(1.005, 9.0, 1600)])
    for name, kw, key in (("full", full, "zs1"), ("full vp", dict(full, vp=True), "zs1_vp")):
        res = ref2d.run(zs, cp, xi_out=xb, **kw)
        P0 = np.concatenate([gold[key + "_P_s"], gold[key + "_P_s_tail"]])
        out[name]["vs_stored_reference"] = dict(
            fixture=f"tests/reference/pdf_{key}.json",
            max_abs_dP_over_peak=float(np.max(np.abs(res.P - P0)) / gold[key + "_P_s"].max()))
        log(f"[bench] {name} vs the stored reference", out[name]["vs_stored_reference"])
    out["workload"] = ("generate_pdf_lnmu(z_s=1.0), Planck, default ln mu grid (3201 nodes), "
                       "reference grids (67 shells x 99 masses, adaptive cells, 961x300 lattice)")
    out["reference_cpu_seconds"] = {"note": "the reference's own records of its z_s = 1 "
                                    "full-config runs, nproc=4, Apple arm64: gauss 15.41 s "
                                    "(build 11.52, lam 0.30, invert 3.59; tests/reference/"
                                    "pdf_zs1.json t_stages), vp 34.91 s (build 11.46, lam 2.86, "
                                    "invert 20.59; pdf_zs1_vp.json)"}
    return out


def surrogate_leg(args, shape, sv, solver_rate, log, plain_rate=None):
    """Surrogate path on an inference batch of fresh LHS cosmologies: timed
    throughput (cheap solve + both networks + interpolation + MFMA inversion +
    diagnostics + gates), then outside the timing every instance is solved by
    the batched solver `sv`, every accepted surrogate PDF is checked against it,
    and the same batch without the networks (the cheap solve alone) shows what
    the networks gain."""
    import torch
    from sgl_analytic_amd import solver as S
    from sgl_analytic_amd import surrogate as SU
    path = SU.default_weights()
    if not os.path.exists(path):
        log(f"[bench] surrogate weights missing at {path}; train with "
            "`python -m sgl_analytic_amd.surrogate`")
        return {"status": "no trained weights"}
    nets = SU.load(path)
    Bs = args.sur_batch
    params = SU.lhs_params(Bs, seed=SEED + 2)
    n_z = S.slice_count(params[:, 9])
    rec = S.cosmo_records(params, n_z)
    x = torch.as_tensor(SU.features(params), dtype=torch.float32, device="cuda")
    inf = SU.Inference(nets, shape, Bs, n_z_max=shape.n_z_max)
    inf.set_records(rec)
    if not args.no_tune:
        log("[bench] surrogate cheap-solve launch shapes", inf.tune())
    inf.run(x)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        p, d, acc, y, dis = inf.run(x)
    torch.cuda.synchronize()
    t_sur = (time.perf_counter() - t0) / reps
    rate = Bs / t_sur
    p_sur = p.cpu().numpy().astype(np.float64)
    acc = acc.cpu().numpy()
    log(f"[bench] surrogate B={Bs}: {t_sur * 1e3:.1f} ms -> {rate:.0f} PDFs/s, accept {acc.mean():.4f}")
    base = {"pdfs_per_s": round(rate, 1), "batch": Bs, "ms_per_batch": round(t_sur * 1e3, 2),
            "n_nets": len(nets), "accept_rate": round(float(acc.mean()), 5),
            "accepted_pdfs_per_s": round(rate * float(acc.mean()), 1), "seed": SEED + 2}
    if args.no_sur_verify:
        return base
    p0, _, _, _, _ = inf.run(x, use_network=False)
    torch.cuda.synchronize()
    p_cheap = p0.cpu().numpy().astype(np.float64)
    del inf
    torch.cuda.empty_cache()
    xo = shape.out_grid()
    L1 = np.zeros(Bs)
    L1c = np.zeros(Bs)
    Bv = sv.B
    for a in range(0, Bs, Bv):
        b = min(Bs, a + Bv)
        r = rec[a:b]
        if b - a < Bv:
            r = np.concatenate([r, np.repeat(r[-1:], Bv - (b - a), 0)])
        pr, _ = sv.run(r)
        torch.cuda.synchronize()
        pr = pr[:b - a].float().cpu().numpy().astype(np.float64)
        L1[a:b] = np.trapezoid(np.abs(p_sur[a:b] - pr), xo, axis=1)
        L1c[a:b] = np.trapezoid(np.abs(p_cheap[a:b] - pr), xo, axis=1)
    n_acc = int(acc.sum())
    ok = L1[acc] <= 1e-3
    blended = Bs / (t_sur + (Bs - n_acc) / solver_rate)
    log(f"[bench] surrogate verified: {int(ok.sum())}/{n_acc} accepted within 1e-3 (max "
        f"{L1[acc].max() if n_acc else float('nan'):.3g}); L1 median with nets "
        f"{np.median(L1):.3g}, without {np.median(L1c):.3g}")
    fp64 = surrogate_fp64_check(args, shape, rec, p_sur, acc, dis.cpu().numpy(), log)
    return {
        **base,
        "blended_pdfs_per_s_with_solver_fallback": round(blended, 1),
        "blended_pdfs_per_s_with_plain_fp64_fallback":
            round(Bs / (t_sur + (Bs - n_acc) / plain_rate), 2) if plain_rate else None,
        "verified": {"against": "batched solver, every instance (itself <= 1e-3 L1 vs FP64)",
                     "accepted": n_acc, "accepted_within_1e3": int(ok.sum()),
                     "accepted_L1_max": float(L1[acc].max()) if n_acc else None,
                     "accepted_L1_median": float(np.median(L1[acc])) if n_acc else None},
        "verified_fp64": fp64,
        "network_gain": {"L1_median_with": float(np.median(L1)),
                         "L1_median_without": float(np.median(L1c)),
                         "L1_p90_with": float(np.quantile(L1, 0.9)),
                         "L1_p90_without": float(np.quantile(L1c, 0.9)),
                         "within_1e3_with": float((L1 <= 1e-3).mean()),
                         "within_1e3_without": float((L1c <= 1e-3).mean())},
    }


def surrogate_fp64_check(args, shape, rec, p_sur, acc, dis, log, chunk=64):
    """Accepted surrogate PDFs against the plain FP64 path: the worst
    --sur-fp64-frac of them by the networks' disagreement, plus
    --sur-fp64-random drawn from the rest (seeded); L1 of each."""
    import torch
    from sgl_analytic_amd import solver as S
    ia = np.flatnonzero(acc)
    if ia.size == 0 or args.sur_fp64_frac <= 0:
        return None
    n_worst = int(np.ceil(args.sur_fp64_frac * ia.size))
    order = ia[np.argsort(-dis[ia], kind="stable")]
    rng = np.random.default_rng(SEED + 3)
    rand = rng.choice(order[n_worst:], size=min(args.sur_fp64_random, order.size - n_worst),
                      replace=False)
    pick = np.concatenate([order[:n_worst], rand])
    xo = shape.out_grid()
    pl = S.Solver(shape, chunk, path="exact")
    l1 = np.zeros(pick.size)
    t0 = time.perf_counter()
    for a in range(0, pick.size, chunk):
        ids = pick[a:a + chunk]
        r = rec[ids]
        if ids.size < chunk:
            r = np.concatenate([r, np.repeat(r[-1:], chunk - ids.size, 0)])
        pr, _ = pl.run(r)
        torch.cuda.synchronize()
        l1[a:a + ids.size] = np.trapezoid(np.abs(p_sur[ids] - pr[:ids.size].cpu().numpy()), xo,
                                          axis=1)
        log(f"[bench] surrogate FP64 check {a + ids.size}/{pick.size}")
    del pl
    torch.cuda.empty_cache()
    out = {"against": "plain FP64 GPU path", "worst_by_disagreement": n_worst,
           "worst_fraction": args.sur_fp64_frac, "random": int(rand.size),
           "L1_max": float(l1.max()), "L1_max_worst": float(l1[:n_worst].max()),
           "L1_max_random": float(l1[n_worst:].max()) if rand.size else None,
           "L1_median": float(np.median(l1)), "within_1e3": int((l1 <= 1e-3).sum()),
           "seconds": round(time.perf_counter() - t0, 1)}
    log("[bench] surrogate FP64 check", out)
    return out


def precision_study(args, B):
    """Both inversion precisions on one batch; every checked member is compared
    with the FP64 plain path (L1, max |dp|, d<mu>, d mass, d P(mu>2))."""
    import torch
    from sgl_analytic_amd import solver as S
    from sgl_analytic_amd._lib import PREC_BF16X3, PREC_FP32
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    params = lhs_params(B)
    n_z = S.slice_count(params[:, 9])
    shape = S.Shape(n_z_max=int(n_z.max()))
    rec = S.cosmo_records(params, n_z)
    x = shape.out_grid()
    rng = np.random.default_rng(SEED + 1)
    idx = np.sort(rng.choice(B, size=min(args.precision_study, B), replace=False))
    out = {"batch": B, "members": len(idx)}
    runs = {}
    for name, pr in (("fp32", PREC_FP32), ("bf16x3", PREC_BF16X3)):
        sv = S.Solver(shape, B, path="binned", precision=pr)
        p, d = sv.run(rec)
        torch.cuda.synchronize()
        runs[name] = (p.float().cpu().numpy().astype(np.float64), d.cpu().numpy())
        del sv
        torch.cuda.empty_cache()
    pl = S.Solver(shape, 1, path="exact")
    rows = {k: [] for k in runs}
    for n, i in enumerate(idx):
        p1, d1 = pl.run(rec[i:i + 1])
        torch.cuda.synchronize()
        pe, de = p1[0].cpu().numpy(), d1[0].cpu().numpy()
        for k, (pb, db) in runs.items():
            rows[k].append((trapz_l1(x, pb[i], pe), float(np.max(np.abs(pb[i] - pe))),
                            abs(db[i, 5] - de[5]), abs(db[i, 0] - de[0]), abs(db[i, 6] - de[6])))
        if n % 16 == 0:
            print(f"[study] {n + 1}/{len(idx)} checked", file=sys.stderr, flush=True)
    for k, r in rows.items():
        a = np.array(r)
        out[k] = {"L1_median": float(np.median(a[:, 0])), "L1_max": float(a[:, 0].max()),
                  "maxabs_max": float(a[:, 1].max()), "dmu_max": float(a[:, 2].max()),
                  "dmass_max": float(a[:, 3].max()), "dtail_max": float(a[:, 4].max())}
    pf, pb = runs["fp32"][0], runs["bf16x3"][0]
    out["fp32_vs_bf16x3_L1_max"] = float(max(trapz_l1(x, pf[i], pb[i]) for i in range(B)))
    print(json.dumps({"precision_study": out}), flush=True)


def stage_times(sv, S):
    """One eager solve with HIP events around every library call (same stream);
    the binning stage is timed as its two halves (sgl_bin_pass, sgl_bin_reduce)."""
    import torch
    L, B, cs, st, P = sv.L, sv.B, ctypes.byref(sv.cs), S._stream(), S._ptr
    names, evs = [], [torch.cuda.Event(enable_timing=True)]
    evs[0].record()

    def mark(name):
        e = torch.cuda.Event(enable_timing=True)
        e.record()
        names.append(name)
        evs.append(e)

    L.sgl_tables_build(P(sv.records), B, cs, P(sv.tables), P(sv.status), st); mark("tables")
    L.sgl_bin_pass(P(sv.tables), B, cs, P(sv.hist), st); mark("bin_pass")
    L.sgl_bin_reduce(P(sv.tables), B, cs, P(sv.hist), st); mark("bin_reduce")
    L.sgl_lnphi_binned(P(sv.tables), P(sv.hist), B, cs, P(sv.lam), P(sv.amp), st); mark("lnphi_binned")
    L.sgl_quadform(P(sv.tables), P(sv.amp), B, cs, 1, P(sv.psi2c), st); mark("quadform")
    L.sgl_psi2_interp(P(sv.psi2c), B, cs, P(sv.psi2), st); mark("psi2_interp")
    L.sgl_phi_pack(P(sv.lam), P(sv.psi2), P(sv.tables), B, cs, sv.precision, P(sv.phi), st); mark("phi_pack")
    if not sv._filon_ready:
        L.sgl_filon_build(P(sv.tables), B, cs, sv.precision, P(sv.E), st)
        sv._filon_ready = True
        mark("filon_build(once)")
    L.sgl_invert_gemm(P(sv.E), P(sv.phi), B, cs, sv.precision, P(sv.p), st); mark("invert_gemm")
    L.sgl_diagnostics(P(sv.p), 0, B, cs, P(sv.diag), st); mark("diagnostics")
    torch.cuda.synchronize()
    return {n: evs[i].elapsed_time(evs[i + 1]) for i, n in enumerate(names)}


def bin_pass_ms(sv, S, reps=3):
    """Average duration of the dominant kernel (k_bin_cells, the only kernel
    sgl_bin_pass launches), HIP events on the library stream around each call."""
    import torch
    L, B, cs, st, P = sv.L, sv.B, ctypes.byref(sv.cs), S._stream(), S._ptr
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        L.sgl_bin_pass(P(sv.tables), B, cs, P(sv.hist), st)
    b.record()
    b.synchronize()
    L.sgl_bin_reduce(P(sv.tables), B, cs, P(sv.hist), st)   # leave the histogram complete
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def roofline(t_bin_ms, shape, B, n_z):
    """Roofline of the dominant kernel, k_bin_cells: VALU-bound (its VALU issue
    is saturated; its four LDS atomics per cell alone would take about half its
    time; it moves ~15 GB per dispatch, 2 % of HBM bandwidth), priced as its
    algorithmic operations per cell against the non-packed FP32 vector peak;
    HBM traffic from the committed FETCH/WRITE counters."""
    cells = float(np.sum(n_z)) * shape.n_m * shape.n_b
    flops = CELL_FLOPS * cells
    ach = flops / (t_bin_ms * 1e-3) / 1e12
    default = (B == 4096 and shape.n_m == 256 and shape.n_b == 512)
    return {"kernel": "k_bin_cells (sgl_bin_pass)", "bound": "valu", "achieved": round(ach, 3),
            "peak": PEAK["fp32"], "unit": "TFLOP/s", "frac": round(ach / PEAK["fp32"], 4),
            "traffic": BIN_TRAFFIC_B4096 if default else None,
            "ms": round(t_bin_ms, 3),
            "algorithmic": f"{CELL_FLOPS} FP ops per cell (source count of the formula) x "
                           f"{cells:.4e} cells",
            "hbm_GBps": round(BIN_TRAFFIC_B4096 / (t_bin_ms * 1e-3) / 1e9, 1) if default else None}


def cpu_baseline(log):
    """The FP64 C oracle (OpenMP) on one z_s = 1 Planck instance of the 64 x 256 x
    512 shape, with the k loop sampled (256 of 8192 nodes) and scaled to a PDF."""
    try:
        from oracle import chain
        from oracle.cosmo import Cosmology
    except Exception as e:  # pragma: no cover
        log("[bench] cpu baseline unavailable:", e)
        return None
    threads = len(os.sched_getaffinity(0))
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    sh = chain.Shape()
    t0 = time.perf_counter()
    cos = Cosmology()
    T = chain.rows(cos, 1.0, sh)
    t_rows = time.perf_counter() - t0
    t0 = time.perf_counter()
    xi, ws, kap, gam = chain.cells(T["rows"], sh)
    t_cells = time.perf_counter() - t0
    kq = sh.k_nodes()[1:]
    ksamp = kq[:: sh.n_k // 256]
    t0 = time.perf_counter()
    chain.lnphi(xi, ws, T["rows"][:, :, 5], ksamp)
    t_lam = (time.perf_counter() - t0) * (sh.n_k / ksamp.size)
    t0 = time.perf_counter()
    chain.invert(sh.out_grid()[::8], sh.k_nodes(), np.ones(sh.n_k + 1, complex))
    t_inv = (time.perf_counter() - t0) * 8
    total = t_rows + t_cells + t_lam + t_inv
    log(f"[bench] cpu oracle: rows {t_rows:.2f}s cells {t_cells:.2f}s lnphi(scaled) {t_lam:.1f}s "
        f"invert(scaled) {t_inv:.2f}s")
    return {"value": round(1.0 / total, 5), "unit": "PDFs/s", "cores": int(os.environ.get(
        "OMP_NUM_THREADS", threads)), "kind": "port",
        "sample": "FP64 C oracle, 1 instance (64x256x512 cells, z_s=1): full tables and cells, "
                  "ln phi on 256 of 8192 k nodes and inversion on 256 of 2048 outputs, "
                  "both scaled linearly to the full grid"}


if __name__ == "__main__":
    main()
