"""Command line: the magnification PDF of one source redshift and cosmology,
written as the reference's JSON record (sgl_analytic/cli.py:24-78).

  python -m sgl_analytic_amd --zs 1
  python -m sgl_analytic_amd --zs 3 --s8 0.75 --Om 0.30 --out output/pdf_zs3_s8lo.json
  python -m sgl_analytic_amd --zs 1 --growth-mode auto --radiation none

The record holds xi = ln mu on linspace(-1, 1, 1601) with P_s = dP/dln mu
there, xi_tail / P_s_tail on linspace(1.005, 9, 1600) (same normalisation),
the moments, diagnostics and clipped moments, and the rest of the
reference's key set (api.reference_schema).  Computed on the MI355X FP64
plain path.
"""
from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np

from .api import EXTENSIONS, P, clipped_moments, run

XI_BODY = np.linspace(-1.0, 1.0, 1601)
XI_TAIL = np.linspace(1.005, 9.0, 1600)


def record(zs, config="full", cosmo=None, closure="gauss", edge="r200"):
    """The reference's output record for one PDF (cli.py:57-72)."""
    xall = np.concatenate([XI_BODY, XI_TAIL])
    p, out = run(zs, config, cosmo=cosmo, xi_out=xall, closure=closure, edge=edge)
    nb = XI_BODY.size
    out.update(xi=XI_BODY.tolist(), P_s=p[:nb].tolist(), xi_tail=XI_TAIL.tolist(),
               P_s_tail=p[nb:].tolist())
    out["moment_interval"] = [out["grid"]["xi_min"], out["grid"]["xi_max"]]
    out["diagnostics"] = dict(grid_mass=float(np.trapezoid(p, xall)),
                              negative_mass=float(np.trapezoid(np.maximum(-p, 0), xall)),
                              minimum_density=float(p.min()))
    out["clipped"] = clipped_moments(XI_BODY, p[:nb])
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--z-s", "--zs", dest="zs", type=float, required=True, help="source redshift")
    ap.add_argument("--h", type=float, default=P["h"])
    ap.add_argument("--Om", type=float, default=P["Om"])
    ap.add_argument("--sigma8", "--s8", dest="s8", type=float, default=P["s8"])
    for name in ("Ob", "ns", "zeq"):
        ap.add_argument(f"--{name}", type=float, default=P[name])
    ap.add_argument("--Ok", "--Omega-K", "--Omega_K", dest="Ok", type=float,
                    default=EXTENSIONS["Ok"], help="curvature: > 0 open, < 0 closed (default flat)")
    ap.add_argument("--w0", type=float, default=EXTENSIONS["w0"],
                    help="dark energy w(a) = w0 + wa (1 - a) (default -1)")
    ap.add_argument("--wa", type=float, default=EXTENSIONS["wa"], help="CPL slope (default 0)")
    ap.add_argument("--growth-mode", "--growth_mode", dest="growth_mode",
                    choices=("auto", "ode", "legacy"), default=EXTENSIONS["growth_mode"],
                    help='growth: "ode" (default), "legacy" (LambdaCDM) or "auto"')
    ap.add_argument("--radiation", choices=("engine", "none"), default=EXTENSIONS["radiation"],
                    help='"engine" (default): Omega_r = Om/(1+zeq); "none": no radiation')
    ap.add_argument("--closure", choices=("gauss", "vp"), default="gauss",
                    help="clustering closure: gauss (two-point term) or vp (positive closure)")
    ap.add_argument("--config", default="full",
                    help='lens model: "full" (default), "halo", "-bias", ... '
                         '(pass minus arms as --config=-bias)')
    ap.add_argument("--edge", choices=("r200", "rvir"), default="r200",
                    help="halo truncation radius (default r200)")
    ap.add_argument("--nproc", type=int, default=1,
                    help="accepted for the reference's interface; the solve runs on one GPU")
    ap.add_argument("--out", default=None, help="output JSON (default output/pdf_zs<zs>.json)")
    a = ap.parse_args(argv)
    if a.nproc < 1:
        ap.error("nproc must be a positive integer")
    try:
        out = record(a.zs, a.config,
                     cosmo=dict(h=a.h, Om=a.Om, s8=a.s8, Ob=a.Ob, ns=a.ns, zeq=a.zeq, Ok=a.Ok,
                                w0=a.w0, wa=a.wa, growth_mode=a.growth_mode,
                                radiation=a.radiation),
                     closure=a.closure, edge=a.edge)
    except ValueError as exc:
        ap.error(str(exc))
    dest = Path(a.out) if a.out else Path("output") / f"pdf_zs{a.zs:g}.json"
    dest.parent.mkdir(parents=True, exist_ok=True)
    with dest.open("w") as stream:
        json.dump(out, stream, allow_nan=False)
    c = out["clipped"]
    print(f"zs={a.zs:g} clipped var={c['var']:.5e} skew={c['skew']:.3f}  "
          f"{out['seconds']:.2f} s -> {dest}")


if __name__ == "__main__":
    main()
