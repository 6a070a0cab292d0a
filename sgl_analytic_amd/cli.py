"""Command line: one magnification PDF written as the reference's JSON record.

  python -m sgl_analytic_amd --zs 1 --config +fil
  python -m sgl_analytic_amd --zs 3 --s8 0.75 --Om 0.30 --config halo --out pdf_zs3.json

The options, their aliases and defaults are the reference's command line
(sgl_analytic/cli.py:24-55); the record is api.run's (the keys of the
reference's run() record) on the body grid linspace(-1, 1, 1601) and the tail
grid linspace(1.005, 9, 1600), inverted together, plus the diagnostics, the
clipped moments and the moment interval (cli.py:57-72).  --edge is added
(pipeline.run's edge).
"""
from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np

from .api import EXTENSIONS, P, clipped_moments, run

BODY = (-1.0, 1.0, 1601)
TAIL = (1.005, 9.0, 1600)
COSMO_KEYS = ("h", "Om", "s8", "Ob", "ns", "zeq", "Ok", "w0", "wa", "growth_mode", "radiation")

# (flags, destination, type, default, choices); default None with type float = required
OPTIONS = (
    (("--z-s", "--zs"), "zs", float, None, None),
    (("--h",), "h", float, P["h"], None),
    (("--Om",), "Om", float, P["Om"], None),
    (("--sigma8", "--s8"), "s8", float, P["s8"], None),
    (("--Ob",), "Ob", float, P["Ob"], None),
    (("--ns",), "ns", float, P["ns"], None),
    (("--zeq",), "zeq", float, P["zeq"], None),
    (("--Ok", "--Omega-K", "--Omega_K"), "Ok", float, EXTENSIONS["Ok"], None),
    (("--w0",), "w0", float, EXTENSIONS["w0"], None),
    (("--wa",), "wa", float, EXTENSIONS["wa"], None),
    (("--growth-mode", "--growth_mode"), "growth_mode", str, EXTENSIONS["growth_mode"],
     ("auto", "ode", "legacy")),
    (("--radiation",), "radiation", str, EXTENSIONS["radiation"], ("engine", "none")),
    (("--closure",), "closure", str, "gauss", ("gauss", "vp")),
    (("--config",), "config", str, "full", None),
    (("--edge",), "edge", str, "r200", ("r200", "rvir")),
    (("--nproc",), "nproc", int, 1, None),
    (("--out",), "out", str, "", None),
)


def parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="python -m sgl_analytic_amd", description=__doc__,
                                 formatter_class=argparse.RawDescriptionHelpFormatter)
    for flags, dest, typ, default, choices in OPTIONS:
        ap.add_argument(*flags, dest=dest, type=typ, choices=choices,
                        required=default is None, default=default)
    return ap


def record(zs, config="full", cosmo=None, closure="gauss", edge="r200") -> dict:
    """One solve on the body and tail grids, split into the record's arrays."""
    body, tail = np.linspace(*BODY), np.linspace(*TAIL)
    p, rec = run(zs, config, cosmo=cosmo, xi_out=np.concatenate([body, tail]), closure=closure,
                 edge=edge)
    p = np.asarray(p, dtype=float)
    rec.update(xi=body.tolist(), P_s=p[:body.size].tolist(), xi_tail=tail.tolist(),
               P_s_tail=p[body.size:].tolist())
    x = np.concatenate([body, tail])
    rec["diagnostics"] = dict(grid_mass=float(np.trapezoid(p, x)),
                              negative_mass=float(np.trapezoid(np.clip(-p, 0, None), x)),
                              minimum_density=float(p.min()))
    rec["clipped"] = clipped_moments(body, p[:body.size])
    rec["moment_interval"] = [rec["grid"]["xi_min"], rec["grid"]["xi_max"]]
    return rec


def main(argv=None):
    ap = parser()
    a = ap.parse_args(argv)
    if a.nproc < 1:
        ap.error("nproc must be a positive integer")
    cosmo = {k: getattr(a, k) for k in COSMO_KEYS}
    try:
        rec = record(a.zs, a.config, cosmo=cosmo, closure=a.closure, edge=a.edge)
    except (ValueError, NotImplementedError) as exc:
        ap.error(str(exc))
    dest = Path(a.out or f"output/pdf_zs{a.zs:g}.json")
    dest.parent.mkdir(parents=True, exist_ok=True)
    dest.write_text(json.dumps(rec, allow_nan=False))
    c = rec["clipped"]
    print(f"zs={a.zs:g} sigma_kappa={rec['sigma_kappa']:.5f} clipped var={c['var']:.5e} "
          f"skew={c['skew']:.3f} {rec['seconds']:.2f} s -> {dest}")
    return rec
