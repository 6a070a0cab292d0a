"""Drop-in API: generate_pdf_lnmu / generate_pdf with the reference's
signatures, defaults, argument checks and error behaviour
(sgl_analytic/api.py:11-113, pipeline.py:36-94, 139-164), computed by the
MI355X kernels.

``run`` (and through it generate_pdf_lnmu / generate_pdf) computes the
reference's own chain, pipeline.run (pipeline.py:114-274): the population
on the reference's shells, engine mass grid and adaptive impact-parameter
cells, binned into the 961 x 300 (kappa, gamma) lattice; the 2-D
characteristic function ln Phi(k_kappa, k_gamma) on the width-relative grid
with the measured k-reach; the Hankel + Fourier inversion to P3(kappa,
gamma); the constant-mu arc integral and the normalisation
(sgl_analytic_amd/ref2d.py over the sgl_ref_* kernels).  Ingredients built
on this chain: spherical truncated NFW halos, "ell" (pseudo-elliptical
halos), "fil" (filaments) and "sub" (subhalos on their hosts, the exact
decorated-host sector; with edge "r200" only) and "bias" (the xi_lin
Limber two-point term, clustering="xilin", the pipeline's default); edge
"r200" (default) or "rvir"; closure "gauss" (the Gaussian pair term) or
"vp" (the positive variance-preserving closure, pos_closure.install).  The
default config="full" therefore runs every ingredient, as the reference does.

``run_kernel_shape`` is the host driver of the kernel-shape chain (64
slices uniform in comoving distance x 256 masses x 512 impact parameters,
the 1-D ln mu jump chain of sgl.Lambda_of_k / P_of_xi, sgl.py:961-1012) that
the batched solver and the surrogate use; it is not the reference's model.

Batched entry point (no reference counterpart): generate_pdf_lnmu_batch.
"""
from __future__ import annotations

import math
import time
import warnings

import numpy as np

P = dict(h=0.674, Om=0.315, s8=0.811, Ob=0.0493, ns=0.965, zeq=3402.0)  # pipeline.py:20
EXTENSIONS = dict(Ok=0.0, w0=-1.0, wa=0.0, growth_mode="ode", radiation="engine")
INGREDIENTS = ("ell", "fil", "sub", "bias")
MODELLED = {"bias", "fil", "ell"}   # ingredients of the kernel-shape chain (run_kernel_shape)
MODELLED_2D = {"fil", "ell", "sub", "bias"}  # ingredients of the reference chain (run)
# the subhalo model the chain builds: the reference's defaults for edge "r200"
# (pipeline.py:171-175, subhalos.py:541-545), as its run record states them
SUB_KW = dict(Mhost_min=1e9, npsi=28, nd=90, nth=48, split=10.0, ntheta=4, nr=40, prune=0.0,
              pair_angle="uniform", sector="exact", exact_kw=dict(method="moments"),
              clump_edge="r200", clump_mass="vir", clump_shape="circular", nphi_c=12)
DEFAULT_SLICES = 64


def _grid(values, lo, hi, count, name):
    """api.py:11-23: a finite, 1-D, strictly increasing grid with >= 2 points."""
    if values is None:
        if isinstance(count, bool) or not isinstance(count, (int, np.integer)) or count < 2:
            raise ValueError(f"n_{name} must be an integer >= 2")
        if not np.isfinite(lo) or not np.isfinite(hi) or lo >= hi:
            raise ValueError(f"{name}_min and {name}_max must be finite and increasing")
        values = np.linspace(lo, hi, count)
    values = np.asarray(values, dtype=float)
    if values.ndim != 1 or values.size < 2 or not np.all(np.isfinite(values)):
        raise ValueError(f"{name} must be a finite one-dimensional grid with >= 2 points")
    if np.any(np.diff(values) <= 0):
        raise ValueError(f"{name} must be strictly increasing")
    return values


def on(cfg):
    """pipeline.py:78-94: ingredient set of a config string or iterable."""
    ing = set(INGREDIENTS)
    if not isinstance(cfg, str):
        s = set(cfg)
        if s - ing:
            raise ValueError(f"unknown ingredient(s) {sorted(s - ing)}; choose from {INGREDIENTS}")
        return s
    if cfg == "halo":
        return set()
    if cfg == "full":
        return ing
    if len(cfg) < 2 or cfg[0] not in "+-" or cfg[1:] not in ing:
        raise ValueError(f"unknown configuration {cfg!r}; use full, halo, or +/-{INGREDIENTS}")
    return {cfg[1:]} if cfg[0] == "+" else ing - {cfg[1:]}


def check_cosmology(cosmo=None):
    """pipeline.make_cosmology's checks (pipeline.py:41-72) and the domain of
    sgl.Cosmology (sgl.py:121-151); returns the completed parameter dict."""
    supplied = dict(cosmo or {})
    if "sigma8" in supplied:
        if "s8" in supplied:
            raise ValueError("supply sigma8 or s8, not both")
        supplied["s8"] = supplied.pop("sigma8")
    cp = {**P, **EXTENSIONS, **supplied}
    bad = set(cp) - set(P) - set(EXTENSIONS)
    if bad:
        raise ValueError(f"unknown cosmology parameter(s) {sorted(bad)}")
    for key in P:
        value = cp[key]
        if not np.isscalar(value) or not np.isfinite(value) or value <= 0:
            raise ValueError(f"{key} must be a finite positive number")
    for key in ("Ok", "w0", "wa"):
        value = cp[key]
        if isinstance(value, bool) or not np.isscalar(value) or not np.isfinite(value):
            raise ValueError(f"{key} must be a finite number")
    if cp["growth_mode"] not in ("auto", "ode", "legacy"):
        raise ValueError("growth_mode must be 'auto', 'ode' or 'legacy'")
    if cp["radiation"] in ("none", False):
        cp["radiation"] = None
    if cp["radiation"] not in ("engine", None):
        raise ValueError("radiation must be 'engine' or None")
    if not 0 < cp["Ob"] < cp["Om"] <= 1:
        raise ValueError("require 0 < Ob < Om <= 1")
    w0, wa = float(cp["w0"]), float(cp["wa"])
    if not -2.0 <= w0 <= 1.0:
        raise ValueError("dark energy requires -2 <= w0, growth_w0 <= 1")
    if w0 + wa >= 0.0:
        raise ValueError("dark energy requires w0 + wa < 0 (matter-dominated early universe)")
    Om, Ok = float(cp["Om"]), float(cp["Ok"])
    Or = Om / (1.0 + float(cp["zeq"])) if cp["radiation"] == "engine" else 0.0
    OL = 1.0 - Om - Ok - Or
    z = 1100.0
    if OL / Om * (1 + z) ** (3 * (w0 + wa)) * math.exp(-3 * wa * z / (1 + z)) > 1e-2:
        raise ValueError("dark energy exceeds de_rec_cap of the matter density at z = 1100")
    zp1 = 1.0 / np.geomspace(1e-6, 1.0, 4096)
    de = OL * zp1 ** (3 * (1 + w0 + wa)) * np.exp(-3 * wa * (1 - 1 / zp1))
    if not np.all(Om * zp1 ** 3 + Or * zp1 ** 4 + Ok * zp1 ** 2 + de > 0):
        raise ValueError("E(z)^2 <= 0 somewhere in 0 <= z <= 1e6: outside the supported domain")
    # growth: "auto" is legacy for LambdaCDM and the ODE otherwise, with the
    # pipeline's warning when it resolves to the ODE (sgl.py:134-140,
    # pipeline.py:166-169)
    lambda_de = w0 == -1.0 and wa == 0.0
    if cp["growth_mode"] == "legacy" and not lambda_de:
        raise ValueError("legacy growth is only valid for LambdaCDM")
    if cp["growth_mode"] == "auto":
        cp["growth_mode"] = "legacy" if lambda_de else "ode"
        if not lambda_de:
            warnings.warn("growth_mode='auto' resolved to 'ode' for non-Lambda dark energy, "
                          "which is not continuous with 'auto' at w = -1; compare against a "
                          "growth_mode='ode' Lambda run.", stacklevel=3)
    return cp


_SOLVERS = {}


def _solver(n_z):
    from .solver import Shape, Solver
    key = ("exact", n_z)
    if key not in _SOLVERS:
        _SOLVERS[key] = Solver(Shape(n_z_max=n_z), 1, path="exact")
    return _SOLVERS[key]


ZINT_MODES = ("engine", "midpoint", "gauss")
CLUSTERING_MODES = ("xilin", "linear", "limber", "closure")


def run(z_s, config="full", cosmo=None, xi_out=None, closure="gauss", edge="r200",
        clustering="xilin", zint="midpoint", zint_n=2, nproc=1):
    """pipeline.run (pipeline.py:114-274) on the device: the reference's 2-D
    chain for one cosmology; returns (P_s on xi_out, info).  info carries the
    keys of the reference's run() record (pipeline.py:243-274).
    The arguments are checked as pipeline.run checks them (pipeline.py:139-164;
    the same ValueErrors).  clustering: "xilin" (the default: the xi_lin pair
    term) or "limber" (the Gaussian large-scale-structure term of
    limber_lss_var, pipeline.py:195-202); without the "bias" ingredient every
    clustering mode but "closure" gives the same chain (no two-point term).
    zint: "midpoint" (the default), "engine" (the engine grid's shells below
    z_s, one node at the upper edge) or "gauss" (the midpoint shells at
    zint_n Gauss-Legendre nodes each, up to 208 nodes in all); engine and
    gauss without subhalos and the vp closure.  Not built,
    NotImplementedError: "linear" with "bias" (the 20 Mpc slice-covariance
    Cox term), "closure", edge "kthr", zint other than midpoint with
    subhalos or the vp closure, subhalos with edge "rvir".
    nproc: the reference's host worker count, checked and recorded (the
    device needs no host pool)."""
    if clustering not in CLUSTERING_MODES:
        raise ValueError(f"clustering must be 'xilin', 'linear', 'limber' or 'closure', "
                         f"got {clustering!r}")
    if not np.isscalar(z_s) or not np.isfinite(z_s) or z_s <= 0:
        raise ValueError("z_s must be a finite positive number")
    if isinstance(nproc, bool) or not isinstance(nproc, (int, np.integer)) or nproc < 1:
        raise ValueError("nproc must be a positive integer")
    if edge not in ("r200", "rvir", "kthr"):
        raise ValueError("edge must be 'r200', 'rvir' or 'kthr'")
    if closure not in ("gauss", "vp"):
        raise ValueError(f"xi_closure must be 'gauss' or 'vp', got {closure!r}")
    s = on(config)
    use_xi = clustering == "xilin" and "bias" in s
    if use_xi and edge == "kthr":
        raise ValueError("clustering='xilin' needs edge='r200' or 'rvir'")
    if closure != "gauss" and not use_xi:
        raise ValueError(f"xi_closure={closure!r} needs xilin clustering and the bias ingredient")
    if zint not in ZINT_MODES:
        raise ValueError(f"zint must be one of {ZINT_MODES}, got {zint!r}")
    if zint == "gauss" and int(zint_n) < 1:
        raise ValueError("zint_n must be >= 1")
    cp = check_cosmology(cosmo)
    if edge == "kthr":
        raise NotImplementedError("edge='kthr' (the engine's Gaussian faint-lens split) is not built")
    if clustering == "closure":
        raise NotImplementedError("clustering='closure' (the lognormal Hermite closure) is not built")
    if clustering == "linear" and "bias" in s:
        raise NotImplementedError("clustering='linear' with the bias ingredient (the slice-"
                                  "covariance Cox term) is not built; use 'xilin' or 'limber'")
    if zint != "midpoint" and ("sub" in s or closure == "vp"):
        raise NotImplementedError(f"zint={zint!r} is built without subhalos and the vp closure")
    if "sub" in s and edge != "r200":
        raise NotImplementedError("subhalos are built with edge='r200' only (clumps truncated at "
                                  "their r200, the reference's pairing for that edge)")
    lss = clustering == "limber" and "bias" in s
    from . import ref2d
    t0 = time.perf_counter()
    pop = ref2d.build(cp, float(z_s), ell="ell" in s, fil="fil" in s, edge=edge, sub="sub" in s,
                      bias=use_xi, vp=closure == "vp", lss=lss, zint=zint, zint_n=int(zint_n))
    t1 = time.perf_counter()
    # xi_out None: the CLI's body and tail grids (pipeline.py:211-215)
    x = (np.concatenate([np.linspace(-1.0, 1.0, 1601), np.linspace(1.005, 9.0, 1600)])
         if xi_out is None else np.asarray(xi_out, dtype=np.float64))
    res = ref2d.invert(pop, x)
    t2 = time.perf_counter()
    info = record_2d(cp, config, s, closure, edge,
                     dict(kbar=pop.kbar, N_expl=pop.N_expl, N_fil=pop.N_fil, n_bins=pop.n_bins,
                          n_points=pop.n_points),
                     dict(sigma=res.sigma, mean=res.mean, var=res.var, skew=res.skew,
                          sigmaDL=res.sigmaDL, grid=res.grid),
                     dict(build=t1 - t0, lam=0.0, invert=t2 - t1), clustering=clustering,
                     lss_var=pop.lss_var, nproc=int(nproc), zint=zint,
                     zint_n=int(zint_n) if zint == "gauss" else None)
    info["zs"] = float(z_s)
    if pop.vp is not None:
        v = pop.vp
        info["pos_closure"] = dict(kind="vp", eta=v.eta, nstate=48, shells=v.n_z,
                                   cells=v.cells, groups=v.n_groups,
                                   target_kappa=v.target_kappa, target_gamma=v.target_gamma,
                                   eval_kernel="MI355X HIP (sgl_ref_vp_lnr)")
    return res.P, info


def record_2d(cp, config, ingredients, closure, edge, pop, res, t_stages, clustering="xilin",
              lss_var=0.0, nproc=1, zint="midpoint", zint_n=None):
    """The run() record of the reference (pipeline.py:243-274; every top-level
    key of its stored PDFs) with the 2-D chain's values: pop = bookkeeping of
    the population (kbar, N_expl, N_fil, n_bins, n_points), res = sigma,
    moments and grid of the inversion, t_stages = stage times."""
    Or = cp["Om"] / (1.0 + cp["zeq"]) if cp["radiation"] == "engine" else 0.0
    de = dict(w0=cp["w0"], wa=cp["wa"], growth_mode=cp["growth_mode"],
              radiation=cp["radiation"], Omega_r=Or)
    return dict(
        t_stages=dict(t_stages), cfg=config if isinstance(config, str) else sorted(ingredients),
        zint=zint, zint_n=zint_n, clustering=clustering,
        xi_kperp=(dict(kmin=1e-4, kmax=1e4, per_decade=[2, 2, 12, 8, 6, 4, 4, 4], mmax=2)
                  if "bias" in ingredients and clustering == "xilin" else None), edge=edge,
        lss_var=float(lss_var),
        kap_floor=1e-8, kbar=float(pop["kbar"]), h=cp["h"], Om=cp["Om"], s8=cp["s8"],
        sigma8=cp["s8"], Ob=cp["Ob"], ns=cp["ns"], zeq=cp["zeq"], Ok=cp["Ok"], w0=cp["w0"],
        wa=cp["wa"], growth_mode=cp["growth_mode"], radiation=cp["radiation"], Omega_r=Or,
        dark_energy=de, nord=1, grid_scale=1.0, split=10.0, nproc=nproc,
        sub_kw=dict(SUB_KW) if "sub" in ingredients else None,
        pair_angle="uniform" if "sub" in ingredients else None, pair_scale=None, k3=False,
        exactS=False, s_taper="cos2",
        sigma_kappa=float(res["sigma"]), kthr=None, N_expl=float(pop["N_expl"]),
        N_fil=float(pop["N_fil"]), V_weak=0.0, mean=float(res["mean"]), var=float(res["var"]),
        skew=float(res["skew"]), sigmaDL=float(res["sigmaDL"]),
        grid={k: v for k, v in res["grid"].items() if isinstance(v, (int, float))},
        tail_box=True, fine=dict(k=[1e-7, 30.0, 480], g=[1e-9, 30.0, 300]), s_is=True,
        xi_clamp=False, xi_linear_m=0.0, xi_closure=closure,
        window="auto" if closure == "vp" else False, n_xi_wide=4000 if closure == "vp" else 500,
        pos_closure=None, direct_product_backend="MI355X HIP FP64 MFMA (libsgl_amd.so)",
        direct_product_fallback=None, s_seed=None, s_seed_effective=0, s_ess=None,
        is_design=None, pk_mod=None, pk_par={},
        provenance=dict(producer="sgl_analytic_amd api.run (MI355X, reference 2-D chain)"),
        ingredients=["halo"] + sorted(ingredients), n_bins=int(pop["n_bins"]),
        n_points=int(pop["n_points"]),
        backend="MI355X HIP (libsgl_amd.so), reference 2-D chain, FP64",
        seconds=float(sum(t_stages.values())))


def run_kernel_shape(z_s, config="full", cosmo=None, xi_out=None, closure="gauss",
                     n_z=DEFAULT_SLICES, edge="r200"):
    """Host driver of the kernel-shape chain (64 slices uniform in comoving
    distance, the 1-D ln mu jump chain) for one cosmology on the plain FP64
    device path; returns (p on xi_out, info).  Not the reference's model: the
    batched solver and the surrogate use this chain.
    edge: halo truncation, "r200" (the reference default) or "rvir"
    (pipeline.run's edge / build_population's trunc, sgl_full.py:1874-1880)."""
    import ctypes

    import torch

    from . import _lib
    from .solver import _ptr, _stream, cosmo_records

    if not np.isscalar(z_s) or not np.isfinite(z_s) or z_s <= 0:
        raise ValueError("z_s must be a finite positive number")
    if edge not in ("r200", "rvir"):
        if edge == "kthr" or isinstance(edge, (int, float)):
            raise NotImplementedError(f"edge={edge!r} is not part of this path (r200, rvir)")
        raise ValueError(f"edge must be 'r200' or 'rvir', got {edge!r}")
    if closure not in ("gauss", "vp"):
        raise ValueError(f"xi_closure must be 'gauss' or 'vp', got {closure!r}")
    s = on(config)
    cp = check_cosmology(cosmo)
    if closure != "gauss" and "bias" not in s:
        raise ValueError(f"xi_closure={closure!r} needs xilin clustering and the bias ingredient")
    explicit = s - MODELLED
    if explicit and config != "full":
        raise NotImplementedError(f"ingredient(s) {sorted(explicit)} are not part of this path")
    sv = _solver(int(n_z))
    sv_shape = sv.shape
    if (("bias" in s) != sv_shape.clustering or ("fil" in s) != sv_shape.filaments
            or edge != sv_shape.edge or ("ell" in s) != sv_shape.ell or closure != "gauss"):
        from dataclasses import replace
        from .solver import Solver
        key = ("exact", int(n_z), "bias" in s, "fil" in s, edge, "ell" in s)
        key += (closure,) if closure != "gauss" else ()
        if key not in _SOLVERS:
            _SOLVERS[key] = Solver(replace(sv_shape, clustering="bias" in s, filaments="fil" in s,
                                           edge=edge, ell="ell" in s), 1, path="exact",
                                   closure=closure)
        sv = _SOLVERS[key]
        sv_shape = sv.shape
    rec = cosmo_records(np.array([[cp["h"], cp["Om"], cp["s8"], cp["Ob"], cp["ns"], cp["zeq"],
                                   cp["Ok"], cp["w0"], cp["wa"], float(z_s)]]), int(n_z),
                        radiation=cp["radiation"], growth_mode=cp["growth_mode"])
    t0 = time.perf_counter()
    sv.run(rec)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    x = np.asarray(xi_out, dtype=np.float64)
    t = torch.as_tensor(x, device=sv.dev)
    p = torch.zeros(x.size, dtype=torch.float64, device=sv.dev)
    _lib.check(sv.L.sgl_invert_fp64_at(_ptr(sv.phi), _ptr(sv.tables), 1, ctypes.byref(sv.cs),
                                       _ptr(t), int(x.size), _ptr(p), _stream()),
               "sgl_invert_fp64_at")
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    d = sv.diag[0].cpu().numpy()
    st = int(sv.status.cpu()[0])
    if st & (_lib.STATUS_VP_TRANSITION | _lib.STATUS_VP_RATE):
        raise _lib.SglError(f"vp closure set-up failed (status {st:#x}): the transition balancing "
                            "did not converge or the rate variance left the 48-node range")
    info = reference_schema(cp, sv_shape, d, config, closure, edge, t_solve=t1 - t0,
                            t_invert=t2 - t1)
    if closure == "vp":
        eta, target = sv.vp_calibration()
        from .solver import vp_groups
        info["pos_closure"] = dict(kind="vp", eta=float(eta[0]), nstate=48, shells=int(n_z),
                                   eval_kernel="hip_chain", groups=int(n_z) * vp_groups(sv_shape),
                                   target_kappa=float(target[0]), target_gamma=None)
    info.update(zs=float(z_s), cfg=config, ingredients=["halo"] + sorted(s & MODELLED),
                clustering="slice-covariance" if "bias" in s else None, closure=closure,
                edge=edge,
                n_slices=int(n_z),
                kernel_grid=dict(xi_min=sv_shape.out_lo, xi_max=sv_shape.out_hi,
                                 n=sv_shape.n_out, mass=float(d[0]), mean=float(d[1]),
                                 var=float(d[2]), skew=float(d[3]), sigmaDL=float(d[4]),
                                 mu_mean=float(d[5]), p_mu_gt_2=float(d[6])),
                h=cp["h"], Om=cp["Om"], s8=cp["s8"], sigma8=cp["s8"], Ob=cp["Ob"], ns=cp["ns"],
                zeq=cp["zeq"], Ok=cp["Ok"], w0=cp["w0"], wa=cp["wa"],
                growth_mode=cp["growth_mode"], radiation=cp["radiation"],
                backend="MI355X HIP (libsgl_amd.so), FP64 plain path")
    return p.cpu().numpy(), info


def reference_schema(cp, shape, diag, config, closure, edge, t_solve, t_invert):
    """Every top-level key of the reference's run() record (pipeline.py:243-274,
    the key set of its stored PDFs), filled with this chain's values.  Keys
    of machinery this chain does not have (the 2-D grid sizing, the
    importance sampler, the subhalo and positive-closure settings, the weak
    arm) are present with None, False or 0 so that readers of the reference's
    files find them; the chain's own grid is under "grid" and "kernel_grid"."""
    Or = cp["Om"] / (1.0 + cp["zeq"]) if cp["radiation"] == "engine" else 0.0
    return dict(
        t_stages=dict(build=float(t_solve), lam=0.0, invert=float(t_invert)),
        zint="uniform-chi midpoint", zint_n=None,
        clustering="slice-covariance" if shape.clustering else None,
        xi_kperp=None, edge=edge, lss_var=0.0, kap_floor=None, kbar=None,
        h=cp["h"], Om=cp["Om"], s8=cp["s8"], Ob=cp["Ob"], ns=cp["ns"], zeq=cp["zeq"],
        Ok=cp["Ok"], w0=cp["w0"], wa=cp["wa"],
        dark_energy=dict(w0=cp["w0"], wa=cp["wa"], growth_w0=cp["w0"], growth_wa=cp["wa"],
                         growth_mode=cp["growth_mode"], radiation=cp["radiation"],
                         Omega_r=Or),
        nord=1, grid_scale=None, split=None, nproc=1, sub_kw=None, pair_angle=None,
        pair_scale=None, k3=None, exactS=False, s_taper=None, sigma_kappa=None, kthr=None,
        N_expl=None, N_fil=None, V_weak=0.0,
        mean=float(diag[1]), var=float(diag[2]), skew=float(diag[3]), sigmaDL=float(diag[4]),
        grid=dict(n_k=shape.n_k, k_min=shape.k_lo, k_max=shape.k_hi, n_m=shape.n_m,
                  n_b=shape.n_b, xi_min=shape.out_lo, xi_max=shape.out_hi, n_out=shape.n_out),
        tail_box=False, fine=None, s_is=False, xi_closure=closure, window=shape.window,
        n_xi_wide=None, pos_closure=None,
        direct_product_backend="MI355X HIP (libsgl_amd.so)", direct_product_fallback=None,
        s_seed=None, s_seed_effective=None, s_ess=None, is_design=None,
        xi_clamp=False, xi_linear_m=0.0, pk_mod=None, pk_par={},
        provenance=dict(producer="sgl_analytic_amd api.run (MI355X, 1-D ln mu jump chain)"),
        seconds=float(t_solve + t_invert))


def generate_pdf_lnmu(z_s, h=0.674, Om=0.315, sigma8=0.811, Ob=0.0493,
                      ns=0.965, zeq=3402.0, lnmu=None, *, Ok=0.0, w0=-1.0,
                      wa=0.0, growth_mode="ode", radiation="engine",
                      closure="gauss", lnmu_min=-1.0, lnmu_max=9.0,
                      n_lnmu=3201, config="full", nproc=1,
                      return_info=False):
    """Return (lnmu, dP_S/dlnmu) as NumPy arrays (api.py:26-84).

    Same arguments and checks as the reference; the PDF comes from the
    reference's 2-D chain on the MI355X (``run``).  The density is not renormalised to the user
    grid; negative residuals stay visible with a RuntimeWarning."""
    x = _grid(lnmu, lnmu_min, lnmu_max, n_lnmu, "lnmu")
    if isinstance(nproc, bool) or not isinstance(nproc, (int, np.integer)) or nproc < 1:
        raise ValueError("nproc must be a positive integer")
    p, out = run(z_s, config, cosmo=dict(h=h, Om=Om, sigma8=sigma8, Ob=Ob, ns=ns, zeq=zeq,
                                         Ok=Ok, w0=w0, wa=wa, growth_mode=growth_mode,
                                         radiation=radiation),
                 xi_out=x, closure=closure)
    p = np.asarray(p, dtype=np.float64)
    if not np.all(np.isfinite(p)):
        raise RuntimeError("PDF inversion returned non-finite values")
    negative_mass = float(np.trapezoid(np.maximum(-p, 0), x))
    out["diagnostics"] = dict(grid_mass=float(np.trapezoid(p, x)), negative_mass=negative_mass,
                              minimum_density=float(p.min()))
    g = out.get("grid", {})
    kg = out.get("kernel_grid", {})
    out["moment_interval"] = [g.get("xi_min", kg.get("xi_min")), g.get("xi_max", kg.get("xi_max"))]
    if negative_mass > 0:
        warnings.warn(f"PDF inversion has negative residuals (integrated mass "
                      f"{negative_mass:.3g}); inspect diagnostics before using tail probabilities.",
                      RuntimeWarning, stacklevel=2)
    return (x, p, out) if return_info else (x, p)


def generate_pdf(z_s, h=0.674, Om=0.315, sigma8=0.811, Ob=0.0493,
                 ns=0.965, zeq=3402.0, mu=None, *, Ok=0.0, w0=-1.0, wa=0.0,
                 growth_mode="ode", radiation="engine", closure="gauss",
                 mu_min=0.3, mu_max=30.0, n_mu=2000, config="full", nproc=1,
                 return_info=False):
    """Return (mu, p_S(mu)) with p(mu) = (dP/dlnmu)/mu (api.py:87-113)."""
    if mu is None:
        if mu_min <= 0 or mu_max <= 0:
            raise ValueError("mu bounds must be positive")
        x = _grid(None, np.log(mu_min), np.log(mu_max), n_mu, "mu")
        mu = np.exp(x)
    else:
        mu = _grid(mu, 0, 1, n_mu, "mu")
        if np.any(mu <= 0):
            raise ValueError("mu must be positive")
        x = np.log(mu)
    result = generate_pdf_lnmu(z_s, h, Om, sigma8, Ob, ns, zeq, x, Ok=Ok, w0=w0, wa=wa,
                               growth_mode=growth_mode, radiation=radiation, closure=closure,
                               config=config, nproc=nproc, return_info=return_info)
    return (mu, result[1] / mu, result[2]) if return_info else (mu, result[1] / mu)


def to_image_plane(xi, P_s):
    """Image-plane dP_I/dln mu from the source-plane PDF on xi = ln mu:
    dP_I is proportional to mu dP_S, normalised on the given grid
    (pipeline.to_image_plane, pipeline.py:106-111)."""
    xi, P_s = np.asarray(xi, float), np.asarray(P_s, float)
    P = np.exp(xi) * P_s
    return P / np.trapezoid(P, xi)


def clipped_moments(xi, P, clip=1.0):
    """Mean, variance and skewness of P restricted to |xi| <= clip and
    renormalised there (pipeline.clipped_moments, pipeline.py:277-284)."""
    xi, P = np.asarray(xi, float), np.asarray(P, float)
    m = np.abs(xi) <= clip
    x, p = xi[m], P[m]
    p = p / np.trapezoid(p, x)
    mean = np.trapezoid(x * p, x)
    var = np.trapezoid((x - mean) ** 2 * p, x)
    return dict(mean=float(mean), var=float(var),
                skew=float(np.trapezoid((x - mean) ** 3 * p, x) / var ** 1.5))


def generate_pdf_lnmu_batch(params, precision="fp32", graph=True, shape=None, *,
                            growth_mode="ode", radiation="engine", closure="gauss"):
    """Batched solve (no reference counterpart): params [B, 10] in the order
    (h, Om, sigma8, Ob, ns, zeq, Ok, w0, wa, z_s); slice count per instance
    ceil(24 (1 + ln(1 + z_s))); growth_mode ("ode", "legacy", "auto") and
    radiation ("engine", None) apply to the whole batch.  Returns (lnmu grid,
    p [B, n_out] float32 numpy, diagnostics [B, 8]).  closure="vp" (the
    positive closure) solves the batch on the plain FP64 path instead (one
    launch per step for a chunk of instances whose closure workspace fits
    VP_BATCH_BYTES; p is float64; precision and graph do not apply)."""
    if closure not in ("gauss", "vp"):
        raise ValueError(f"closure must be 'gauss' or 'vp', got {closure!r}")
    from ._lib import PREC_BF16X3, PREC_FP32
    from .solver import Shape, Solver, cosmo_records, slice_count
    params = np.atleast_2d(np.asarray(params, dtype=np.float64))
    modes = set()
    for row in params:
        cp = check_cosmology(dict(h=row[0], Om=row[1], s8=row[2], Ob=row[3], ns=row[4],
                                  zeq=row[5], Ok=row[6], w0=row[7], wa=row[8],
                                  growth_mode=growth_mode, radiation=radiation))
        modes.add(cp["growth_mode"])
        if not np.isfinite(row[9]) or row[9] <= 0:
            raise ValueError("z_s must be a finite positive number")
    if len(modes) > 1:
        raise ValueError("growth_mode='auto' resolves differently within the batch; "
                         "split it by dark-energy model")
    mode = modes.pop() if modes else "ode"
    n_z = slice_count(params[:, 9])
    sh = shape or Shape(n_z_max=int(n_z.max()))
    prec = PREC_FP32 if precision == "fp32" else PREC_BF16X3
    rec = cosmo_records(params, n_z, radiation=None if radiation in ("none", False, None)
                        else radiation, growth_mode=mode)
    if closure == "vp":
        return sh.out_grid(), *_vp_batch(sh, rec)
    sv = Solver(sh, params.shape[0], path="binned", precision=prec)
    sv.set_records(rec)
    if graph:
        sv.capture()
    p, d = sv.run()
    import torch
    torch.cuda.synchronize()
    return sh.out_grid(), p.cpu().numpy(), d.cpu().numpy()


VP_BATCH_BYTES = 32 << 30     # device workspace of one vp chunk (zpart dominates)


def _vp_batch(sh, rec):
    """Plain-path solves with the positive closure in chunks whose closure
    workspace stays within VP_BATCH_BYTES; returns (p float64, diag)."""
    import ctypes

    import torch

    from . import _lib
    from .solver import Solver
    L = _lib.load()
    per = int(L.sgl_vp_bytes(ctypes.byref(sh.c()), 1))
    chunk = max(1, min(rec.shape[0], VP_BATCH_BYTES // max(per, 1)))
    ps, ds = [], []
    sv = None
    for i in range(0, rec.shape[0], chunk):
        r = rec[i:i + chunk]
        if sv is None or sv.B != r.shape[0]:
            sv = Solver(sh, r.shape[0], path="exact", closure="vp")
        p, d = sv.run(r)
        torch.cuda.synchronize()
        st = int(sv.status.cpu()[0])
        if st & (_lib.STATUS_VP_TRANSITION | _lib.STATUS_VP_RATE):
            raise _lib.SglError(f"vp closure set-up failed (status {st:#x})")
        ps.append(p.cpu().numpy())
        ds.append(d.cpu().numpy())
    return np.concatenate(ps), np.concatenate(ds)
