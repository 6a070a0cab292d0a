"""Launch-shape autotuner for the batched solver.

Four knobs of the library choose launch shapes (``sgl_set_tuning``):

- ``bin_threads``: threads per workgroup of the binning kernel (one
  workgroup per (slice, instance); more threads give more cells in flight
  but share the same 75 KB of LDS histogram);
- ``lnphi_block``: k nodes per workgroup of the binned exponent kernel
  (the block also stages that many bin records through LDS per pass);
- ``gemm_bm``: rows of the Filon matrix per inversion GEMM tile (64 or 128;
  the larger tile halves the re-reads of phi at the price of occupancy);
- ``bin_schedule``: 0 = one binning workgroup per (instance, slice), the
  hardware dispatcher refilling the CUs; 1 = persistent workgroups (as many
  as the device holds) taking (instance, slice) items from a work queue.

The knobs are tuned one at a time, in this order (the schedule after the
binning block size it runs with): every
candidate runs its stage on the solver's own buffers and shape (records
already set), timed with HIP events on the current stream, best of ``reps``
after one warm-up.  Every candidate yields the same histogram and spectrum
bit for bit (the histogram sums are fixed point; the other kernels keep
their reduction order), so tuning never changes results.  Results are cached per (device name, shape, B, precision) in
memory; ``apply`` sets the knobs for a later solve (a captured graph keeps the
launch shapes it was captured with, so tune before ``Solver.capture``).
"""
from __future__ import annotations

import ctypes
import dataclasses

import torch

from . import _lib
from .solver import Solver, _ptr, _stream

KNOBS = {
    "bin_threads": (_lib.TUNE_BIN_THREADS, (256, 512, 1024)),
    "lnphi_block": (_lib.TUNE_LNPHI_BLOCK, (128, 256)),
    "gemm_bm": (_lib.TUNE_GEMM_BM, (64, 128)),
    "bin_schedule": (_lib.TUNE_BIN_SCHEDULE, (0, 1)),
}

_CACHE: dict = {}


def current() -> dict:
    """Knob values the library holds now (as launch sizes)."""
    L = _lib.load()
    return {name: vals[L.sgl_get_tuning(idx)] for name, (idx, vals) in KNOBS.items()}


def apply(choice: dict) -> None:
    L = _lib.load()
    for name, size in choice.items():
        idx, vals = KNOBS[name]
        _lib.check(L.sgl_set_tuning(idx, vals.index(size)), f"sgl_set_tuning({name})")


def _key(s: Solver):
    return (torch.cuda.get_device_name(s.dev), dataclasses.astuple(s.shape), s.B, s.precision)


def _stage(s: Solver, name: str):
    """The launch that a knob affects, on the solver's buffers."""
    L, B, cs, st = s.L, s.B, ctypes.byref(s.cs), _stream()
    if name in ("bin_threads", "bin_schedule"):
        return lambda: _lib.check(L.sgl_bin_cells(_ptr(s.tables), B, cs, _ptr(s.hist), st),
                                  "sgl_bin_cells")
    if name == "lnphi_block":
        return lambda: _lib.check(L.sgl_lnphi_binned(_ptr(s.tables), _ptr(s.hist), B, cs,
                                                     _ptr(s.lam), _ptr(s.amp), st),
                                  "sgl_lnphi_binned")
    return lambda: _lib.check(L.sgl_invert_gemm(_ptr(s.E), _ptr(s.phi), B, cs, s.precision,
                                                _ptr(s.p), st), "sgl_invert_gemm")


def _time(fn, reps: int) -> float:
    fn()
    best = float("inf")
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        best = min(best, a.elapsed_time(b))
    return best


def tune(s: Solver, reps: int = 3, knobs=None) -> dict:
    """Pick the fastest launch shape of each knob for solver ``s``.

    ``s`` must be a batched solver with its records set.  Returns
    ``{"choice": {knob: size}, "ms": {knob: {size: ms}}}`` and leaves the
    choice applied.  Cached per device, shape, batch and precision."""
    if s.path != "binned":
        raise ValueError("the tuner times the batched path")
    key = _key(s)
    if key in _CACHE:
        apply(_CACHE[key]["choice"])
        return _CACHE[key]
    s.run()                          # tables, histogram, spectrum, E, phi all valid
    torch.cuda.synchronize()
    choice, ms = current(), {}
    for name in (knobs or KNOBS):
        fn = _stage(s, name)
        ms[name] = {}
        for size in KNOBS[name][1]:
            apply({name: size})
            ms[name][size] = _time(fn, reps)
        choice[name] = min(ms[name], key=ms[name].get)
        apply({name: choice[name]})
    s.run()                          # leave the solver's outputs from the chosen shapes
    torch.cuda.synchronize()
    _CACHE[key] = {"choice": dict(choice), "ms": ms}
    return _CACHE[key]
