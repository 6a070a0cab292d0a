"""CPU FP64 oracle of the magnification-PDF hot path.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The product (``sgl_analytic_amd``) never imports, links or runs
anything here, and fails loudly when its HIP library is missing.

Modules: cosmo (sgl.Cosmology restatement), nfw (projected / truncated NFW),
chain (kernel-shape chain + ctypes binding of chain.c), build (gcc recipe).
"""
