"""Write tests/golden/kernel_shape_oracle.npz: the FP64 oracle's PDFs at the
full kernel shape (64 slices uniform in comoving distance x 256 masses x 512
impact parameters, 8192 log-k nodes in [1e-2, 1e6], 2048 ln mu nodes on
[-0.5, 3]; halos + the slice-covariance two-point term, the B-spline window;
Planck cosmology) at z_s = 0.5, 1, 3, 10 with the Gaussian closure and at
z_s = 0.3 with the positive (vp) closure, with their FP64 moments.

These are the reference vectors of the GPU tests of that shape
(tests/test_gpu_kernel_shape.py): the oracle takes minutes per PDF on the
host, the device milliseconds.  Run in the build container:
    python tests/golden/make_kernel_shape.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import chain  # noqa: E402

CASES = (("zs0p5", 0.5, "gauss"), ("zs1", 1.0, "gauss"), ("zs3", 3.0, "gauss"),
         ("zs10", 10.0, "gauss"), ("zs0p3_vp", 0.3, "vp"))


def main():
    out = {}
    for key, zs, closure in CASES:
        t0 = time.time()
        r = chain.pdf(zs, chain.Shape(), closure=closure)
        out[key + "_p"] = r.p
        m = r.info["moments"]
        out[key + "_mom"] = np.array([m[k] for k in sorted(m)])
        if closure == "vp":
            out[key + "_eta"] = np.array([r.info["eta"], r.info["vp_target"]])
        print(f"{key}: {time.time() - t0:.0f} s", flush=True)
    out["moment_keys"] = np.array(sorted(m))
    out["x"] = chain.Shape().out_grid()
    np.savez_compressed(os.path.join(HERE, "kernel_shape_oracle.npz"), **out)
    print("wrote kernel_shape_oracle.npz")


if __name__ == "__main__":
    main()
