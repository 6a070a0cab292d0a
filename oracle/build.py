"""Build recipe of the oracle's C restatement (checker only).

TEST INFRASTRUCTURE.  ``python -m oracle.build`` compiles oracle/chain.c into
oracle/_ref/liboracle_chain.so with gcc (-O2 -fopenmp, no fast-math: the
compensated forms must not be re-associated).  The reference itself is pure
Python + Numba, so there is no reference C source to compile here; _ref/
holds only this restatement's library.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
LIB = os.path.join(OUT_DIR, "liboracle_chain.so")


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    src = os.path.join(HERE, "chain.c")
    tmp = LIB + ".tmp"
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-D_GNU_SOURCE",
           "-o", tmp, src, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
