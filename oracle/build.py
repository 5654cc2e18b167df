"""Build recipe for the oracle's C restatement (test infrastructure only).

Compiles oracle/lm_oracle.c with gcc into oracle/_build/liblm_oracle.so.
The reference (ali-ramadhan/lagrangian-microbes) is pure Python whose native
arithmetic lives in third-party packages that are not vendored (parcels
2.0.0beta2, scipy's cKDTree), so there is no reference C source to compile
into oracle/_ref.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "lm_oracle.c")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liblm_oracle.so")

CFLAGS = ["-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-Wall"]


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = ["gcc", *CFLAGS, "-o", tmp, SRC, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
