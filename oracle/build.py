"""Build recipe for the CPU oracle (test infrastructure only).

Compiles oracle/permanent_oracle.c with gcc into oracle/liboracle.so. The reference
(theboss) is pure Python, so there is no reference source to compile into
oracle/_ref: the oracle is a restatement, pinned by the checks listed in
oracle/oracle.py.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "permanent_oracle.c")
LIB = os.path.join(HERE, "liboracle.so")


def build(verbose: bool = False) -> str:
    cmd = [
        "gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared",
        "-fno-fast-math", "-ffp-contract=off",
        # numpy multiplies complex128 with the plain (a*c - b*d, a*d + b*c) formula; so does
        # the oracle (no C99 Annex G inf/nan recovery calls).
        "-fcx-limited-range",
        SRC, "-o", LIB + ".tmp", "-lm",
    ]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
