"""Writes tests/golden/glynn_large.json: oracle-grade Glynn permanents (double-double
accumulation, re-seeded column sums) of seeded Haar blocks at the full sizes of the
single-large-permanent workload (n = 24, 28, 32). Each matrix is Haar(2n, seed)[rows, :n]
with rows a seeded sorted n-subset of range(2n); only seeds and results are stored,
the matrices are rebuilt by the test from the seeds.

Run: python tests/golden/make_glynn_fixtures.py   (about 5 minutes on 8 cores)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as orc  # noqa: E402


def block(n: int, b: int) -> np.ndarray:
    seed = 1000 * n + b
    U = orc.haar_unitary(2 * n, seed)
    rows = np.sort(np.random.default_rng(seed).choice(2 * n, n, replace=False))
    return U[rows, :n]


def main():
    spec = {24: 16, 28: 16, 32: 8}
    out = {"generator": "oracle.perm_glynn_batch(mode=1)", "matrix": "Haar(2n, 1000n+b)[sorted rng(1000n+b).choice(2n, n), :n]",
           "cases": []}
    for n, count in spec.items():
        A = np.stack([block(n, b) for b in range(count)])
        vals = orc.perm_glynn_batch(A, 1, threads=0)
        for b, v in enumerate(vals):
            out["cases"].append({"n": n, "b": b, "re": float(v.real), "im": float(v.imag)})
        print(n, "done", flush=True)
    with open(os.path.join(HERE, "glynn_large.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
