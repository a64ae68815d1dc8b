"""Writes tests/golden/glynn_large.json: oracle-grade Glynn permanents (double-double
accumulation, re-seeded column sums) of seeded Haar blocks at the full sizes of the
single-large-permanent workload (n = 24, 28, 32). Each matrix is Haar(2n, seed)[rows, :n]
with rows a seeded sorted n-subset of range(2n); only seeds and results are stored,
the matrices are rebuilt by the test from the seeds.

Cases already in the file are kept (the values are deterministic); only missing (n, b)
pairs are computed.

Run: python tests/golden/make_glynn_fixtures.py   (about 10 minutes on 8 cores from empty)
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
    spec = {24: 16, 28: 16, 32: 16}
    path = os.path.join(HERE, "glynn_large.json")
    have = {}
    if os.path.exists(path):
        with open(path) as f:
            have = {(c["n"], c["b"]): c for c in json.load(f)["cases"]}
    out = {"generator": "oracle.perm_glynn_batch(mode=1)", "matrix": "Haar(2n, 1000n+b)[sorted rng(1000n+b).choice(2n, n), :n]",
           "cases": []}
    for n, count in spec.items():
        todo = [b for b in range(count) if (n, b) not in have]
        if todo:
            vals = orc.perm_glynn_batch(np.stack([block(n, b) for b in todo]), 1, threads=0)
            for b, v in zip(todo, vals):
                have[(n, b)] = {"n": n, "b": b, "re": float(v.real), "im": float(v.imag)}
        out["cases"].extend(have[(n, b)] for b in range(count))
        print(n, "done", len(todo), "computed", flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
