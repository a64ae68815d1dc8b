"""Writes tests/golden/glynn_large.json: reference-grade Glynn permanents of seeded Haar
blocks at the full sizes of the single-large-permanent workload (n = 24, 28, 32), from the
oracle's 80-bit walk (mode 2: long double column sums and products, compensated sums).
Each matrix is Haar(2n, seed)[rows, :n] with rows a seeded sorted n-subset of range(2n);
only seeds and results are stored, the matrices are rebuilt by the test from the seeds.
re_mode1/im_mode1 keep the FP64 walk with double-double accumulation (mode 1) that the
file held before, whose own error at these sizes is what `re`/`im` now measure.

Cases already computed in mode 2 are kept (the values are deterministic); the others are
computed.

Run: python tests/golden/make_glynn_fixtures.py   (about 35 minutes on 8 cores from empty)
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
    out = {"generator": "oracle.perm_glynn_batch(mode=2)", "matrix": "Haar(2n, 1000n+b)[sorted rng(1000n+b).choice(2n, n), :n]",
           "cases": []}
    for n, count in spec.items():
        todo = [b for b in range(count) if have.get((n, b), {}).get("mode") != 2]
        for b in todo:  # one matrix at a time, each parallel over its chunks
            A = block(n, b)
            v = orc.perm_glynn_batch(A[None], 2, threads=0)[0]
            old = have.get((n, b))
            if old is not None and "re_mode1" not in old:
                m1 = complex(old["re"], old["im"])
            elif old is not None:
                m1 = complex(old["re_mode1"], old["im_mode1"])
            else:
                m1 = orc.perm_glynn_batch(A[None], 1)[0]
            have[(n, b)] = {"n": n, "b": b, "mode": 2, "re": float(v.real), "im": float(v.imag),
                            "re_mode1": float(m1.real), "im_mode1": float(m1.imag)}
            print(n, b, f"mode1 rel. deviation {abs(m1 - v) / abs(v):.2e}", flush=True)
            with open(path, "w") as f:  # keep what is done if interrupted
                json.dump({**out, "cases": [have[k] for k in sorted(have)]}, f, indent=1)
        out["cases"].extend(have[(n, b)] for b in range(count))
        print(n, "done", len(todo), "computed", flush=True)
    with open(path, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
