"""Autotuner for the chunk size of the Gray-walk kernels.

Every kernel here cuts a permanent's Gray enumeration into aligned chunks of 2^c ranks,
one chunk per lane. c sets the work per lane between O(n^2) re-seeds and per-chunk
reductions, and, with the batch, the number of waves per permanent: larger chunks
amortise the seeding, smaller ones give small batches enough waves and shorten the
tail of the launch. The best c depends on the size (k or n), the batch and the machine,
so it is measured: for each size the tuner times the kernel at candidate chunk sizes
around the built-in plan's choice, on seeded synthetic inputs of the workload's shape,
and stores the fastest in the library's tuning table (tb_tune_set). Every later call
that does not set a chunk size itself uses the table, including the batched sampler,
whose captured graphs are keyed by the resolved chunk sizes.

Entries are kept per (kernel, size, batch bucket): the chunk that serves one permanent
best is not the one that serves thousands. The table is process-wide; save()/load()
keep it across processes.
"""
import json
import math
import time
from typing import Dict, Iterable, List, Optional

import numpy as np

from . import _device as D
from . import _lib as L
from . import ops

KERNELS = {"sweep": L.TB_TUNE_SWEEP, "glynn": L.TB_TUNE_GLYNN, "grad": L.TB_TUNE_GRAD}


def batch_bucket(batch: int) -> int:
    """The library's batch bucket of a launch: ceil(log2(batch)), capped at the last."""
    b = 0
    while b < L.TB_TUNE_BATCH_BUCKETS - 1 and (1 << b) < batch:
        b += 1
    return b


def set_entry(kernel: str, size: int, chunk_bits: int, batch: int = 0) -> None:
    """Tuning entry for `kernel` at `size` for launches of `batch` items (their bucket);
    batch 0 sets every bucket of the size."""
    L.check(L.load().tb_tune_set(KERNELS[kernel], int(size), int(batch), int(chunk_bits)))


def get_entry(kernel: str, size: int, batch: int) -> int:
    rc = L.load().tb_tune_get(KERNELS[kernel], int(size), int(batch))
    L.check(min(rc, 0))
    return rc


def reset() -> None:
    """Back to the built-in plans for every kernel, size and batch."""
    for name in KERNELS:
        for size in range(L.TB_TUNE_MAX_SIZE + 1):
            set_entry(name, size, 0)


def table() -> Dict[str, Dict[str, int]]:
    """Non-zero entries of the tuning table: {kernel: {"size/bucket": chunk_bits}}, where
    bucket b holds batches in (2^(b-1), 2^b]."""
    out: Dict[str, Dict[str, int]] = {}
    for name in KERNELS:
        ent = {}
        for s in range(L.TB_TUNE_MAX_SIZE + 1):
            for b in range(L.TB_TUNE_BATCH_BUCKETS):
                c = get_entry(name, s, 1 << b)
                if c:
                    ent[f"{s}/{b}"] = c
        out[name] = ent
    return out


def save(path: str) -> None:
    with open(path, "w") as f:
        json.dump({"arch": "gfx950", "table": table()}, f, indent=1, sort_keys=True)


def load(path: str) -> None:
    with open(path) as f:
        data = json.load(f)
    for name, ent in data["table"].items():
        for key, c in ent.items():
            size, bucket = (int(v) for v in key.split("/"))
            set_entry(name, size, int(c), 1 << bucket)


# ------------------------------------------------------------------ built-in plans
# Mirrors of the library's built-in chunk choices, used only to centre the candidates.
def sweep_builtin(k: int, items: int) -> int:
    nb = k - 1
    want = 0
    while (max(items, 1) << want) < 4096 and want < 30:
        want += 1
    c = max(min(11, nb - 6 - want), min(5, nb - 6))
    return min(max(c, min(3, nb)), nb, 16)


def glynn_builtin(n: int, batch: int) -> int:
    """The FP64 kernels' built-in chunk (two waves per SIMD): ~8192 waves, then up to two
    sizes larger while the rounds of 2 x 1024 waves stay as full, scored by round fill
    times the share of a chunk left after seeding, (n - 2) 2 / (5 2^c)."""
    nb = n - 1
    want = 0
    while (batch << want) < 8192 and want < 30:
        want += 1

    def clamp(c):
        return max(2, min(c, 12, nb - 8))

    c0 = clamp(max(nb - 6 - want, min(7, nb - 8)))
    best, best_c = -1.0, c0
    for c in (c0, c0 + 1, c0 + 2):
        if clamp(c) != c:
            continue
        rounds = batch * 2.0 ** (nb - c - 6) / 2048
        score = rounds / math.ceil(rounds) * (1 - (n - 2) * 2.0 / (5 * 2.0 ** c))
        if score > best + 1e-9:
            best, best_c = score, c
    return best_c


def grad_columns(n: int) -> int:
    """Columns per sweep of the gradient kernel at n (8 for 14 <= n <= 26)."""
    return 8 if 14 <= n <= 26 else 4


def grad_static_bits(n: int) -> int:
    return 2 if grad_columns(n) == 8 else (4 if n <= 19 else 3)


def grad_builtin(n: int, batch: int) -> int:
    """The gradient's built-in chunk: ~8192 waves, then up to two sizes larger while the
    main launch's rounds of 2 x 1024 waves stay as full, scored by round fill times the
    share of a chunk left after seeding, 4n(n - 1) / ((8n + 18G - 30) 2^c)."""
    nb = n - 1
    g = grad_columns(n)
    groups = (n + g - 1) // g
    main = n // g if g == 8 and 0 < n % 8 <= 4 else groups  # the 4-column tail apart
    want = 0
    while (batch * groups << want) < 8192 and want < 30:
        want += 1
    low = grad_static_bits(n)
    if main != groups:  # the tail launch's 4-column groups need their static bits too
        low = max(low, 4 if n <= 19 else 3)
    c0 = min(max(min(nb - 6 - want, 11), low), nb - 7)
    best, best_c = -1.0, c0
    for c in range(c0, min(c0 + 2, 11, nb - 7) + 1):
        rounds = batch * main * 2.0 ** (nb - c - 6) / 2048
        score = rounds / math.ceil(rounds) * (1 - 4.0 * n * (n - 1) / ((8 * n + 18 * g - 30) * 2.0 ** c))
        if score > best + 1e-9:
            best, best_c = score, c
    return best_c


def _candidates(centre: int, lo: int, hi: int, spread: int) -> List[int]:
    return [c for c in range(centre - spread, centre + spread + 1) if lo <= c <= hi]


def _time(fn, reps: int) -> float:
    T = D.torch()
    fn()  # warm (and allocate the workspace for this plan)
    T.cuda.synchronize()
    e0, e1 = T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    T.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


# A candidate replaces the built-in choice only when it is faster by more than this
# fraction, so timing noise does not flip the table between runs.
MARGIN = 0.03


def _pick(kernel: str, size: int, batch: int, cands: Iterable[int], builtin: int, fn,
          reps: int) -> dict:
    times = {}
    for c in cands:
        set_entry(kernel, size, c, batch)
        times[c] = _time(fn, reps)
    best = min(times, key=times.get)
    if builtin in times and times[best] > (1.0 - MARGIN) * times[builtin]:
        best = builtin
    set_entry(kernel, size, best, batch)
    return {"chunk_bits": best, "seconds": {c: t for c, t in sorted(times.items())}}


def _haar_block(rows: int, cols: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((rows, cols)) + 1j * rng.standard_normal((rows, cols))) / np.sqrt(2)
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def tune_sweep(ks: Iterable[int], items: int = 4096, spread: int = 2, reps: int = 3,
               seed: int = 0, min_k: int = 8) -> Dict[int, dict]:
    """C&C minors sweep at k particles over `items` permanents (the entry of that batch's
    bucket). Returns {k: {"chunk_bits", "seconds", "items"}}."""
    out = {}
    for k in ks:
        if k < min_k or k > L.TB_SWEEP_MAX_K:
            continue  # sub-millisecond launches: the built-in plan stands
        V = _haar_block(k, k, seed + k)[: k - 1, :].T.copy()  # particle-major k x (k-1)
        Vd = D.to_device_c128(np.ascontiguousarray(np.broadcast_to(V, (items, k, k - 1))))
        c0 = sweep_builtin(k, items)
        cands = _candidates(c0, 3, min(k - 1, 16), spread)
        out[k] = _pick("sweep", k, items, cands, c0, lambda: ops.sweep_minors_batch(Vd, k), reps)
        out[k]["items"] = items
    return out


def tune_sampler(m: int, n: int, samples: int, **kw) -> Dict[int, dict]:
    """The sweeps of a batched GCC-B run with `samples` samples and up to n particles."""
    del m  # the sweep cost does not depend on the number of modes
    return tune_sweep(range(2, n + 1), items=samples, **kw)


def tune_glynn(ns: Iterable[int], batch: int, spread: int = 2, reps: int = 2,
               seed: int = 0) -> Dict[int, dict]:
    """FP64 Glynn at n over a batch of `batch` matrices."""
    out = {}
    for n in ns:
        if n < 16:
            continue
        A = D.to_device_c128(np.stack([_haar_block(2 * n, 2 * n, seed + 31 * n + b)[:n, :n]
                                       for b in range(batch)]))
        c0 = glynn_builtin(n, batch)
        cands = _candidates(c0, 2, min(12, n - 9), spread)
        out[n] = _pick("glynn", n, batch, cands, c0, lambda: ops.glynn_batch(A), reps)
    return out


def tune_grad(ns: Iterable[int], batch: int, spread: int = 2, reps: int = 3,
              seed: int = 0) -> Dict[int, dict]:
    """Single-sweep gradient (14 <= n <= 26) over a batch of `batch` matrices."""
    out = {}
    for n in ns:
        if not 14 <= n <= 26:
            continue
        A = D.to_device_c128(np.stack([_haar_block(2 * n, 2 * n, seed + 37 * n + b)[:n, :n]
                                       for b in range(batch)]))
        u = grad_static_bits(n)
        c0 = grad_builtin(n, batch)
        cands = _candidates(c0, u, min(11, n - 8), spread)
        out[n] = _pick("grad", n, batch, cands, c0, lambda: ops.perm_grad_batch(A), reps)
    return out


def timed(fn, *args, **kw):
    """(result, wall seconds) of one tuning call."""
    t0 = time.perf_counter()
    r = fn(*args, **kw)
    return r, time.perf_counter() - t0
