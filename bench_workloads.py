"""bench.py's measured runs of the secondary workloads (single large permanents,
differentiation, GCC-A, uniform-loss GCC-B, interferometer fitting, outcome histogram,
BOBS and exact distributions). Each returns one JSON-able record. Kept beside bench.py,
outside the theboss_amd package, because it drives the oracle.

`checker` is the CPU oracle module passed in by the caller (bench.py); it produces
seeded inputs and reference values, and its restatements of the reference's loops are
timed as the CPU baselines reported beside the GPU figures. It never computes a GPU
figure.
"""
import collections
import json
import math
import os
import time

import numpy as np

from theboss_amd import _device as D
from theboss_amd import _lib as L
from theboss_amd import ops

FP64_PEAK = 78.6e12
FP32_PEAK = 157.3e12
ROOT = os.path.dirname(os.path.abspath(__file__))


def _events():
    T = D.torch()
    return T.cuda.Event(enable_timing=True), T.cuda.Event(enable_timing=True)


def _time(fn, reps=1, warm_s=0.0):
    """Mean seconds per call over reps calls, after one warm-up call, or after warm-up
    calls filling warm_s seconds (a GPU idle before raises its clock over the first tens
    of milliseconds of work, which short launches would otherwise be timed in)."""
    T = D.torch()
    t_end = time.perf_counter() + warm_s
    fn()
    T.cuda.synchronize()
    while time.perf_counter() < t_end:
        fn()
        T.cuda.synchronize()
    e0, e1 = _events()
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    T.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3 / reps


def clock_during(fn, seconds):
    """Run fn (which queues kernels on the current stream) with a one-wave clock probe on a
    second stream that spans `seconds` of it; returns (fn's result, average shader clock in
    GHz over that span, from the probe's shader-cycle and real-time counter deltas)."""
    T = D.torch()
    lib = L.load()
    hz = int(lib.tb_realtime_hz())
    L.check(min(hz, 0))
    buf = T.zeros(2, dtype=T.int64, device="cuda")
    side = T.cuda.Stream()  # default priority
    T.cuda.synchronize()
    res = fn()
    with T.cuda.stream(side):
        L.check(lib.tb_clock_probe(D.ptr(buf), max(1, int(seconds * hz)), D.stream_ptr()))
    T.cuda.synchronize()
    cycles, ticks = (int(v) for v in buf.cpu().tolist())
    return res, cycles / ticks * hz / 1e9


class GuanFlops:
    """Counts, inside the with-block, the multiplicity-aware algorithmic flops of every
    tb_guan_perm_batch and tb_guan_perm_outer launch the calculators make: per item (pair)
    prod_d (e_d + 1) Guan codes, each costing 2|T| + 6(sum t - 1) + 2 (|T| complex adds
    into the sums of the product side's distinct modes, sum t - 1 complex multiplies, the
    accumulation), with e the enumerated side's multiplicities and t the product side's.
    The count is summed on the device and read once at the end."""

    def __enter__(self):
        T = D.torch()
        self._orig = ops.guan_perm_batch
        self._acc = T.zeros((), dtype=T.float64, device="cuda")
        self.launches = 0
        self.items = 0

        def counted(U, out_modes, out_mult, in_modes, in_mult, formula=L.TB_FORMULA_RYSER):
            codes = (in_mult.to(T.float64) + 1).prod(dim=1)
            distinct = (out_mult > 0).sum(dim=1).to(T.float64)
            parts = out_mult.sum(dim=1).to(T.float64)
            self._acc += (codes * (2 * distinct + 6 * (parts - 1) + 2)).sum()
            self.launches += 1
            self.items += int(in_mult.shape[0])
            return self._orig(U, out_modes, out_mult, in_modes, in_mult, formula)

        self._orig_outer = ops.guan_perm_outer

        def counted_outer(U, out_modes, out_mult, in_modes, in_mult, formula=L.TB_FORMULA_RYSER):
            # pair (j, i) enumerates the T row j when it has fewer codes than the S row i
            f64 = T.float64
            codes_t, codes_s = (out_mult.to(f64) + 1).prod(dim=1), (in_mult.to(f64) + 1).prod(dim=1)
            dist_t, dist_s = (out_mult > 0).sum(dim=1).to(f64), (in_mult > 0).sum(dim=1).to(f64)
            parts_t, parts_s = out_mult.sum(dim=1).to(f64), in_mult.sum(dim=1).to(f64)
            enum_s = codes_s[None, :] * (2 * dist_t[:, None] + 6 * (parts_t[:, None] - 1) + 2)
            enum_t = codes_t[:, None] * (2 * dist_s[None, :] + 6 * (parts_s[None, :] - 1) + 2)
            self._acc += T.where(codes_t[:, None] < codes_s[None, :], enum_t, enum_s).sum()
            self.launches += 1
            self.items += int(out_mult.shape[0]) * int(in_mult.shape[0])
            return self._orig_outer(U, out_modes, out_mult, in_modes, in_mult, formula)

        ops.guan_perm_batch = counted
        ops.guan_perm_outer = counted_outer
        return self

    def __exit__(self, *exc):
        ops.guan_perm_batch = self._orig
        ops.guan_perm_outer = self._orig_outer
        self.flops = float(self._acc.item())
        return False


def sweep_small(checker):
    """The C&C minors sweep as the drop-in submatrices calculator and the plain GCC-B path
    launch it: one item, and small batches, at k = 20 (the batch-aware built-in plan cuts
    the chunks until the launch has ~4096 waves, not below 2^5 Gray steps per lane)."""
    T = D.torch()
    lib = L.load()
    k = 20
    V = checker.haar_unitary(2 * k, 2020)[: k - 1, :k].T.copy()  # particle-major k x (k-1)
    ref = ops.sweep_minors_batch(D.to_device_c128(V[None]), k).cpu().numpy()[0]
    rows = []
    for B in (1, 16, 256, 4096):
        Vd = D.to_device_c128(np.ascontiguousarray(np.broadcast_to(V, (B, k, k - 1))))
        reps = 200 if B <= 16 else 20
        out = {}
        t = _time(lambda: out.setdefault("m", ops.sweep_minors_batch(Vd, k)), reps=reps, warm_s=0.2)
        wpi = lib.tb_sweep_workspace_size(k, B) // (B * k * 16)
        fl = B * 2 ** (k - 1) * (10 * k - 12)
        got = out["m"].cpu().numpy()
        rows.append({"batch": B, "waves_per_item": int(wpi), "waves": int(B * wpi),
                     "ms": t * 1e3, "tflops": fl / t / 1e12, "frac_fp64_peak": fl / t / FP64_PEAK,
                     "max_rel_diff_vs_one_item": float(np.max(np.abs(got - ref[None])) / np.max(np.abs(ref)))})
    return {"workload": "C&C minors sweep, k=20, batches 1..4096 (built-in plans)", "results": rows}


def grad_executed_flops(n):
    """Flops the single-sweep gradient executes per n x n matrix: per Gray step of a group of
    G columns a row flip (2n), n + 2G - 5 complex multiplies (6 each), G running sums (2
    each) and G corrector FMAs (4 each), 8n + 18G - 30 flops, over 2^(n-1) steps; groups of
    8 columns, the last 1..4 columns (n mod 8) in one group of 4, else ceil(n/8) groups of 8
    (the last overlapping its predecessor)."""
    full, rest = divmod(n, 8)
    groups = [8] * full + ([4] if 0 < rest <= 4 else [8] if rest else [])
    return 2 ** (n - 1) * sum(8 * n + 18 * g - 30 for g in groups)


def glynn_flops(n):
    """Executed flops of one n x n Glynn permanent: 2^(n-1) Gray steps, each n complex
    column adds (2n), n - 1 complex multiplies (6 each) and the signed accumulation (2),
    8n - 4 per step."""
    return (8 * n - 4) * 2 ** (n - 1)


def glynn_instructions_per_code(n):
    """FP64 instructions per Gray code of the kernel that runs n (FP64, generic matrices).
    Paired walk (14 <= n <= 39): per pair of codes one row flip (2n) and two products
    of two folded chains each, joined in the accumulation (2 (4n - 2)): 5n - 2 per code.
    Plain walk: 2n column adds and n - 1 complex multiplies of 4 instructions, the last
    one fused into the accumulation: 6n - 4."""
    return 5 * n - 2 if 14 <= n <= 39 else 6 * n - 4


def glynn_issue_ceiling(n, clock_ghz):
    """Fraction of the FP64 peak (2.4 GHz, every instruction an FMA) that the Glynn kernel
    can reach when its instructions issue back to back: the 8n - 4 flops of a Gray code
    against its instructions, 2 peak flops each, at the measured clock."""
    return (8 * n - 4) / (2 * glynn_instructions_per_code(n)) * clock_ghz / 2.4


def _glynn_blocks(checker, n, B=16):
    mats = []
    for b in range(B):
        seed = 1000 * n + b
        U = checker.haar_unitary(2 * n, seed)
        r = np.sort(np.random.default_rng(seed).choice(2 * n, n, replace=False))
        mats.append(U[r, :n])
    return np.stack(mats)


def _cpu_cores():
    return max(1, min(16, os.cpu_count() or 1))


def glynn_cpu_rates(checker, n0=24, npy=13):
    """Per (Gray step x column) costs of the reference's Glynn loop on the host: the
    oracle's C walk in the reference's arithmetic (mode 0), OpenMP over 16 blocks of
    n0 x n0 (one block per thread), and the pure-Python walk on one n_py x n_py block.
    Returns (C thread-seconds, Python seconds) per step-column and the sample notes."""
    cores = _cpu_cores()
    A = _glynn_blocks(checker, n0)
    t0 = time.perf_counter()
    checker.perm_glynn_batch(A, mode=0, threads=cores)
    dt = time.perf_counter() - t0
    c_rate = dt * min(cores, len(A)) / (len(A) * 2 ** (n0 - 1) * n0)
    Ap = _glynn_blocks(checker, npy, 1)[0]
    t0 = time.perf_counter()
    checker.perm_glynn_py(Ap)
    py_rate = (time.perf_counter() - t0) / (2 ** (npy - 1) * npy)
    note_c = (f"oracle C Glynn walk in the reference's arithmetic on 16 blocks n={n0} "
              f"({dt:.2f} s on {min(cores, len(A))} threads), per Gray step and column")
    note_py = (f"pure-Python restatement of the reference's Glynn loop on one n={npy} block "
               f"({py_rate * 1e9:.0f} ns per Gray step and column, 1 core)")
    return c_rate, py_rate, cores, note_c, note_py


def glynn(checker, ns=(24, 28, 32), precisions=("fp64", "fp32_colsum"), cpu=True):
    """16 collision-free n x n blocks of Haar(2n) per n in {24, 28, 32}; FP64 against the
    golden oracle values, FP32 column sums against FP64. With cpu, the reference's loop
    timed on the host (C restatement and pure Python) is scaled to each n: the walk costs
    a fixed time per Gray step and column."""
    with open(os.path.join(ROOT, "tests", "golden", "glynn_large.json")) as f:
        gold = json.load(f)["cases"]
    rows = []
    for n in ns:
        B = 16
        A = D.to_device_c128(_glynn_blocks(checker, n, B))
        fl = B * glynn_flops(n)
        rec = {"n": n, "batch": B, "flops": fl}
        for prec, name in ((L.TB_PREC_FP64, "fp64"), (L.TB_PREC_FP32_COLSUM_ANY_N, "fp32_colsum")):
            if name not in precisions:
                continue
            out = {}
            reps = 1 if n >= 30 else 20  # short launches: average many back to back
            t = _time(lambda: out.setdefault("v", ops.glynn_batch(A, prec)), reps=reps,
                      warm_s=0.2)
            vals = out["v"].cpu().numpy()
            rec[name] = {"ms": t * 1e3, "tflops": fl / t / 1e12,
                         "frac_fp64_peak": fl / t / FP64_PEAK}
            if prec == L.TB_PREC_FP32_COLSUM_ANY_N:
                rec[name]["frac_fp32_peak"] = fl / t / FP32_PEAK
            ref = {c["b"]: complex(c["re"], c["im"]) for c in gold if c["n"] == n}
            errs = [abs(vals[b] - ref[b]) / abs(ref[b]) for b in ref]
            rec[name]["max_rel_err_vs_oracle"] = max(errs)
            rec[name]["checked_against"] = f"{len(errs)} golden oracle values"
            if prec == L.TB_PREC_FP64:
                # the same launch again with the clock probe spanning 80% of it
                _, ghz = clock_during(lambda: ops.glynn_batch(A, prec), 0.8 * t)
                rec[name]["clock_ghz"] = ghz
                rec[name]["issue_bound_ceiling_frac"] = glynn_issue_ceiling(n, ghz)
        rows.append(rec)
    # The CPU lines come after the GPU timings: a GPU left idle during seconds of host work
    # starts the next short launches at a low clock.
    if cpu:
        c_rate, py_rate, cores, note_c, note_py = glynn_cpu_rates(checker)
        for rec in rows:
            n, B = rec["n"], rec["batch"]
            work = B * 2 ** (n - 1) * n  # Gray steps x columns of the batch
            wall_c = work * c_rate / min(cores, B)
            wall_py = work * py_rate
            rec["cpu_baseline"] = {"value": B / wall_c, "unit": "permanents/s", "cores": min(cores, B),
                                   "kind": "port", "est_seconds": wall_c, "sample": note_c}
            rec["cpu_python_estimate"] = {"value": B / wall_py, "unit": "permanents/s", "cores": 1,
                                          "kind": "port", "est_seconds": wall_py, "sample": note_py}
            for name in ("fp64", "fp32_colsum"):
                if name in rec:
                    rec[name]["gpu_vs_cpu_baseline"] = wall_c / (rec[name]["ms"] * 1e-3)
    return {"workload": "single large permanents (Glynn, 16 per n)", "results": rows}


def glynn_fp32(checker, sizes=((16, 4096), (20, 4096), (23, 512))):
    """The FP32 column-sum variant where it is meant to be used (n <= 23, batches whose
    launch lasts milliseconds), against the FP64 kernel on the same blocks: time, fraction
    of the FP32 and FP64 vector peaks ((8n - 4) 2^(n-1) flops per permanent), shader clock and
    the max relative deviation from the FP64 result."""
    rows = []
    for n, B in sizes:
        base = np.stack([checker.haar_unitary(2 * n, 7000 + n + b)[:n, :n] for b in range(16)])
        A = D.to_device_c128(np.ascontiguousarray(np.resize(base, (B, n, n))))
        fl = B * glynn_flops(n)
        rec = {"n": n, "batch": B, "flops": fl}
        for prec, name in ((L.TB_PREC_FP32_COLSUM, "fp32_colsum"), (L.TB_PREC_FP64, "fp64")):
            out = {}
            t = _time(lambda: out.setdefault("v", ops.glynn_batch(A, prec)), reps=10, warm_s=0.2)
            _, ghz = clock_during(lambda: [ops.glynn_batch(A, prec) for _ in range(5)], 0.8 * 5 * t)
            rec[name] = {"ms": t * 1e3, "tflops": fl / t / 1e12, "frac_fp32_peak": fl / t / FP32_PEAK,
                         "frac_fp64_peak": fl / t / FP64_PEAK, "clock_ghz": ghz}
            rec[name + "_values"] = out["v"].cpu().numpy()
        v32, v64 = rec.pop("fp32_colsum_values"), rec.pop("fp64_values")
        rec["fp32_max_rel_diff_vs_fp64"] = float(np.max(np.abs(v32 - v64) / np.abs(v64)))
        rows.append(rec)
    return {"workload": "Glynn FP32 column sums at n = 16, 20, 23 (FP64 alongside)", "results": rows}


def grad_cpu(checker, mats):
    """The reference's route to a permanent and all its n^2 minors, timed on the host over
    one block per thread: the forward by the oracle's C Glynn walk in the reference's
    arithmetic, the minors by n C&C Ryser submatrices walks (the reference's calculator,
    one per deleted row: perm(A without row i, column j) for all j)."""
    import concurrent.futures as cf

    cores = _cpu_cores()
    sample = mats[:cores]
    n = sample[0].shape[0]

    def one(A):
        checker.perm_glynn_matrix(A, 0)
        ones = np.ones(n, dtype=np.int64)
        for i in range(n):
            t = ones.copy()
            t[i] = 0
            checker.cc_submatrices(A, ones, t)

    checker.lib()
    t0 = time.perf_counter()
    with cf.ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(one, sample))
    dt = time.perf_counter() - t0
    return {"value": len(sample) / dt, "unit": "matrices/s (permanent and all n^2 minors)",
            "cores": cores, "kind": "port",
            "sample": f"{len(sample)} of the blocks, one per thread ({dt:.2f} s): the oracle's C "
                      f"Glynn walk for the permanent and {n} C&C Ryser submatrices walks for the "
                      f"minors, as the reference computes them"}


def grad(checker, cpu=True, tune=True):
    """64 n=20 blocks of Haar(40): forward permanents + all n^2 minors (backward). With
    tune, the gradient kernel's chunk size is first tuned for this shape (as the sampler's
    are for theirs); with cpu, the reference's route to the same numbers is timed on the
    host beside it."""
    n, B = 20, 64
    mats = []
    for b in range(B):
        r = np.sort(np.random.default_rng(300 + b).choice(40, n, replace=False))
        mats.append(checker.haar_unitary(40, 300 + b)[r, :n])
    tuned = None
    if tune:
        from theboss_amd import autotune

        tuned = autotune.tune_grad([n], B, spread=4, reps=3)[n]
    from theboss_amd.autograd import permanent

    T = D.torch()
    A = D.to_device_c128(np.stack(mats))
    # 20 repetitions: the host work before the first launch after the start event is then
    # a small share of the measured time (the passes themselves queue back to back)
    t_fwd = _time(lambda: ops.glynn_batch(A), reps=20)  # forward alone (no grad)
    t_bwd = _time(lambda: ops.perm_grad_batch(A), reps=20)
    Ag = A.clone().requires_grad_(True)
    seed = T.ones(B, dtype=T.complex128, device=A.device)

    def fwd_bwd():
        Ag.grad = None
        permanent(Ag).backward(seed)

    t_pass = _time(fwd_bwd, reps=20)
    # average shader clock over 20 back-to-back gradient launches
    _, ghz = clock_during(lambda: [ops.perm_grad_batch(A) for _ in range(20)], 0.8 * 20 * t_bwd)
    cpu_rec = grad_cpu(checker, mats) if cpu else None  # after the GPU timings
    f_fwd = B * glynn_flops(n)
    # All n^2 minors by one sweep: 2^(n-1) (2n^2 + 20n) (n column adds, ~3n multiplies for
    # leave-one-out products, n^2 signed accumulations per step).
    f_grad = B * 2 ** (n - 1) * (2 * n * n + 20 * n)
    f_bwd_exec = B * grad_executed_flops(n)
    return {"workload": "permanent + gradient, 64 x n=20",
            "autograd_fwd_bwd_ms": t_pass * 1e3,
            "cpu_baseline": cpu_rec,
            "autotune": tuned,
            "gpu_vs_cpu_baseline": (B / t_pass) / cpu_rec["value"] if cpu_rec else None,
            "frac_fp64_peak_executed": f_bwd_exec / t_pass / FP64_PEAK,
            "frac_fp64_peak_executed_kernels": f_bwd_exec / t_bwd / FP64_PEAK,
            "frac_fp64_peak_effective": (f_fwd + f_grad) / t_pass / FP64_PEAK,
            "clock_ghz": ghz,
            "perm_only_ms": t_fwd * 1e3, "perm_and_grad_kernel_ms": t_bwd * 1e3,
            "flops_executed": f_bwd_exec, "flops_effective": f_fwd + f_grad,
            "note": "autograd pass: the forward of permanent() on a tensor that requires grad "
                    "runs the gradient launch, which returns perm and all n^2 minors (one Gray "
                    "sweep per group of 8 columns); the backward scales the kept minors. "
                    "executed: the flops the sweeps perform, per matrix 2^(n-1) (8n + 114) for "
                    "each group of 8 columns and 2^(n-1) (8n + 42) for the group of 4 that takes "
                    "the last n mod 8 <= 4 columns (n = 20: 2 of 8 and 1 of 4) "
                    "(the FLOP counter reads ~1.2x this: FMAs with a +-1 or +-2 operand "
                    "count 2); _kernels: over the gradient launches alone, without the autograd "
                    "host work. effective: algorithm-equivalent flops, a Glynn forward (8n - 4) 2^(n-1) "
                    "plus the all-minors count 2^(n-1)(2n^2 + 20n) per matrix, which charges "
                    "more than the single sweep executes (not a roofline fraction)"}


def exact(checker, m=12, n=6, eta=0.5):
    """Exact uniform-loss output distribution, m = 12, n = 6 photons in modes 0..5,
    eta = 0.5: all C(m + n, n) = 18,564 outcomes with at most n photons, every
    (outcome, lossy input) permanent by one batched Ryser launch per particle number.
    CPU line: the oracle's restatement of the reference's per-outcome loop
    (bs_exact_distribution_with_uniform_losses.py), C permanents, one core."""
    from theboss_amd.distribution_calculators import (BosonSamplingExperimentConfiguration,
                                           BSDistributionCalculatorWithUniformLosses)
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator

    U = checker.haar_unitary(m, 1200 + m)
    inp = [1] * n + [0] * (m - n)
    cfg = BosonSamplingExperimentConfiguration(U, inp, n, m, 0, n, eta)

    def run():
        calc = BSDistributionCalculatorWithUniformLosses(cfg, GPURyserPermanentCalculator(U))
        return np.asarray(calc.calculate_distribution())

    T = D.torch()
    # warm-up run (library, workspaces), which also counts the launches' flops: the
    # distribution is deterministic, so the timed run makes the same launches, uncounted
    with GuanFlops() as gf:
        run()
    T.cuda.synchronize()
    t0 = time.perf_counter()
    dist = run()
    T.cuda.synchronize()
    dt = time.perf_counter() - t0
    perms = sum(math.comb(l + m - 1, l) * math.comb(n, l) for l in range(1, n + 1))
    t0 = time.perf_counter()
    ref = np.asarray(checker.uniform_losses_distribution(U, inp, eta, checker.perm_ryser))
    dt_cpu = time.perf_counter() - t0
    return {"workload": f"exact uniform-loss distribution, m={m}, n={n}, eta={eta}",
            "outcomes": len(dist), "permanents": perms, "seconds": dt,
            "guan_launches": gf.launches, "guan_items": gf.items, "guan_flops": gf.flops,
            "permanents_per_s": perms / dt, "sum": float(dist.sum()),
            "max_abs_diff_vs_oracle": float(np.max(np.abs(dist - ref))),
            "cpu_baseline": {"value": perms / dt_cpu, "unit": "permanents/s", "cores": 1,
                             "kind": "port", "seconds": dt_cpu,
                             "sample": "the whole distribution by the oracle's per-outcome loop "
                                       "(C Ryser permanents, Python loop), 1 core"},
            "gpu_vs_cpu_baseline": dt_cpu / dt,
            "note": "end to end from Python: outcome enumeration, host index lists and the "
                    "per-particle-number launches"}


def gcca(checker):
    """GCC-A, m=20, n=10 photons in modes 0..9, 200 samples, Ryser on the GPU (the
    samples advance in lock-step: one all-pairs launch per particle number); plus one
    10x10 permanent by Ryser and Glynn."""
    from theboss_amd.permanent_calculators import GPUGlynnGrayPermanentCalculator, GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import GeneralizedCliffordsSimulationStrategy

    m, n, S = 20, 10, 200
    U = checker.haar_unitary(m, 1)
    inp = [1] * n + [0] * (m - n)
    # warm-up run of the full size with other draws: kernels loaded (a first use of a
    # kernel loads its code object, tens of ms) and device memory pools grown (the first
    # 200-sample run in a process takes ~4x longer than later ones)
    GeneralizedCliffordsSimulationStrategy(GPURyserPermanentCalculator(U),
                                           rng=np.random.default_rng(1)).simulate(inp, S)
    strat = GeneralizedCliffordsSimulationStrategy(GPURyserPermanentCalculator(U),
                                                   rng=np.random.default_rng(0))
    D.torch().cuda.synchronize()
    t0 = time.perf_counter()
    samples = strat.simulate(inp, S)
    D.torch().cuda.synchronize()
    dt = time.perf_counter() - t0
    # the flop count, outside the timed run: the same draws again (same seed, so the same
    # layers and launches) with the counter on every launch
    with GuanFlops() as gf:
        again = GeneralizedCliffordsSimulationStrategy(GPURyserPermanentCalculator(U),
                                                       rng=np.random.default_rng(0)).simulate(inp, S)
    assert again == samples, "the counted run drew other samples than the timed run"
    ones = [1] * n + [0] * (m - n)
    r = GPURyserPermanentCalculator(U, ones, ones).compute_permanent()
    g = GPUGlynnGrayPermanentCalculator(U, ones, ones).compute_permanent()
    # CPU line of the reference's speed class: the layer pmf with the checker's pure-Python
    # Ryser walk (the reference's loop structure), timed on a few chains of one layer per
    # particle number and scaled by the number of layers of each particle number this run
    # computed (the memo of the reference computes the same layers).
    counts = collections.Counter(sum(key) + 1 for key in strat.pmfs)
    per_k = collections.defaultdict(list)
    rng = np.random.default_rng(1)
    spent, chains = 0.0, 0
    while spent < 6.0 and chains < 4:
        prefix = [0] * m
        for k in range(1, n + 1):
            t1 = time.perf_counter()
            pmf = checker.gcca_layer_pmf(U, inp, prefix, perm=checker.perm_ryser_py)
            per_k[k].append(time.perf_counter() - t1)
            prefix[int(rng.choice(m, p=pmf / pmf.sum()))] += 1
        spent += sum(per_k[k][-1] for k in range(1, n + 1))
        chains += 1
    est = sum(float(np.mean(per_k[k])) * c for k, c in counts.items())
    cpu = {"value": S / est, "unit": "samples/s", "cores": 1, "kind": "port",
           "sample": f"{chains} chains of the {n} layer pmfs with a pure-Python Ryser walk, "
                     f"{spent:.1f} s, scaled by this run's {len(strat.pmfs)} layers"}
    return {"workload": "GCC-A m=20 n=10, 200 samples (Ryser, lock-step layers)",
            "samples_per_s": S / dt, "seconds": dt, "layers_computed": len(strat.pmfs),
            "guan_launches": gf.launches, "guan_items": gf.items, "guan_flops": gf.flops,
            "cpu_python_baseline": cpu, "gpu_vs_cpu_python": (S / dt) / cpu["value"],
            "ryser_vs_glynn_10x10_rel": abs(r - g) / abs(g),
            "all_samples_have_n_photons": all(sum(s_) == n for s_ in samples)}


def lossy_cpu(checker, U, n, eta, l0=18):
    """CPU line of the uniform-loss sampler. A sample's cost is dominated by the binomial
    tail of its photon number l (the k-th step sweeps 2^k terms), so a bounded sample
    cannot show the mean: the oracle's C GCC-B sampler is timed at l = l0 (one sample per
    thread) and scaled to every l by the C&C work sum_{k <= l} 2^k (10k - 12), then
    averaged over Binomial(n, eta)."""
    cores = _cpu_cores()
    order = np.tile(np.arange(n, dtype=np.int32), (cores, 1))
    unif = np.random.default_rng(1).random((cores, n))
    steps = np.full(cores, l0, dtype=np.int32)
    t0 = time.perf_counter()
    checker.gccb_sample_batch(U, order, unif, steps, threads=cores)
    t_l0 = time.perf_counter() - t0  # wall time of one sample per thread

    def work(l):
        return sum(2 ** k * (10 * k - 12) for k in range(2, l + 1))

    mean = sum(math.comb(n, l) * eta ** l * (1 - eta) ** (n - l) * t_l0 * work(l) / work(l0)
               for l in range(n + 1))
    return {"value": cores / mean, "unit": "samples/s", "cores": cores, "kind": "port",
            "est_seconds_per_sample_per_thread": mean,
            "sample": f"oracle C GCC-B sampler, {cores} samples of l = {l0} photons on {cores} "
                      f"threads ({t_l0:.2f} s), scaled by the C&C work to each l and averaged "
                      f"over Binomial({n}, {eta})"}


def lossy(checker, cpu=True):
    """Uniform-loss GCC-B, m=80, n=40 photons in modes 0..39, eta=0.5, 4096 samples."""
    T = D.torch()
    m, n, S, eta = 80, 40, 4096, 0.5
    Uh = checker.haar_unitary(m, 5)
    U = D.to_device_c128(Uh)
    parts = D.to_device_i32(np.arange(n))
    out = T.empty((S, m), dtype=T.int32, device=U.device)
    ops.gccb_sample(U, parts, S, 1, eta, out=out)  # capture + warm
    T.cuda.synchronize()
    t = _time(lambda: ops.gccb_sample(U, parts, S, 2, eta, out=out), reps=2)
    l = out.sum(1).cpu().numpy()
    cpu_rec = lossy_cpu(checker, Uh, n, eta) if cpu else None
    return {"workload": "uniform-loss GCC-B m=80 n=40 eta=0.5, 4096 samples",
            "samples_per_s": S / t, "ms_per_batch": t * 1e3, "mean_photons": float(l.mean()),
            "max_photons": int(l.max()), "binomial_mean": n * eta, "cpu_baseline": cpu_rec,
            "gpu_vs_cpu_baseline": (S / t) / cpu_rec["value"] if cpu_rec else None}


def fit(checker):
    """Clements mesh m=16 (120 units) fitted by Adam to 10000 GPU samples (n=8) drawn
    from a seeded target unitary: 500 steps, learning rate 0.1 decayed to 0.001 along a
    half cosine, best of 4 random starts."""
    from theboss_amd.interferometer import clements_unitary, fit_clements, negative_log_likelihood, sample_index_tensors

    T = D.torch()
    m, n, S = 16, 8, 10000
    g = T.Generator().manual_seed(7)
    th = (T.rand(120, generator=g, dtype=T.float64) * math.pi / 2).cuda()
    ph = (T.rand(120, generator=g, dtype=T.float64) * 2 * math.pi).cuda()
    target = clements_unitary(th, ph, m).contiguous()
    out, _ = ops.gccb_sample(target, D.to_device_i32(np.arange(n)), S, 3)
    samples = out.cpu().numpy()
    rows, lf = sample_index_tensors(samples, list(range(n)))
    with T.no_grad():
        nll_target = float(negative_log_likelihood(
            target, T.from_numpy(rows).cuda(), T.arange(n, device="cuda"), T.from_numpy(lf).cuda()))
    steps, restarts, ends = 500, 4, []
    t0 = time.perf_counter()
    for r in range(restarts):  # the restarts of fit_clements, one by one to record each end
        _, _, h = fit_clements(samples, list(range(n)), m, steps=steps, lr=0.1, lr_end=0.001,
                               seed=1 + r)
        ends.append(h[-1])
        if r == 0 or h[-1] < hist[-1]:
            hist = h
    T.cuda.synchronize()
    dt = time.perf_counter() - t0
    return {"workload": "Clements m=16 fit by Adam on 10000 samples (n=8)",
            "adam_steps": steps, "restarts": restarts,
            "ms_per_step": 1e3 * dt / (steps * restarts),
            "nll_start": hist[0], "nll_end": hist[-1], "nll_end_per_start": ends,
            "nll_of_target": nll_target}


def hist(checker):
    """Device outcome histogram (sample-based distributions): 2^24 GCC-B samples of n=10
    photons in m=20 modes drawn by the device sampler (16 calls of 2^20), counted over the
    C(29, 10) = 20,030,010 outcomes of generate_possible_states(10, 20)."""
    T = D.torch()
    m, n, S = 20, 10, 1 << 24
    U = D.to_device_c128(checker.haar_unitary(m, 7))
    parts = D.to_device_i32(np.arange(n))
    samples = T.empty((S, m), dtype=T.int32, device=U.device)
    chunk = 1 << 20
    for i in range(S // chunk):
        ops.gccb_sample(U, parts, chunk, 100 + i, out=samples[i * chunk:(i + 1) * chunk])
    # Reference result through the tensor wrapper, then the bare launch timed on
    # preallocated counters (they accumulate over the repetitions).
    counts, outside = ops.outcome_histogram(samples, n)
    lib = L.load()
    cnt = T.zeros(ops.outcome_count(n, m), dtype=T.int32, device=samples.device)
    stream = D.stream_ptr(samples.device)
    t = _time(lambda: L.check(lib.tb_outcome_histogram(D.ptr(samples), m, S, m, n, 0, D.ptr(cnt),
                                                       None, stream)), reps=3)
    read = S * m * 4
    algo = read + S * 4  # one 4-byte counter update per sample
    cpu, cpu_py = _hist_cpu_lines(checker, samples[: 1 << 20].cpu().numpy(), counts, n, m, S)
    return {"workload": "outcome histogram, 2^24 samples, m=20, n=10 (20,030,010 outcomes)",
            "ms": t * 1e3, "samples_per_s": S / t,
            "algorithmic_bytes": algo, "achieved_GBps": algo / t / 1e9,
            "frac_hbm_peak": algo / t / 8e12,
            "counted": int(counts.sum().item()), "outside": outside,
            "distinct_outcomes_hit": int((counts > 0).sum().item()),
            "cpu_baseline": cpu, "cpu_python_estimate": cpu_py}


def _hist_cpu_lines(checker, host_samples, counts, n, m, S):
    """CPU lines of the outcome histogram.
    * cpu_baseline: the same rank-and-count on the host with numpy (each sample's index
      in the reference's outcome order from a binomial table, then bincount), on the first
      2^20 samples, checked against the device counts of those samples' ranks.
    * cpu_python_estimate: the reference's loop (each sample compared with the outcomes
      in order until one matches, bs_sample_based_distribution_calculator.py:92-106, as the
      oracle restates it) costs a fixed time per comparison; that cost is timed on 200
      samples of 4 photons in the same 20 modes and scaled by this run's mean scan length
      (sum over outcomes of count x (index + 1))."""
    comb = np.zeros((n + m + 1, m + 1), dtype=np.int64)
    for a in range(n + m + 1):
        for b in range(min(a, m) + 1):
            comb[a, b] = math.comb(a, b)
    x = host_samples.astype(np.int64)
    t0 = time.perf_counter()
    rank = np.zeros(len(x), dtype=np.int64)
    rem = x.sum(axis=1)
    for i in range(m - 1):
        r = rem - x[:, i] - 1
        ok = r >= 0
        rank[ok] += comb[(r + m - i - 1)[ok], m - i - 1]
        rem = rem - x[:, i]
    host_counts = np.bincount(rank, minlength=math.comb(n + m - 1, n))
    dt = time.perf_counter() - t0
    cpu = {"value": len(x) / dt, "unit": "samples/s", "cores": 1, "kind": "port",
           "sample": f"first {len(x)} samples: numpy rank from a binomial table and bincount "
                     f"({dt:.2f} s)",
           "distinct_ranks": int((host_counts > 0).sum()),
           "ranks_counted_on_device": bool(np.all(counts.cpu().numpy()[host_counts > 0] > 0))}
    small = checker.generate_possible_states(4, m)
    rng = np.random.default_rng(5)
    picks = rng.integers(0, len(small), 200)
    probe = [small[i] for i in picks]
    t0 = time.perf_counter()
    checker.sample_frequencies(probe, small, len(probe))
    per_cmp = (time.perf_counter() - t0) / float(np.sum(picks + 1))
    c = counts.cpu().numpy().astype(np.float64)
    mean_scan = float(np.sum(c * (np.arange(len(c)) + 1.0)) / max(c.sum(), 1.0))
    est = S * mean_scan * per_cmp
    cpu_py = {"value": S / est, "unit": "samples/s", "cores": 1, "kind": "port",
              "est_seconds": est,
              "sample": f"{per_cmp * 1e6:.2f} us per outcome comparison (200 samples of 4 photons "
                        f"in {m} modes), times a mean scan of {mean_scan:.3g} outcomes per sample"}
    return cpu, cpu_py


def bobs(checker):
    """BOBS non-uniform-loss approximation: m=16 Haar modes with transmissivity 0.5..0.95
    by output mode, 8 photons in modes 0..7, 4 approximated modes, 4096 samples in one
    device call; against one sample per call (the reference's per-sample loop)."""
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import NonuniformLossesApproximationStrategy

    T = D.torch()
    m, n, S, k = 16, 8, 4096, 4
    U = checker.haar_unitary(m, 31)
    eta = np.linspace(0.5, 0.95, m)
    lossy = np.sqrt(eta)[:, None] * U
    init = [1] * n + [0] * (m - n)
    strat = NonuniformLossesApproximationStrategy(GPURyserPermanentCalculator(lossy), k, seed=3)
    strat.simulate(init, S)  # capture + warm
    T.cuda.synchronize()
    t0 = time.perf_counter()
    samples = strat.simulate(init, S)
    T.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    one = NonuniformLossesApproximationStrategy(GPURyserPermanentCalculator(lossy), k, seed=4)
    one.simulate(init, 1)
    T.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(16):
        one.simulate(init, 1)
    T.cuda.synchronize()
    t_one = (time.perf_counter() - t0) / 16
    # CPU line: the reference simulates each sample as GCC-B on the 2m-mode dilation of its
    # own interferometer with the photons left by the extracted uniform loss; the oracle's
    # C sampler does the same work on a seeded 2m-mode unitary with photon numbers drawn
    # from the same binomial, 16384 samples over the host threads.
    cores = _cpu_cores()
    rng = np.random.default_rng(9)
    Sc = 16384
    steps = rng.binomial(n, 1.0 - strat._uniform_losses, Sc).astype(np.int32)
    order = np.stack([rng.permutation(n) for _ in range(Sc)]).astype(np.int32)
    W = checker.haar_unitary(2 * m, 32)
    t0 = time.perf_counter()
    checker.gccb_sample_batch(W, order, rng.random((Sc, n)), steps, threads=cores)
    t_cpu = time.perf_counter() - t0
    cpu = {"value": Sc / t_cpu, "unit": "samples/s", "cores": cores, "kind": "port",
           "sample": f"{Sc} GCC-B samples on a {2 * m}-mode unitary by the oracle's C sampler, "
                     f"photons ~ Binomial({n}, {1.0 - strat._uniform_losses:.3f}) ({t_cpu:.2f} s)"}
    return {"workload": "BOBS non-uniform losses, m=16, n=8, 4 approximated modes, 4096 samples",
            "samples_per_s": S / t_batch, "ms_per_batch": t_batch * 1e3,
            "one_per_call_samples_per_s": 1.0 / t_one, "batched_vs_one_per_call": t_one * S / t_batch,
            "mean_photons": float(np.mean([sum(s_) for s_ in samples])),
            "cpu_baseline": cpu, "gpu_vs_cpu_baseline": (S / t_batch) / cpu["value"]}
