"""GPU parity at the full sizes of the five workloads the permanent path is measured on:

* GCC-A, m = 20 modes, n = 10 photons, 200 samples (every layer pmf against the oracle),
  and one 10 x 10 permanent by Ryser, Glynn and brute force;
* single large permanents, 16 blocks per n in {24, 28, 32}: the FP32 column-sum variant
  against the golden oracle values, and the rule that keeps n >= 24 in FP64;
* permanent + gradient of 64 n = 20 blocks: central finite differences through autograd;
* batched GCC-B at m = 400, n = 20 with 4096 samples: bit-exact against the same call in
  slices and against the oracle sampler on a strided subset;
* uniform-loss GCC-B at m = 80, n = 40, eta = 0.5 (forced photon numbers up to 26 against
  the oracle sampler, up to 31 against Glynn permanents; a full 4096-sample Philox run)
  and the Clements-mesh fit at m = 16, n = 8 on 10,000 samples.
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
REL = 1e-10


@pytest.fixture(scope="module")
def T():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from theboss_amd import _lib

    _lib.load()
    return torch


def dev_c(T, a):
    return T.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


def dev_i(T, a):
    return T.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


def _forced(rng, S, n):
    order = np.stack([rng.permutation(n) for _ in range(S)]).astype(np.int32)
    return order, rng.random((S, n))


def _picks_from_trace(U, order, unif, trace, s, steps):
    """Output mode of each step of sample s, from the device's own traced minors: the
    Laplace expansion over all modes and the inverse-CDF draw of the step."""
    m = U.shape[0]
    picks = []
    for k in range(1, steps + 1):
        amp = U[:, order[s, :k]] @ trace[s, k - 1, :k]
        cum = np.cumsum(np.abs(amp) ** 2)
        picks.append(min(int(np.searchsorted(cum, unif[s, k - 1] * cum[-1], side="right")), m - 1))
    return picks


# --------------------------------------------------------------------------- GCC-A
def test_gcca_m20_n10_200_samples_every_layer_pmf(T, oracle):
    """generalized_cliffords_simulation_strategy.py:136-247 at m = 20, n = 10 photons in
    modes 0..9: every layer the strategy memoised while drawing 200 samples equals the
    oracle's layer pmf (C restatement, pinned against the Python one and against the exact
    output distribution) to 1e-10 of the layer's largest entry."""
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import GeneralizedCliffordsSimulationStrategy

    m, n, S = 20, 10, 200
    U = oracle.haar_unitary(m, 1)
    inp = [1] * n + [0] * (m - n)
    strat = GeneralizedCliffordsSimulationStrategy(GPURyserPermanentCalculator(U),
                                                   rng=np.random.default_rng(0))
    samples = strat.simulate(inp, S)
    assert len(samples) == S and all(sum(x) == n for x in samples)
    keys = list(strat.pmfs)
    assert len(keys) > 500  # every prefix of every sample, memoised once
    print(f"GCC-A m={m} n={n}: {len(keys)} layers")
    ref = oracle.gcca_layer_pmf_batch(U, inp, [list(k) for k in keys])
    got = np.stack([np.asarray(strat.pmfs[k], dtype=np.float64) for k in keys])
    err = np.max(np.abs(got - ref), axis=1) / np.max(ref, axis=1)
    assert err.max() <= REL, (keys[int(np.argmax(err))], err.max())


def test_gcca_single_10x10_permanent_ryser_glynn_classic(T, oracle):
    """The workload's single 10 x 10 check: perm(U[0:10, 0:10]) by the GPU Ryser (Guan)
    calculator, the GPU Glynn calculator and brute force (10! terms), as the reference
    cross-checks its calculators (tests/test_bs_permanent_calculators.py:73-86)."""
    from theboss_amd.permanent_calculators import (GPUGlynnGrayPermanentCalculator,
                                                   GPURyserPermanentCalculator)

    U = oracle.haar_unitary(20, 1)
    ones = [1] * 10 + [0] * 10
    r = GPURyserPermanentCalculator(U, ones, ones).compute_permanent()
    g = GPUGlynnGrayPermanentCalculator(U, ones, ones).compute_permanent()
    ref = oracle.perm_classic_matrix(U[:10, :10])
    for v in (r, g):
        assert abs(v - ref) <= REL * abs(ref), (v, ref)


# --------------------------------------------------------------------------- single large
def _golden(oracle, n):
    with open(os.path.join(HERE, "golden", "glynn_large.json")) as f:
        cases = [c for c in json.load(f)["cases"] if c["n"] == n]
    mats = []
    for c in cases:
        seed = 1000 * n + c["b"]
        U = oracle.haar_unitary(2 * n, seed)
        rows = np.sort(np.random.default_rng(seed).choice(2 * n, n, replace=False))
        mats.append(U[rows, :n])
    return np.stack(mats), np.array([complex(c["re"], c["im"]) for c in cases])


# FP32 column sums with FP64 accumulation: max relative error over the 16 golden blocks,
# measured on MI355X (7.4e-5, 3.3e-3, 4.3e-3), bounded at about 3x.
FP32_BOUND = {24: 2.2e-4, 28: 1.0e-2, 32: 1.3e-2}


@pytest.mark.parametrize("n", [24, 28, 32])
def test_glynn_fp32_colsum_error_vs_golden(T, oracle, n):
    """The single-large-permanent workload's FP32 leg, 16 golden blocks per n: the FP32
    column-sum kernel (forced at every n) against the oracle, per-n bound; the default
    FP32 request at these n runs FP64 and meets the 1e-10 bar."""
    from theboss_amd import _lib, ops

    A, ref = _golden(oracle, n)
    assert len(ref) == 16
    Ad = dev_c(T, A)
    f32 = ops.glynn_batch(Ad, _lib.TB_PREC_FP32_COLSUM_ANY_N).cpu().numpy()
    err = np.abs(f32 - ref) / np.abs(ref)
    print(f"fp32-colsum n={n}: max rel err {err.max():.2e}, median {np.median(err):.2e}")
    assert err.max() <= FP32_BOUND[n]
    promoted = ops.glynn_batch(Ad, _lib.TB_PREC_FP32_COLSUM).cpu().numpy()
    assert np.max(np.abs(promoted - ref) / np.abs(ref)) <= REL


def test_fp32_colsum_rule_boundary_n23_n24(T, oracle):
    """FP32 column sums are used up to n = 23 only: at n = 23 the FP32 request runs the
    FP32 kernel (same values as the forced variant, not the FP64 ones); at n = 24 it runs
    the FP64 kernel (bit-identical to the FP64 request), and so does the GLYNN_FP32
    calculator."""
    from theboss_amd import _lib, ops
    from theboss_amd.permanent_calculators import (BSPermanentCalculatorFactory,
                                                   PermanentCalculatorType)

    for n in (23, 24):
        A = np.stack([oracle.haar_unitary(2 * n, 6000 + 7 * n + b)[:n, :n] for b in range(4)])
        Ad = dev_c(T, A)
        f64 = ops.glynn_batch(Ad, _lib.TB_PREC_FP64).cpu().numpy()
        f32 = ops.glynn_batch(Ad, _lib.TB_PREC_FP32_COLSUM).cpu().numpy()
        anyn = ops.glynn_batch(Ad, _lib.TB_PREC_FP32_COLSUM_ANY_N).cpu().numpy()
        ref = oracle.perm_glynn_batch(A, 1)
        assert np.max(np.abs(f64 - ref) / np.abs(ref)) <= REL
        assert not np.array_equal(anyn, f64)  # the FP32 kernel really ran
        if n == 23:
            assert np.array_equal(f32, anyn)
            assert np.max(np.abs(f32 - ref) / np.abs(ref)) <= FP32_SMALL_BOUND[23]
        else:
            assert np.array_equal(f32, f64)
    U = oracle.haar_unitary(48, 6100)
    ones = [1] * 24 + [0] * 24
    calc = BSPermanentCalculatorFactory(U, ones, ones,
                                        PermanentCalculatorType.GLYNN_FP32).generate_calculator()
    ref64 = BSPermanentCalculatorFactory(U, ones, ones,
                                         PermanentCalculatorType.GLYNN).generate_calculator()
    assert calc.compute_permanent() == ref64.compute_permanent()


# FP32 column sums below the rule's limit: max relative error of these 8 Haar blocks
# against the oracle, bounded at about 3x the value measured on MI355X. The error varies
# from block to block with the cancellation of the Glynn sum: over 64 blocks (seeds
# 5000 + n + b) the median / 90th percentile / max measured 9e-7 / 1.8e-6 / 4.7e-6 (n = 8,
# one lane per matrix), 7.8e-7 / 2.3e-6 / 6.5e-6 (n = 12, plain walk), and with the paired
# walk 1.7e-6 / 5.9e-6 / 2.9e-5 (n = 16), 4.0e-6 / 1.1e-5 / 2.1e-5 (n = 20) and
# 9.3e-6 / 2.6e-5 / 8.1e-5 (n = 23).
FP32_SMALL_BOUND = {8: 6e-6, 12: 6e-6, 16: 5e-5, 20: 7e-5, 23: 9e-5}


@pytest.mark.parametrize("n", sorted(FP32_SMALL_BOUND))
def test_glynn_fp32_colsum_error_table_small_n(T, oracle, n):
    from theboss_amd import _lib, ops

    A = np.stack([oracle.haar_unitary(2 * n, 5000 + n + b)[:n, :n] for b in range(8)])
    ref = oracle.perm_glynn_batch(A, 1)
    got = ops.glynn_batch(dev_c(T, A), _lib.TB_PREC_FP32_COLSUM).cpu().numpy()
    err = np.abs(got - ref) / np.abs(ref)
    print(f"fp32-colsum n={n}: max rel err {err.max():.2e}, median {np.median(err):.2e}")
    assert err.max() <= FP32_SMALL_BOUND[n]


# --------------------------------------------------------------------------- gradient
def test_grad_n20_64_blocks_central_differences(T, oracle):
    """Differentiation workload (64 blocks n = 20 of Haar(40)) through the autograd
    Function: d/dA of L = -sum_b log|perm(A_b)|^2 against central differences with
    h = 1e-6 on the real and imaginary part of 16 sampled entries, each L evaluated by
    the oracle's double-double Glynn."""
    from theboss_amd.autograd import permanent

    n, B, h = 20, 64, 1e-6
    blocks = []
    for b in range(B):
        rows = np.sort(np.random.default_rng(300 + b).choice(40, n, replace=False))
        blocks.append(oracle.haar_unitary(40, 300 + b)[rows, :n])
    A0 = np.stack(blocks)
    A = dev_c(T, A0).requires_grad_(True)
    p = permanent(A)
    loss = -T.log(p.real ** 2 + p.imag ** 2).sum()
    loss.backward()
    g = A.grad.cpu().numpy()
    rng = np.random.default_rng(20)
    picks = [(int(rng.integers(B)), int(rng.integers(n)), int(rng.integers(n))) for _ in range(16)]
    pert = []
    for b, i, j in picks:
        for d in (h, -h, 1j * h, -1j * h):
            M = A0[b].copy()
            M[i, j] += d
            pert.append(M)
    vals = oracle.perm_glynn_batch(np.stack(pert), 1)
    L = -np.log(np.abs(vals) ** 2).reshape(len(picks), 4)
    for (b, i, j), l4 in zip(picks, L):
        fd = complex((l4[0] - l4[1]) / (2 * h), (l4[2] - l4[3]) / (2 * h))
        assert abs(g[b, i, j] - fd) <= 1e-6 * max(1.0, abs(g[b, i, j])), (b, i, j, g[b, i, j], fd)


# --------------------------------------------------------------------------- GCC-B, m=400
def test_gccb_m400_n20_4096_samples_slices_and_oracle(T, oracle):
    """The benchmark workload with forced draws: 4096 samples at m = 400, n = 20 (16384
    sweep tasks per step at k = 20 on the 8192-block persistent grid, several tasks per
    block) equal the same draws run as 16 calls of 256 samples, and a strided subset of 64
    samples equals the oracle's GCC-B sampler. The chunk sizes of the 4096-sample plan are
    passed to the sliced calls, so both sum the minors in the same order."""
    from theboss_amd import autotune, ops

    m, n, S = 400, 20, 4096
    U = oracle.haar_unitary(m, 20240611)
    rng = np.random.default_rng(4096)
    order, unif = _forced(rng, S, n)
    Ud, parts = dev_c(T, U), dev_i(T, np.arange(n))
    chunks = [0, 0] + [autotune.sweep_builtin(k, S) for k in range(2, n + 1)]
    full, _ = ops.gccb_sample(Ud, parts, S, 1, 1.0, dev_i(T, order), T.from_numpy(unif).cuda())
    full = full.cpu().numpy()
    assert np.all(full.sum(axis=1) == n)
    parts_out = []
    for lo in range(0, S, 256):
        o, _ = ops.gccb_sample(Ud, parts, 256, 1, 1.0, dev_i(T, order[lo:lo + 256]),
                               T.from_numpy(unif[lo:lo + 256].copy()).cuda(), chunk_bits=chunks)
        parts_out.append(o.cpu().numpy())
    assert np.array_equal(np.concatenate(parts_out), full)
    idx = np.arange(0, S, 64)
    expect = oracle.gccb_sample_batch(U, order[idx], unif[idx])
    assert np.array_equal(full[idx], expect)


# --------------------------------------------------------------------------- uniform loss
def test_uniform_loss_m80_n40_forced_steps_vs_oracle(T, oracle):
    """Uniform-loss GCC-B (generalized_cliffords_b_uniform_losses_simulation_strategy.py:
    50-121) at m = 80, n = 40, eta = 0.5 with forced photon numbers l in {0, 1, 12, 20, 24,
    26} (the binomial's bulk and upper tail): samples bit-exact against the oracle sampler;
    the l = 26 sample's minors at k = 22 and 26 against oracle-grade Glynn minors."""
    from theboss_amd import ops

    m, n, eta = 80, 40, 0.5
    U = oracle.haar_unitary(m, 5)
    steps = np.array([0, 1, 12, 20, 24, 26, 20, 12], dtype=np.int32)
    S = len(steps)
    order, unif = _forced(np.random.default_rng(80), S, n)
    out, trace = ops.gccb_sample(dev_c(T, U), dev_i(T, np.arange(n)), S, 1, eta,
                                 dev_i(T, order), T.from_numpy(unif).cuda(), dev_i(T, steps),
                                 trace=True)
    out, tr = out.cpu().numpy(), trace.cpu().numpy()
    assert np.array_equal(out.sum(axis=1), steps)
    assert np.array_equal(out, oracle.gccb_sample_batch(U, order, unif, steps))
    s = 5  # l = 26
    picks = _picks_from_trace(U, order, unif, tr, s, 26)
    assert np.array_equal(np.bincount(picks, minlength=m), out[s])
    for k in (22, 26):
        cur = np.zeros(m, int)
        cur[order[s, :k]] = 1
        r = np.bincount(picks[:k - 1], minlength=m)
        ref = oracle.cc_submatrices_glynn(U, cur, r)
        want = ref[order[s, :k]]
        got = tr[s, k - 1, :k]
        assert np.max(np.abs(got - want) / np.abs(want)) <= REL, k


def test_uniform_loss_m80_n40_step31_minors_vs_glynn(T, oracle):
    """Photon numbers beyond what the oracle sampler reaches (l = 31, 2^30 Gray steps in
    the last sweep): the traced minors of steps 28 and 31 equal the permanents of the same
    (k-1) x (k-1) effective matrices by the FP64 Glynn kernel (pinned up to n = 32 by the
    golden oracle values), and the Laplace expansion of each step reproduces its pick."""
    from theboss_amd import ops

    m, n, eta = 80, 40, 0.5
    U = oracle.haar_unitary(m, 5)
    order, unif = _forced(np.random.default_rng(31), 1, n)
    steps = np.array([31], dtype=np.int32)
    out, trace = ops.gccb_sample(dev_c(T, U), dev_i(T, np.arange(n)), 1, 1, eta,
                                 dev_i(T, order), T.from_numpy(unif).cuda(), dev_i(T, steps),
                                 trace=True)
    out, tr = out.cpu().numpy(), trace.cpu().numpy()
    assert out.sum() == 31
    picks = _picks_from_trace(U, order, unif, tr, 0, 31)
    assert np.array_equal(np.bincount(picks, minlength=m), out[0])
    for k in (28, 31):
        rows = picks[:k - 1]
        mats = []
        for nu in range(k):
            cols = np.delete(order[0, :k], nu)
            mats.append(U[np.ix_(rows, cols)])
        ref = ops.glynn_batch(dev_c(T, np.stack(mats))).cpu().numpy()
        got = tr[0, k - 1, :k]
        assert np.max(np.abs(got - ref) / np.abs(ref)) <= REL, k


def test_uniform_loss_m80_n40_4096_samples_device_rng(T, oracle):
    """The uniform-loss workload as benchmarked (device Philox draws, 4096 samples): every
    sample is a valid occupation with at most 40 photons, and the photon numbers follow
    Binomial(40, 0.5) within the reference's TVD bound (quantum_computations_utilities.py:
    95-127) and 5 standard errors of the mean."""
    from theboss_amd import ops

    m, n, eta, S = 80, 40, 0.5, 4096
    U = oracle.haar_unitary(m, 5)
    out, _ = ops.gccb_sample(dev_c(T, U), dev_i(T, np.arange(n)), S, 2024, eta)
    out = out.cpu().numpy()
    assert out.min() >= 0
    l = out.sum(axis=1)
    assert l.max() <= n
    emp = np.bincount(l, minlength=n + 1) / S
    binom = np.array([math.comb(n, k) * eta ** k * (1 - eta) ** (n - k) for k in range(n + 1)])
    assert 0.5 * np.abs(emp - binom).sum() < oracle.tvd_error_bound(n + 1, S, 1e-4)
    assert abs(l.mean() - n * eta) < 5 * math.sqrt(n * eta * (1 - eta) / S)


# --------------------------------------------------------------------------- fitting
def test_clements_fit_m16_n8_10000_samples(T, oracle):
    """Interferometer fitting at its configured size: a Clements mesh of m = 16 modes (120
    beam splitters) is fitted by Adam, through the permanent gradient kernel, to 10,000
    samples of n = 8 photons drawn by the GPU sampler from a seeded target mesh. 500 Adam
    steps with the learning rate decayed from 0.1 to 0.001, best of 4 random starts (some
    starts settle in a basin near 11.8 nats): the negative log-likelihood ends within 0.1
    nat of the target's own (11.02)."""
    from theboss_amd import ops
    from theboss_amd.interferometer import (clements_unitary, fit_clements,
                                            negative_log_likelihood, sample_index_tensors)

    m, n, S = 16, 8, 10000
    g = T.Generator().manual_seed(7)
    th = (T.rand(120, generator=g, dtype=T.float64) * math.pi / 2).cuda()
    ph = (T.rand(120, generator=g, dtype=T.float64) * 2 * math.pi).cuda()
    target = clements_unitary(th, ph, m).contiguous()
    out, _ = ops.gccb_sample(target, T.arange(n, dtype=T.int32, device="cuda"), S, 3)
    samples = out.cpu().numpy()
    assert np.all(samples.sum(axis=1) == n)
    rows, lf = sample_index_tensors(samples, list(range(n)))
    with T.no_grad():
        nll_target = float(negative_log_likelihood(
            target, T.from_numpy(rows).cuda(), T.arange(n, device="cuda"), T.from_numpy(lf).cuda()))
    _, _, hist = fit_clements(samples, list(range(n)), m, steps=500, lr=0.1, lr_end=0.001,
                              seed=1, restarts=4)
    print(f"Clements fit: NLL {hist[0]:.3f} -> {hist[-1]:.3f} (target {nll_target:.3f})")
    assert hist[-1] < hist[0] - 1.0
    assert hist[-1] < nll_target + 0.1
