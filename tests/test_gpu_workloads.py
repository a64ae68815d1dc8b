"""GPU tests of the workloads built on the kernels: full-size Glynn against golden oracle
values, the differentiable permanent, GCC-A, lossy GCC-B strategies and the
interferometer fit."""
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


# ------------------------------------------------------------- full-size Glynn parity
def _golden_block(oracle, n, b):
    seed = 1000 * n + b
    U = oracle.haar_unitary(2 * n, seed)
    rows = np.sort(np.random.default_rng(seed).choice(2 * n, n, replace=False))
    return U[rows, :n]


@pytest.mark.parametrize("n", [24, 28, 32])
def test_glynn_fp64_full_size_golden(T, oracle, n):
    from theboss_amd import ops

    with open(os.path.join(HERE, "golden", "glynn_large.json")) as f:
        cases = [c for c in json.load(f)["cases"] if c["n"] == n]
    A = np.stack([_golden_block(oracle, n, c["b"]) for c in cases])
    got = ops.glynn_batch(dev_c(T, A)).cpu().numpy()
    for g, c in zip(got, cases):
        ref = complex(c["re"], c["im"])
        assert abs(g - ref) <= REL * abs(ref), (n, c["b"], g, ref, abs(g - ref) / abs(ref))


# ------------------------------------------------------------- differentiation
def test_autograd_matches_central_differences(T, oracle):
    from theboss_amd.autograd import permanent

    n = 6
    A0 = oracle.haar_unitary(12, 3)[:n, :n]
    A = dev_c(T, A0).requires_grad_(True)
    p = permanent(A)
    loss = -T.log(p.real ** 2 + p.imag ** 2)
    loss.backward()
    g = A.grad.cpu().numpy()
    h = 1e-6

    def L(M):
        v = oracle.perm_glynn_matrix(M, 1)
        return -math.log(abs(v) ** 2)

    for i in range(n):
        for j in range(n):
            E = np.zeros((n, n), complex)
            E[i, j] = h
            d_re = (L(A0 + E) - L(A0 - E)) / (2 * h)
            d_im = (L(A0 + 1j * E) - L(A0 - 1j * E)) / (2 * h)
            assert abs(g[i, j] - complex(d_re, d_im)) <= 1e-6 * max(1.0, abs(g[i, j])), (i, j)


def test_grad_rows_equal_reference_submatrices_calculator(T, oracle):
    """Row i of d perm / dA = C&C minors with output rows = all rows but i (the reference
    submatrices calculator's definition), 64 n=20 blocks of Haar(40) as in the
    differentiation workload (checked on 4 of them)."""
    from theboss_amd import ops

    n = 20
    blocks = []
    for b in range(64):
        rows = np.sort(np.random.default_rng(300 + b).choice(40, n, replace=False))
        blocks.append(oracle.haar_unitary(40, 300 + b)[rows, :n])
    A = np.stack(blocks)
    perm, G = ops.perm_grad_batch(dev_c(T, A))
    G = G.cpu().numpy()
    perm = perm.cpu().numpy()
    for b in (0, 21, 42, 63):
        Gref = oracle.perm_minors(A[b])  # n^2 independent Glynn permanents, double-double
        assert np.max(np.abs(G[b] - Gref)) <= REL * np.max(np.abs(Gref)), b
        for i in (0, 7, 19):
            t = np.ones(n, int)
            t[i] = 0
            # The calculator indexes U[output][input]: outputs = rows of A, inputs = columns.
            # Its Ryser walk carries ~1e-10 error of its own at this size (see oracle.py).
            ref = oracle.cc_submatrices(A[b], np.ones(n, int), t)
            assert np.max(np.abs(G[b, i] - ref)) <= 1e-9 * np.max(np.abs(ref)), (b, i)
        lap = np.array([np.sum(A[b, i] * G[b, i]) for i in range(n)])
        assert np.max(np.abs(lap - perm[b])) <= REL * abs(perm[b])


def test_autograd_forward_same_with_and_without_grad(T, oracle):
    """With requires_grad the forward runs the gradient launch (perm from the Laplace
    expansion of the minors); without, the Glynn kernel. Both meet the oracle bar."""
    from theboss_amd.autograd import permanent

    n = 16
    As = np.stack([oracle.haar_unitary(2 * n, 40 + b)[:n, :n] for b in range(4)])
    ref = oracle.perm_glynn_batch(As, 1)
    p0 = permanent(dev_c(T, As)).cpu().numpy()
    A = dev_c(T, As).requires_grad_(True)
    p1 = permanent(A)
    p1.real.sum().backward()
    p1 = p1.detach().cpu().numpy()
    for got in (p0, p1):
        assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-10
    G = np.stack([oracle.perm_minors(a) for a in As])
    assert np.max(np.abs(A.grad.cpu().numpy() - G.conj())) <= 1e-10 * np.max(np.abs(G))


def test_autograd_batch_shapes(T, oracle):
    from theboss_amd.autograd import permanent

    A = dev_c(T, np.stack([oracle.haar_unitary(8, s)[:4, :4] for s in range(6)]).reshape(2, 3, 4, 4))
    A.requires_grad_(True)
    p = permanent(A)
    assert p.shape == (2, 3)
    p.abs().sum().backward()
    assert A.grad.shape == A.shape


# ------------------------------------------------------------- GCC-A
def test_gcca_layer_pmfs_match_oracle(T, oracle):
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import GeneralizedCliffordsSimulationStrategy

    m = 6
    U = oracle.haar_unitary(m, 8)
    inp = [1, 2, 0, 1, 0, 0]
    strat = GeneralizedCliffordsSimulationStrategy(GPURyserPermanentCalculator(U),
                                                   rng=np.random.default_rng(0))
    strat.simulate(inp, 20)
    assert strat.pmfs
    for key, pmf in strat.pmfs.items():
        ref = oracle.gcca_layer_pmf(U, inp, list(key))
        assert np.allclose(pmf, ref, rtol=1e-10, atol=1e-15), key


def test_gcca_sampling_statistics(T, oracle):
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import GeneralizedCliffordsSimulationStrategy

    m, S = 4, 3000
    U = oracle.haar_unitary(m, 9)
    inp = [1, 1, 0, 0]
    strat = GeneralizedCliffordsSimulationStrategy(GPURyserPermanentCalculator(U),
                                                   rng=np.random.default_rng(1))
    samples = strat.simulate(inp, S)
    states = oracle.generate_possible_states(2, m)
    freq = np.array([sum(1 for s in samples if s == st) for st in states]) / S
    exact = np.array([oracle.fixed_losses_probability(U, inp, st, oracle.perm_ryser) for st in states])
    assert 0.5 * np.abs(freq - exact).sum() < oracle.tvd_error_bound(len(states), S, 1e-4)


# ------------------------------------------------------------- lossy GCC-B
def test_uniform_loss_strategy_matches_exact_distribution(T, oracle):
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import GeneralizedCliffordsBUniformLossesSimulationStrategy

    m, eta, S = 4, 0.6, 20000
    U = oracle.haar_unitary(m, 12)
    inp = [1, 1, 1, 0]
    strat = GeneralizedCliffordsBUniformLossesSimulationStrategy(GPURyserPermanentCalculator(U), eta, seed=3)
    samples = strat.simulate(inp, S)
    assert all(isinstance(s, np.ndarray) and s.dtype == np.int64 for s in samples)  # as the reference
    states = oracle.generate_possible_states(3, m, losses=True)
    idx = {st: i for i, st in enumerate(states)}
    freq = np.zeros(len(states))
    for s in samples:
        freq[idx[tuple(int(x) for x in s)]] += 1
    freq /= S
    exact = np.array(oracle.uniform_losses_distribution(U, inp, eta, oracle.perm_ryser))
    assert 0.5 * np.abs(freq - exact).sum() < oracle.tvd_error_bound(len(states), S, 1e-4)


def test_lossy_network_strategy_known_answer(T, oracle):
    """The reference's deterministic known-answer distribution (permutation matrix,
    eta = 0.8), sampled through the 2m expanded-space strategy."""
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import LossyNetworksGeneralizedCliffordsSimulationStrategy

    with open(os.path.join(HERE, "golden", "exact_distribution_kat.json")) as f:
        kat = json.load(f)
    U = np.asarray(kat["interferometer_real"], dtype=np.complex128) * math.sqrt(kat["uniform_transmissivity"])
    strat = LossyNetworksGeneralizedCliffordsSimulationStrategy(GPURyserPermanentCalculator(U), seed=4)
    S = 20000
    samples = strat.simulate(kat["initial_state"], S)
    states = oracle.generate_possible_states(3, 5, losses=True)
    idx = {st: i for i, st in enumerate(states)}
    freq = np.zeros(len(states))
    for s in samples:
        freq[idx[tuple(int(x) for x in s)]] += 1
    freq /= S
    expect = np.asarray(kat["expected_uniform_losses_distribution"])
    assert 0.5 * np.abs(freq - expect).sum() < oracle.tvd_error_bound(len(states), S, 1e-4)


# ------------------------------------------------------------- interferometer fitting
def test_clements_fit_reduces_nll(T, oracle):
    from theboss_amd import ops
    from theboss_amd.interferometer import clements_unitary, fit_clements

    m, n, S = 6, 3, 2000
    g = T.Generator().manual_seed(0)
    th = (T.rand(15, generator=g, dtype=T.float64) * math.pi / 2).cuda()
    ph = (T.rand(15, generator=g, dtype=T.float64) * 2 * math.pi).cuda()
    target = clements_unitary(th, ph, m)
    assert T.allclose(target.conj().T @ target, T.eye(m, dtype=T.complex128, device="cuda"), atol=1e-12)
    out, _ = ops.gccb_sample(target.contiguous(), T.arange(n, dtype=T.int32, device="cuda"), S, 11)
    samples = out.cpu().numpy()
    _, _, hist = fit_clements(samples, list(range(n)), m, steps=60, lr=0.05, seed=1)
    assert hist[-1] < hist[0] - 0.05, hist[::10]
    # restarts: the run with the lowest final NLL among the starts seed, seed + 1, ...
    ends = [fit_clements(samples, list(range(n)), m, steps=20, lr=0.1, lr_end=0.01, seed=s)[2][-1]
            for s in (4, 5, 6)]
    th3, ph3, h3 = fit_clements(samples, list(range(n)), m, steps=20, lr=0.1, lr_end=0.01, seed=4,
                                restarts=3)
    assert len(h3) == 20 and abs(h3[-1] - min(ends)) <= 1e-9 * abs(min(ends)), (h3[-1], ends)
    with pytest.raises(ValueError):
        fit_clements(samples, list(range(n)), m, steps=1, restarts=0)


# ------------------------------------------------------------- exact distributions
def test_exact_distribution_known_answer_on_gpu(T, oracle):
    """tests/test_exact_distribution_calculator.py:75-142 of the reference, with every
    permanent from the GPU Chin-Huh and Ryser calculators."""
    from theboss_amd.distribution_calculators import (
        BosonSamplingExperimentConfiguration, BSDistributionCalculatorWithFixedLosses,
        BSDistributionCalculatorWithUniformLosses)
    from theboss_amd.permanent_calculators import BSPermanentCalculatorFactory, PermanentCalculatorType

    with open(os.path.join(HERE, "golden", "exact_distribution_kat.json")) as f:
        kat = json.load(f)
    U = np.asarray(kat["interferometer_real"], dtype=np.complex128)
    cfg = BosonSamplingExperimentConfiguration(
        interferometer_matrix=U, initial_state=kat["initial_state"], number_of_modes=5,
        initial_number_of_particles=3, number_of_particles_lost=2, number_of_particles_left=1,
        uniform_transmissivity=kat["uniform_transmissivity"])
    for ctype in (PermanentCalculatorType.CHIN_HUH, PermanentCalculatorType.RYSER):
        calc = BSPermanentCalculatorFactory(None, None, None, ctype).generate_calculator()
        dist = BSDistributionCalculatorWithUniformLosses(cfg, calc).calculate_distribution()
        assert np.allclose(dist, kat["expected_uniform_losses_distribution"])
        fixed = BSDistributionCalculatorWithFixedLosses(cfg, calc).calculate_distribution()
        assert abs(sum(fixed) - 1.0) < kat["probability_sum_tolerance"]


def test_exact_distribution_haar_matches_oracle(T, oracle):
    from theboss_amd.distribution_calculators import (
        BosonSamplingExperimentConfiguration, BSDistributionCalculatorWithUniformLosses)
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator

    U = oracle.haar_unitary(4, 77)
    init = [2, 1, 0, 0]
    cfg = BosonSamplingExperimentConfiguration(U, init, 3, 4, 0, 3, 0.6)
    dist = BSDistributionCalculatorWithUniformLosses(cfg, GPURyserPermanentCalculator(U)).calculate_distribution()
    ref = oracle.uniform_losses_distribution(U, init, 0.6, oracle.perm_ryser)
    assert np.allclose(dist, ref, rtol=1e-10, atol=1e-14)
    assert abs(sum(dist) - 1) < 1e-12
