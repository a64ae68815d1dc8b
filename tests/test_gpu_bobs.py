"""Per-sample interferometers in the device sampler and the BOBS approximation strategies
built on them (run on an MI355X)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from theboss_amd import _lib

    _lib.load()
    return torch


def test_per_sample_sampler_matches_oracle_with_forced_draws(T, oracle):
    """Each sample has its own (lossy) interferometer, its own particles (collisions
    allowed, 0..5 of them) and forced draws: outputs bit-exact against the oracle sampler
    run sample by sample."""
    from theboss_amd import _device as D
    from theboss_amd import ops

    rng = np.random.default_rng(41)
    S, m, n = 40, 8, 5
    Us = np.stack([oracle.haar_unitary(m, 500 + s) * (1.0 if s % 3 else 0.8) for s in range(S)])
    steps = rng.integers(0, n + 1, S).astype(np.int32)
    parts = np.zeros((S, n), dtype=np.int32)
    order = np.zeros((S, n), dtype=np.int32)
    for s in range(S):
        parts[s, : steps[s]] = np.sort(rng.integers(0, m, steps[s]))
        order[s] = np.r_[rng.permutation(steps[s]), np.arange(steps[s], n)]
    unif = rng.random((S, n))
    dev = D.device()
    out, _ = ops.gccb_sample(D.to_device_c128(Us, dev), D.to_device_i32(parts, dev), S, 1,
                             forced_order=D.to_device_i32(order, dev),
                             forced_uniforms=T.from_numpy(unif).to(dev),
                             forced_steps=D.to_device_i32(steps, dev))
    out = out.cpu().numpy()
    for s in range(S):
        modes = parts[s][order[s]]  # particle modes in the order they enter
        ref = oracle.gccb_sample_batch(Us[s], modes[None, :], unif[s][None, :], steps=[steps[s]])
        assert np.array_equal(out[s], ref[0]), s
        assert out[s].sum() == steps[s]


def test_per_sample_sampler_rejects_bad_arguments(T, oracle):
    from theboss_amd import _device as D
    from theboss_amd import ops

    dev = D.device()
    Us = D.to_device_c128(np.stack([oracle.haar_unitary(4, s) for s in range(3)]), dev)
    parts = D.to_device_i32(np.zeros((3, 2), dtype=np.int32), dev)
    with pytest.raises(ValueError):
        ops.gccb_sample(Us, parts, 3, 1)  # per-sample lists need forced_steps
    with pytest.raises(ValueError):
        ops.gccb_sample(Us, parts, 4, 1, forced_steps=D.to_device_i32(np.ones(4, np.int32), dev))


def _exact_uniform_loss(oracle, U, init, eta):
    m, n = len(init), sum(init)
    outs = oracle.generate_possible_states(n, m, losses=True)
    return outs, np.array(oracle.uniform_losses_distribution(U, init, eta))


def _tvd_check(samples, outs, exact, p_err=1e-4):
    import collections

    from theboss_amd.quantum_computations_utilities import (
        count_total_variation_distance, count_tv_distance_error_bound_of_experiment_results)

    c = collections.Counter(samples)
    freq = np.array([c[o] / len(samples) for o in outs])
    assert sum(c[o] for o in outs) == len(samples)  # every sample is a listed outcome
    tvd = count_total_variation_distance(freq, exact)
    bound = count_tv_distance_error_bound_of_experiment_results(len(outs), len(samples), p_err)
    assert tvd <= bound, (tvd, bound)


def test_lossy_state_bobs_exact_level_matches_uniform_loss_distribution(T, oracle):
    """hierarchy_level = m (nothing approximated): the distribution is the exact
    uniform-loss one. The strategy permutes U's columns for every sample
    (lossy_state_approximated_simulation_strategy.py:333-345), which leaves the
    distribution unchanged only for an input that occupies every mode alike, so that is
    the input here."""
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import LossyStateApproximationSimulationStrategy

    m, eta = 5, 0.6
    U = oracle.haar_unitary(m, 77)
    init = [1, 1, 1, 1, 1]
    strat = LossyStateApproximationSimulationStrategy(GPURyserPermanentCalculator(U), eta, m, seed=5)
    samples = strat.simulate(init, 60000)
    outs, exact = _exact_uniform_loss(oracle, U, init, eta)
    _tvd_check(samples, outs, exact)


def test_nonuniform_bobs_without_approximation_on_uniform_losses(T, oracle):
    """A uniformly lossy sqrt(eta) U with no approximated modes: the extracted uniform loss
    is 1 - eta and the samples follow the exact uniform-loss distribution."""
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import NonuniformLossesApproximationStrategy

    m, eta = 4, 0.7
    U = oracle.haar_unitary(m, 78)
    init = [1, 1, 0, 1]
    strat = NonuniformLossesApproximationStrategy(GPURyserPermanentCalculator(np.sqrt(eta) * U), 0, seed=6)
    samples = strat.simulate(init, 60000)
    outs, exact = _exact_uniform_loss(oracle, U, init, eta)
    _tvd_check(samples, outs, exact)


def test_nonuniform_bobs_block_diagonal_losses(T, oracle):
    """Two independent 2-mode blocks with transmissivities 0.5 and 0.2 (the reference's
    non-uniform losses test): the joint distribution is the product of the blocks'
    uniform-loss distributions."""
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import NonuniformLossesApproximationStrategy

    B1, B2 = oracle.haar_unitary(2, 11), oracle.haar_unitary(2, 12)
    etas = (0.5, 0.2)
    U = np.zeros((4, 4), complex)
    U[:2, :2] = np.sqrt(etas[0]) * B1
    U[2:, 2:] = np.sqrt(etas[1]) * B2
    init = [1, 1, 1, 1]
    strat = NonuniformLossesApproximationStrategy(GPURyserPermanentCalculator(U), 0, seed=7)
    samples = strat.simulate(init, 60000)
    o1, p1 = _exact_uniform_loss(oracle, B1, [1, 1], etas[0])
    o2, p2 = _exact_uniform_loss(oracle, B2, [1, 1], etas[1])
    outs = [tuple(a) + tuple(b) for a in o1 for b in o2]
    exact = np.array([x * y for x in p1 for y in p2])
    _tvd_check(samples, outs, exact)


def test_bobs_approximate_levels_run(T, oracle):
    """Approximated modes (symmetrisation, QFT and phases) produce samples with the right
    photon counts for every approximation level."""
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import (LossyStateApproximationSimulationStrategy,
                                                   NonuniformLossesApproximationStrategy)

    m = 6
    U = oracle.haar_unitary(m, 80)
    init = [1, 1, 1, 1, 0, 0]
    for h in range(m + 1):
        s = LossyStateApproximationSimulationStrategy(GPURyserPermanentCalculator(U), 1.0, h, seed=h)
        for smp in s.simulate(init, 50):
            assert sum(smp) == 4 and len(smp) == m
    for k in range(m + 1):
        s = NonuniformLossesApproximationStrategy(GPURyserPermanentCalculator(0.9 * U), k, seed=k)
        for smp in s.simulate(init, 50):
            assert sum(smp) <= 4 and len(smp) == m


def test_nonuniform_bobs_approximation_within_paper_bound(T, oracle):
    """Approximating k modes: the TVD to the exact uniform-loss distribution stays within
    the BOBS bound eta^2 (n - h) / 2 + eta (1 - eta) / 2 with h = m - k (the reference's
    test_lossy_approximate_sampling_accuracy) plus the statistical bound."""
    import collections

    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.quantum_computations_utilities import (
        count_total_variation_distance, count_tv_distance_error_bound_of_experiment_results)
    from theboss_amd.simulation_strategies import NonuniformLossesApproximationStrategy

    m, eta, k = 4, 0.5, 2
    U = oracle.haar_unitary(m, 79)
    init = [1, 1, 1, 1]
    n, h = sum(init), m - k
    strat = NonuniformLossesApproximationStrategy(GPURyserPermanentCalculator(np.sqrt(eta) * U), k, seed=8)
    samples = strat.simulate(init, 40000)
    outs, exact = _exact_uniform_loss(oracle, U, init, eta)
    c = collections.Counter(samples)
    freq = np.array([c[o] / len(samples) for o in outs])
    bound = eta ** 2 * (n - h) / 2 + eta * (1 - eta) / 2
    stat = count_tv_distance_error_bound_of_experiment_results(len(outs), len(samples), 1e-4)
    assert count_total_variation_distance(freq, exact) <= bound + stat
