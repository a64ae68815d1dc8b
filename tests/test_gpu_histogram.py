"""The device outcome histogram and the sample-based distribution calculator on an
MI355X: counts bit-exact against the oracle's restatement of the reference's counting
loop on the same samples; frequencies of GPU-drawn samples within the TVD bound of the
exact distribution."""
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


def dev_i(T, a):
    return T.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


def _random_samples(rng, S, m, n, lossy):
    """Occupations of n (or fewer, if lossy) particles in m modes, plus a few rows that are
    in no list (too many particles, a negative entry)."""
    rows = []
    for _ in range(S):
        l = int(rng.integers(0, n + 1)) if lossy else n
        rows.append(np.bincount(rng.integers(0, m, l), minlength=m))
    rows = np.array(rows, dtype=np.int32)
    rows[::97, 0] += n + 1
    rows[5::101, -1] = -1
    return rows


@pytest.mark.parametrize("m,n,losses", [(6, 3, False), (6, 3, True), (10, 5, False), (20, 10, False),
                                        (8, 8, True), (1, 4, False)])
def test_histogram_matches_reference_counting(T, oracle, m, n, losses):
    from theboss_amd import ops

    rng = np.random.default_rng(m * 100 + n)
    S = 3000
    smp = _random_samples(rng, S, m, n, losses)
    counts, outside = ops.outcome_histogram(dev_i(T, smp), n, losses=losses)
    outcomes = oracle.generate_possible_states(n, m, losses=losses)
    ref = oracle.sample_frequencies([tuple(r) for r in smp], outcomes, 1) if len(outcomes) < 5000 else None
    counts = counts.cpu().numpy()
    assert len(counts) == len(outcomes)
    in_list = sum(1 for r in smp if (r >= 0).all() and (r.sum() == n or (losses and r.sum() <= n)))
    assert counts.sum() == in_list and outside == S - in_list
    if ref is not None:
        assert np.array_equal(counts, ref.astype(np.int64))
    else:  # large list: check the nonzero bins through the rank of each distinct sample
        from theboss_amd.states import outcome_rank

        import collections

        c = collections.Counter(tuple(r) for r in smp)
        for row, k in c.items():
            r = outcome_rank(row, n, losses)
            if r >= 0:
                assert counts[r] == k


def test_sample_based_calculator_user_outcomes_match_reference(T, oracle):
    """Fixed host samples through a stub strategy: the calculator's frequencies for a
    user outcome list (with a repeated entry and an impossible one) equal the oracle's."""
    from theboss_amd.distribution_calculators import (BSSampleBasedDistributionCalculator,
                                                      BosonSamplingExperimentConfiguration)

    m, n = 5, 3
    rng = np.random.default_rng(11)
    host = [tuple(int(x) for x in np.bincount(rng.integers(0, m, int(rng.integers(1, n + 1))), minlength=m))
            for _ in range(5000)]

    class Fixed:
        def simulate(self, input_state, samples_number=1):
            return host[:samples_number]

    outcomes = oracle.generate_possible_states(n, m, losses=True)[3:40] + [(3, 0, 0, 0, 0), (9, 0, 0, 0, 0)]
    cfg = BosonSamplingExperimentConfiguration(oracle.haar_unitary(m, 1), [1, 1, 1, 0, 0], n, m, 0, n)
    calc = BSSampleBasedDistributionCalculator(cfg, Fixed(), outcomes=outcomes)
    got = calc.calculate_distribution()
    ref = oracle.sample_frequencies(host, outcomes, 5000)
    assert np.allclose(got, ref, rtol=0, atol=0)


def test_sample_based_distribution_of_gpu_sampler_within_tvd_bound(T, oracle):
    """Batched GCC-B samples (device) -> device histogram -> frequencies within the TVD
    bound (p = 1e-4) of the exact distribution of the oracle."""
    from theboss_amd.distribution_calculators import (BSSampleBasedDistributionCalculator,
                                                      BosonSamplingExperimentConfiguration)
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.quantum_computations_utilities import (
        count_total_variation_distance, count_tv_distance_error_bound_of_experiment_results)
    from theboss_amd.simulation_strategies import BatchedGCCBSimulationStrategy

    m, n = 6, 3
    U = oracle.haar_unitary(m, 21)
    init = [1, 1, 1, 0, 0, 0]
    cfg = BosonSamplingExperimentConfiguration(U, init, n, m, 0, n)
    strategy = BatchedGCCBSimulationStrategy(GPURyserPermanentCalculator(U), seed=17)
    calc = BSSampleBasedDistributionCalculator(cfg, strategy)
    freq = calc.calculate_approximate_distribution(200000)
    exact = [oracle.fixed_losses_probability(U, init, o) for o in oracle.generate_possible_states(n, m)]
    tvd = count_total_variation_distance(freq, exact)
    bound = count_tv_distance_error_bound_of_experiment_results(len(exact), 200000, 1e-4)
    assert abs(sum(freq) - 1) < 1e-12
    assert tvd <= bound, (tvd, bound)
