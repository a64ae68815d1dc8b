"""The reference's GCC strategy suite (tests/test_gcc_strategies.py:21-70 on
tests/gcc_based_strategies_tests_base.py:28-110 and
tests/simulation_strategies_tests_common.py:81-146, 210-237), run on an MI355X through the
strategy factory: the 5-mode permutation matrix, standard [1,1,1,1,0] and binned
[2,1,1,0,0] inputs, lossless and uniformly lossy (eta = 0.5), sample frequencies within the
TVD bound (p = 1e-4) of the exact distribution. The reference draws 1000 samples; these
tests draw 4000 (a tighter bound)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SAMPLES = 4000
P_ERR = 1e-4
PERMUTATION = np.array([[0, 0, 1, 0, 0], [1, 0, 0, 0, 0], [0, 0, 0, 1, 0], [0, 0, 0, 0, 1],
                        [0, 1, 0, 0, 0]], dtype=complex)
STANDARD, BINNED = [1, 1, 1, 1, 0], [2, 1, 1, 0, 0]


@pytest.fixture(scope="module")
def T():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from theboss_amd import _lib

    _lib.load()
    return torch


def _config(state, lossy):
    from theboss_amd.distribution_calculators import BosonSamplingExperimentConfiguration

    n = sum(state)
    lost = 2 if lossy else 0
    return BosonSamplingExperimentConfiguration(
        interferometer_matrix=PERMUTATION, initial_state=list(state), initial_number_of_particles=n,
        number_of_modes=len(state), number_of_particles_lost=lost, number_of_particles_left=n - lost,
        uniform_transmissivity=0.5 if lossy else 1)


def _factory(cfg, strategy):
    from theboss_amd.permanent_calculators import BSPermanentCalculatorFactory
    from theboss_amd.simulation_strategies import SimulationStrategyFactory

    calc = BSPermanentCalculatorFactory(PERMUTATION, None, None).generate_calculator()
    return SimulationStrategyFactory(cfg, calc, strategy), calc


def _within_bound(exact_calc, strategy_obj, cfg, outcomes):
    from theboss_amd.distribution_calculators import BSSampleBasedDistributionCalculator
    from theboss_amd.quantum_computations_utilities import (
        count_total_variation_distance, count_tv_distance_error_bound_of_experiment_results)

    exact = exact_calc.calculate_probabilities_of_outcomes(outcomes)
    freq = BSSampleBasedDistributionCalculator(cfg, strategy_obj, outcomes=outcomes) \
        .calculate_approximate_distribution(SAMPLES)
    tvd = count_total_variation_distance(exact, freq)
    bound = count_tv_distance_error_bound_of_experiment_results(len(outcomes), SAMPLES, P_ERR)
    assert abs(sum(exact) - 1) < 1e-9
    assert tvd <= bound, (tvd, bound)


@pytest.mark.parametrize("state", [STANDARD, BINNED], ids=["standard", "binned"])
@pytest.mark.parametrize("stype", ["GCC", "LOSSY_NET_GCC", "UNIFORM_LOSSES_GCC"])
def test_lossless_sampling_accuracy(T, state, stype):
    """gcc_based_strategies_tests_base.py:28-57: exact calculator with fixed (no) losses."""
    from theboss_amd.distribution_calculators import BSDistributionCalculatorWithFixedLosses
    from theboss_amd.simulation_strategies import StrategyType
    from theboss_amd.states import generate_possible_states

    cfg = _config(state, lossy=False)
    factory, calc = _factory(cfg, StrategyType[stype])
    exact = BSDistributionCalculatorWithFixedLosses(cfg, calc)
    outcomes = generate_possible_states(sum(state), len(state))
    _within_bound(exact, factory.generate_strategy(), cfg, outcomes)


@pytest.mark.parametrize("state", [STANDARD, BINNED], ids=["standard", "binned"])
@pytest.mark.parametrize("stype", ["LOSSY_NET_GCC", "UNIFORM_LOSSES_GCC"])
def test_uniform_losses_sampling_accuracy(T, state, stype):
    """gcc_based_strategies_tests_base.py:59-110: the uniform-loss exact calculator; the
    lossy-network strategy gets the matrix scaled in place by sqrt(eta) (:89-92)."""
    from theboss_amd.distribution_calculators import BSDistributionCalculatorWithUniformLosses
    from theboss_amd.simulation_strategies import StrategyType

    cfg = _config(state, lossy=True)
    factory, calc = _factory(cfg, StrategyType[stype])
    exact = BSDistributionCalculatorWithUniformLosses(cfg, calc)
    if stype == "LOSSY_NET_GCC":
        factory.bs_permanent_calculator.matrix *= pow(0.5, 0.5)
    _within_bound(exact, factory.generate_strategy(), cfg, exact.get_outcomes_in_proper_order())


def test_haar_average_probability(T, oracle):
    """simulation_strategies_tests_common.py:319-368: averaged over Haar matrices, every
    outcome's frequency approaches 1 / (number of outcomes) (the Haar average of the output
    state is maximally mixed on the symmetric subspace). 100 matrices x 1000 samples: the
    spread of the average is about 0.1 / dim, the tolerance 0.6 / dim."""
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import GeneralizedCliffordsSimulationStrategy
    from theboss_amd.states import generate_possible_states

    state = [1, 1, 1, 0, 0]
    outcomes = generate_possible_states(sum(state), len(state))
    index = {o: i for i, o in enumerate(outcomes)}
    freq = np.zeros(len(outcomes))
    for seed in range(100):
        strat = GeneralizedCliffordsSimulationStrategy(
            GPURyserPermanentCalculator(oracle.haar_unitary(5, 200 + seed)), rng=np.random.default_rng(seed))
        for s in strat.simulate(state, 1000):
            freq[index[tuple(s)]] += 1
    freq /= freq.sum()
    dim = len(outcomes)
    assert np.abs(freq - 1 / dim).max() < 0.6 / dim
