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


def _within_bound(exact_calc, strategy_obj, cfg, outcomes, approximation_bound=0.0):
    from theboss_amd.distribution_calculators import BSSampleBasedDistributionCalculator
    from theboss_amd.quantum_computations_utilities import (
        count_total_variation_distance, count_tv_distance_error_bound_of_experiment_results)

    exact = exact_calc.calculate_probabilities_of_outcomes(outcomes)
    freq = BSSampleBasedDistributionCalculator(cfg, strategy_obj, outcomes=outcomes) \
        .calculate_approximate_distribution(SAMPLES)
    tvd = count_total_variation_distance(exact, freq)
    bound = count_tv_distance_error_bound_of_experiment_results(len(outcomes), SAMPLES, P_ERR)
    assert abs(sum(exact) - 1) < 1e-9
    assert tvd <= approximation_bound + bound, (tvd, approximation_bound, bound)


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


@pytest.mark.parametrize("state", [STANDARD, BINNED], ids=["standard", "binned"])
@pytest.mark.parametrize("lossy", [False, True], ids=["lossless", "lossy"])
def test_bobs_exact_level_sampling_accuracy(T, state, lossy):
    """tests/test_bobs_strategy.py:20-54: BOBS with hierarchy_level = m (nothing
    approximated) samples the exact distribution; lossy runs scale the matrix by
    sqrt(eta)."""
    from theboss_amd.distribution_calculators import (BSDistributionCalculatorWithFixedLosses,
                                                      BSDistributionCalculatorWithUniformLosses)
    from theboss_amd.simulation_strategies import StrategyType
    from theboss_amd.states import generate_possible_states

    cfg = _config(state, lossy)
    cfg.hierarchy_level = cfg.number_of_modes
    factory, calc = _factory(cfg, StrategyType.BOBS)
    if lossy:
        exact = BSDistributionCalculatorWithUniformLosses(cfg, calc)
        outcomes = exact.get_outcomes_in_proper_order()
        factory.bs_permanent_calculator.matrix *= pow(0.5, 0.5)
    else:
        exact = BSDistributionCalculatorWithFixedLosses(cfg, calc)
        outcomes = generate_possible_states(sum(state), len(state))
    _within_bound(exact, factory.generate_strategy(), cfg, outcomes)


def test_bobs_approximate_sampling_accuracy(T):
    """tests/test_bobs_strategy.py:56-90: 3 of 5 modes approximated (hierarchy level 2),
    eta = 0.5: TVD within the BOBS bound eta^2 (n - h) / 2 + eta (1 - eta) / 2 plus the
    statistical bound. As in the reference (gcc_based_strategies_tests_base.py:74-86), the
    exact calculator is built while the configuration holds [1, 1, 1, 1, 0] (it keeps a
    deep copy) and the strategy then samples the input [0, 0, 3, 1, 0]."""
    from theboss_amd.distribution_calculators import BSDistributionCalculatorWithUniformLosses
    from theboss_amd.simulation_strategies import StrategyType

    cfg = _config(STANDARD, lossy=True)
    cfg.hierarchy_level = 2
    factory, calc = _factory(cfg, StrategyType.BOBS)
    exact = BSDistributionCalculatorWithUniformLosses(cfg, calc)
    cfg.initial_state = [0, 0, 3, 1, 0]
    factory.bs_permanent_calculator.matrix *= pow(0.5, 0.5)
    eta, n = 0.5, cfg.initial_number_of_particles
    bound = eta ** 2 / 2 * (n - cfg.hierarchy_level) + eta * (1 - eta) / 2
    _within_bound(exact, factory.generate_strategy(), cfg, exact.get_outcomes_in_proper_order(), bound)


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


@pytest.mark.parametrize("state", [[1, 1, 1, 0], [3, 1, 2, 0]], ids=["standard", "binned"])
@pytest.mark.parametrize("kind", ["fixed", "uniform"])
def test_distribution_calculators_sum_to_one(T, oracle, state, kind):
    """tests/test_distribution_calculators.py:56-99: Haar 4-mode matrix, 2 particles lost
    (fixed) or eta = 0.5 (uniform), Ryser: the distribution sums to 1."""
    from theboss_amd.distribution_calculators import (BosonSamplingExperimentConfiguration,
                                                      BSDistributionCalculatorWithFixedLosses,
                                                      BSDistributionCalculatorWithUniformLosses)
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator

    U = oracle.haar_unitary(4, 300)
    cfg = BosonSamplingExperimentConfiguration(
        interferometer_matrix=U, initial_state=state, initial_number_of_particles=sum(state),
        number_of_modes=4, number_of_particles_lost=2, number_of_particles_left=sum(state) - 2,
        uniform_transmissivity=0.5)
    calc = GPURyserPermanentCalculator(U, input_state=state)
    cls = BSDistributionCalculatorWithFixedLosses if kind == "fixed" else BSDistributionCalculatorWithUniformLosses
    assert np.isclose(sum(cls(cfg, calc).calculate_distribution()), 1, rtol=0, atol=1e-12)
