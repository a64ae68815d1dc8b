"""Exact lossy boson-sampling distributions with every permanent of a distribution
computed in one batched GPU launch.

Restates theboss's distribution calculators:
* BosonSamplingExperimentConfiguration (bs_distribution_calculator_interface.py:11-26);
* BSDistributionCalculatorWithFixedLosses (bs_distribution_calculator_with_fixed_losses.py:22-167):
    p(t) = l!/prod t! * sum_{lossy inputs s'} mult(s') |perm(U[t, s'])|^2 / prod s'! / l! / C(n, l)
  with mult(s') = prod_i C(s_i, s_i - s'_i) and l = sum(t);
* BSDistributionCalculatorWithUniformLosses (bs_exact_distribution_with_uniform_losses.py:19-105):
    p(t) = eta^l C(n, l) (1 - eta)^(n - l) * p_fixed(t), p(0) = (1 - eta)^n.
The reference maps outcomes over a multiprocessing.Pool (:62-69); here the (outcome,
lossy input) pairs of a whole call go to the calculator's compute_permanent_batch.
"""
import math
from copy import deepcopy
from dataclasses import dataclass
from typing import Iterable, List, Sequence

import numpy as np

from .states import as_occupation, generate_lossy_n_particle_input_states, generate_possible_states


@dataclass
class BosonSamplingExperimentConfiguration:
    interferometer_matrix: Sequence[Sequence[complex]]
    initial_state: Sequence[int]
    initial_number_of_particles: int
    number_of_modes: int
    number_of_particles_lost: int
    number_of_particles_left: int
    uniform_transmissivity: float = 1
    network_simulation_strategy: object = None
    hierarchy_level: int = 0


class BSDistributionCalculatorWithFixedLosses:
    def __init__(self, configuration: BosonSamplingExperimentConfiguration,
                 permanent_calculator) -> None:
        self.configuration = deepcopy(configuration)
        self._permanent_calculator = permanent_calculator

    def get_outcomes_in_proper_order(self) -> List[tuple]:
        return generate_possible_states(self.configuration.number_of_particles_left,
                                        self.configuration.number_of_modes)

    def calculate_distribution(self) -> List[float]:
        return self.calculate_probabilities_of_outcomes(self.get_outcomes_in_proper_order())

    def calculate_probabilities_of_outcomes(self, outcomes: Iterable[Iterable[int]]) -> List[float]:
        return list(_fixed_losses_probabilities(self.configuration, self._permanent_calculator,
                                                list(outcomes)))


def _as_outcome_array(outcomes, m: int) -> np.ndarray:
    """Outcomes as a B x m int64 array (the per-vector checks of as_occupation, at once)."""
    a = np.asarray(outcomes, dtype=np.float64).reshape(-1, m) if len(outcomes) else \
        np.zeros((0, m))
    r = np.round(a)
    if not np.all(r == a):
        raise ValueError("occupation numbers must be integers")
    if np.any(r < 0):
        raise ValueError("occupation numbers must be non-negative")
    return r.astype(np.int64)


def _fixed_losses_probabilities(cfg, calc, outcomes) -> np.ndarray:
    """bs_distribution_calculator_with_fixed_losses.py:108-151 for many outcomes: per
    particle number, every (outcome, lossy input) permanent in one launch (all pairs,
    gathered on the device, when the calculator offers it), then the reference's weights
    as array arithmetic."""
    initial = [int(x) for x in as_occupation(cfg.initial_state)]
    n0, m = sum(initial), len(initial)
    calc.matrix = cfg.interferometer_matrix
    outs = _as_outcome_array(outcomes, m)
    result = np.zeros(len(outs))
    if len(outs) == 0:
        return result
    fact = np.array([float(math.factorial(i)) for i in range(max(n0, int(outs.max())) + 1)])
    init = np.asarray(initial, dtype=np.int64)
    left_of = outs.sum(axis=1)
    for left in np.unique(left_of):
        left = int(left)
        idxs = np.nonzero(left_of == left)[0]
        lossy = np.asarray(generate_lossy_n_particle_input_states(initial, left),
                           dtype=np.int64).reshape(-1, m)
        mult = np.array([math.prod(math.comb(int(a), int(a - b)) for a, b in zip(init, s))
                         for s in lossy], dtype=np.float64)
        norm = np.prod(fact[lossy], axis=1)
        if left == 0:
            p = np.full(len(idxs), float(np.sum(mult / norm)))
        else:
            perms = calc.compute_permanent_outer(lossy, outs[idxs]) \
                if hasattr(calc, "compute_permanent_outer") else None
            if perms is not None:  # (outcomes x lossy inputs) device tensor
                sq = (perms.real ** 2 + perms.imag ** 2).cpu().numpy()
            else:
                ins = [tuple(s) for _ in idxs for s in lossy]
                pairs_out = [tuple(outs[i]) for i in idxs for _ in lossy]
                pr = calc.compute_permanent_batch(ins, pairs_out).reshape(len(idxs), len(lossy))
                sq = pr.real ** 2 + pr.imag ** 2
            p = (sq / norm[None, :] * mult[None, :]).sum(axis=1)
        p /= math.factorial(int(lossy[-1].sum()))
        p /= math.comb(n0, left)
        result[idxs] = p * math.factorial(left) / np.prod(fact[outs[idxs]], axis=1)
    return result


class BSDistributionCalculatorWithUniformLosses(BSDistributionCalculatorWithFixedLosses):
    def __init__(self, configuration: BosonSamplingExperimentConfiguration,
                 permanent_calculator) -> None:
        super().__init__(configuration, permanent_calculator)
        self.weights = self._initialize_weights()
        self.weightless = False

    def set_weightless(self, weightless: bool) -> None:
        self.weights = self._initialize_weights() if not weightless else [1 for _ in self.weights]
        self.weightless = weightless

    def _initialize_weights(self) -> List[float]:
        n = self.configuration.initial_number_of_particles
        eta = self.configuration.uniform_transmissivity
        return [eta ** l * math.comb(n, l) * (1.0 - eta) ** (n - l) for l in range(n + 1)]

    def get_outcomes_in_proper_order(self) -> List[tuple]:
        return generate_possible_states(self.configuration.initial_number_of_particles,
                                        self.configuration.number_of_modes, losses=True)

    def calculate_probabilities_of_outcomes(self, outcomes: Iterable[Iterable[int]]) -> List[float]:
        """bs_exact_distribution_with_uniform_losses.py:62-97: the empty outcome has weight
        eta^0 C(n, 0) (1 - eta)^n; the others are fixed-loss probabilities weighted by
        their particle number's binomial weight."""
        outs = _as_outcome_array(list(outcomes), self.configuration.number_of_modes)
        probs = np.zeros(len(outs))
        left = outs.sum(axis=1)
        probs[left == 0] = self.weights[0]
        nonempty = np.nonzero(left > 0)[0]
        if len(nonempty):
            fixed = _fixed_losses_probabilities(self.configuration, self._permanent_calculator,
                                                outs[nonempty])
            probs[nonempty] = fixed * np.asarray(self.weights, dtype=np.float64)[left[nonempty]]
        return list(probs)


class BSSampleBasedDistributionCalculator:
    """Outcome frequencies of a sampling strategy
    (bs_sample_based_distribution_calculator.py:18-113).

    The reference scans its outcome list for every sample (:92-106). Here the samples go
    to the device histogram (tb_outcome_histogram), which ranks each sample directly in the
    generate_possible_states order; a user-supplied outcome list is mapped onto those
    ranks, and, as in the reference, only the first occurrence of a repeated outcome
    collects its count. Strategies with a device sampler (sample_tensor) keep the samples
    on the GPU; others are simulated on the host and uploaded.

    As in the reference, calculate_distribution() and calculate_probabilities_of_outcomes()
    draw calculate_approximate_distribution's default of 5000 samples, whatever
    samples_number holds (:68-75, 77).
    """

    def __init__(self, experiment_configuration: BosonSamplingExperimentConfiguration,
                 strategy, samples_number: int = 5000, outcomes=None) -> None:
        self._configuration = experiment_configuration
        self._strategy = strategy
        self._outcomes = outcomes
        self._samples_number = samples_number

    @property
    def configuration(self) -> BosonSamplingExperimentConfiguration:
        return self._configuration

    @configuration.setter
    def configuration(self, new_configuration: BosonSamplingExperimentConfiguration) -> None:
        self._configuration = new_configuration

    @property
    def strategy(self):
        return self._strategy

    @strategy.setter
    def strategy(self, new_strategy) -> None:
        self._strategy = new_strategy

    @property
    def outcomes(self):
        return self._outcomes

    @outcomes.setter
    def outcomes(self, new_outcomes) -> None:
        self._outcomes = new_outcomes

    @property
    def samples_number(self) -> int:
        return self._samples_number

    @samples_number.setter
    def samples_number(self, new_samples_number: int) -> None:
        self._samples_number = new_samples_number

    def calculate_distribution(self) -> List[float]:
        return self.calculate_approximate_distribution()

    def calculate_probabilities_of_outcomes(self, outcomes: Iterable[Iterable[int]]) -> List[float]:
        self._outcomes = outcomes
        return self.calculate_approximate_distribution()

    def get_outcomes_in_proper_order(self):
        return self._outcomes

    def _sample_tensor(self, samples_number: int):
        from . import _device as D

        T = D.torch()
        state = self._configuration.initial_state
        if hasattr(self._strategy, "sample_tensor"):
            out, _ = self._strategy.sample_tensor(state, samples_number)
            return out
        samples = self._strategy.simulate(state, samples_number)
        host = np.asarray([as_occupation(s) for s in samples], dtype=np.int32).reshape(
            len(samples), -1)
        return T.from_numpy(host).to(D.device())

    def calculate_approximate_distribution(self, samples_number: int = 5000) -> List[float]:
        from . import ops
        from .states import outcome_rank

        cfg = self._configuration
        samples = self._sample_tensor(samples_number)
        if self._outcomes is None:
            counts, _ = ops.outcome_histogram(samples, cfg.number_of_particles_left)
            return list((counts.cpu().numpy() / samples_number).astype(np.float64))
        outcomes = [tuple(int(x) for x in as_occupation(o)) for o in self._outcomes]
        top = max([sum(o) for o in outcomes] + [0])
        counts, _ = ops.outcome_histogram(samples, top, losses=True)
        counts = counts.cpu().numpy()
        probs = np.zeros(len(outcomes), dtype=np.float64)
        seen = set()
        for j, o in enumerate(outcomes):
            r = outcome_rank(o, top, losses=True)
            if r >= 0 and o not in seen and len(o) == samples.shape[1]:
                probs[j] = counts[r]
            seen.add(o)
        return list(probs / samples_number)
