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
        outcomes = [tuple(int(x) for x in as_occupation(o)) for o in outcomes]
        return list(_fixed_losses_probabilities(self.configuration, self._permanent_calculator,
                                                outcomes))


def _fixed_losses_probabilities(cfg, calc, outcomes) -> np.ndarray:
    """One batched permanent launch for all (outcome, lossy input) pairs."""
    initial = [int(x) for x in as_occupation(cfg.initial_state)]
    n0 = sum(initial)
    calc.matrix = cfg.interferometer_matrix
    groups = {}
    for idx, t in enumerate(outcomes):
        groups.setdefault(sum(t), []).append(idx)
    result = np.zeros(len(outcomes))
    for left, idxs in groups.items():
        lossy = generate_lossy_n_particle_input_states(initial, left)
        mult = np.array([math.prod(math.comb(initial[i], initial[i] - s[i]) for i in range(len(s)))
                         for s in lossy], dtype=np.float64)
        norm = np.array([math.prod(math.factorial(c) for c in s) for s in lossy], dtype=np.float64)
        ins = [s for _ in idxs for s in lossy]
        outs = [outcomes[i] for i in idxs for _ in lossy]
        perms = calc.compute_permanent_batch(ins, outs).reshape(len(idxs), len(lossy))
        sub = (perms.real ** 2 + perms.imag ** 2) / norm[None, :] * mult[None, :]
        p = sub.sum(axis=1)
        p /= math.factorial(sum(lossy[-1]))
        p /= math.comb(n0, left)
        for row, i in enumerate(idxs):
            t = outcomes[i]
            result[i] = p[row] * math.factorial(left) / math.prod(math.factorial(c) for c in t)
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
        outcomes = [tuple(int(x) for x in as_occupation(o)) for o in outcomes]
        probs = np.zeros(len(outcomes))
        nonempty = [i for i, t in enumerate(outcomes) if sum(t) > 0]
        for i, t in enumerate(outcomes):
            if sum(t) == 0:
                probs[i] = self.weights[0]
        if nonempty:
            fixed = _fixed_losses_probabilities(self.configuration, self._permanent_calculator,
                                                [outcomes[i] for i in nonempty])
            for j, i in enumerate(nonempty):
                probs[i] = fixed[j] * self.weights[sum(outcomes[i])]
        return list(probs)
