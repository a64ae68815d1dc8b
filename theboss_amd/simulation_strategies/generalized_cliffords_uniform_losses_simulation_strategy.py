"""GCC-A with uniform losses.

Restates theboss's GeneralizedCliffordsUniformLossesSimulationStrategy
(theboss/simulation_strategies/generalized_cliffords_uniform_losses_simulation_strategy.py:19-173):
every input particle survives with probability eta (it is skipped when random() >= eta,
:142-152) and the survivors are drawn from the GCC-A layers. Each layer computed also fills
the lossy output distribution (:154-173): the entry of the output r + e_o is
pmf[o] * l! / prod (r + e_o)_i!, times the binomial weight eta^l C(n, l) (1 - eta)^(n - l)
with l = |r + e_o| (:82-88), and the vacuum entry is the l = 0 weight (:88). Entries follow
generate_possible_states(n, m, losses=True).

Here all samples advance in lock-step: a round draws the survival of one particle per
sample, then the kept samples draw their next output mode from their prefixes' pmfs, which
come from one batched launch per particle number.
"""
from math import comb, factorial
from typing import List, Optional, Sequence

import numpy as np

from ..states import generate_possible_states, outcome_rank
from .generalized_cliffords_simulation_strategy import GeneralizedCliffordsSimulationStrategy


class GeneralizedCliffordsUniformLossesSimulationStrategy(GeneralizedCliffordsSimulationStrategy):
    #: samples per lock-step batch while filling a distribution up to an accuracy
    ACCURACY_BATCH = 256

    def __init__(self, bs_permanent_calculator, transmissivity: float = 0,
                 rng: Optional[np.random.Generator] = None) -> None:
        self._transmissivity = transmissivity
        self.distribution: List[float] = []
        self.unweighted_distribution: List[float] = []
        self._possible_outputs: List[tuple] = []
        self._binomial_weights: List[float] = []
        self.missing_values_in_distribution = False
        super().__init__(bs_permanent_calculator, rng)

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[np.ndarray]:
        """Samples as int64 arrays (:34-51)."""
        self._initialize_simulation(input_state)
        return [np.array(k, dtype=np.int64) for k in self._lock_step(max(int(samples_number), 0))]

    def _initialize_simulation(self, input_state: Sequence[int]) -> None:
        """:53-88: memo reset, the lossy outcome list, the distributions (-1 marks missing
        entries when missing_values_in_distribution is set) and the binomial weights."""
        super()._initialize_simulation(input_state)
        n, m = self.number_of_input_photons, len(self.input_state)
        initializer = -1 if self.missing_values_in_distribution else 0
        self._possible_outputs = generate_possible_states(n, m, losses=True)
        self.distribution = [initializer for _ in self._possible_outputs]
        self.unweighted_distribution = [initializer for _ in self._possible_outputs]
        eta = self._transmissivity
        self._binomial_weights = [pow(eta, left) * float(comb(n, left)) * pow(1 - eta, n - left)
                                  for left in range(n + 1)]
        self.distribution[0] = self._binomial_weights[0]

    def compute_distribution_up_to_accuracy(self, input_state: Sequence[int],
                                            accuracy: float = 1.0) -> List[float]:
        """:90-107: sample until the filled entries sum to the accuracy."""
        self._initialize_simulation(input_state)
        while not np.isclose(max(accuracy - sum(self.distribution), 0), 0):
            self._lock_step(self.ACCURACY_BATCH)
        return self.distribution

    def compute_unweighted_distribution_up_to_accuracy(self, input_state: Sequence[int],
                                                       accuracy: float = 1.0) -> List[float]:
        """:109-128: as above for the per-particle-number distributions, whose entries
        sum to n when complete."""
        self._initialize_simulation(input_state)
        while not np.isclose(max(accuracy - sum(self.unweighted_distribution) / sum(self.input_state), 0), 0):
            self._lock_step(self.ACCURACY_BATCH)
        return self.unweighted_distribution

    def _particles_kept(self, samples: int) -> np.ndarray:
        return self._rng.random(samples) < self._transmissivity

    def _calculate_layers(self, keys) -> None:
        super()._calculate_layers(keys)
        n = self.number_of_input_photons
        for key in keys:
            for output, p in zip(self.possible_outputs[key], self.pmfs[key]):
                out = [int(x) for x in output]
                i = outcome_rank(out, n, losses=True)
                value = float(p) * factorial(sum(out))
                for c in out:
                    value /= factorial(c)
                self.unweighted_distribution[i] = value
                self.distribution[i] = value * self._binomial_weights[sum(out)]
