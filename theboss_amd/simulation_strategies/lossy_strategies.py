"""Lossy GCC-B strategies.

* GeneralizedCliffordsBUniformLossesSimulationStrategy restates
  theboss/simulation_strategies/generalized_cliffords_b_uniform_losses_simulation_strategy.py:24-121:
  the number of surviving photons l ~ Binomial(n, eta) is drawn by CDF threshold
  (:56-85), then l GCC-B steps run. Here the whole batch runs in the device sampler.
* LossyNetworksGeneralizedCliffordsSimulationStrategy restates
  lossy_networks_generalized_cliffords_simulation_strategy.py:24-88: a lossy m x m matrix
  is embedded into a 2m x 2m one (boson_sampling_utilities.py:287-342), GCC-B runs on the
  2m modes with the input padded by m empty modes, and the extra modes are trimmed.
"""
from typing import List, Sequence, Tuple

import numpy as np

from ..states import as_occupation, prepare_interferometer_matrix_in_expanded_space
from .batched_gccb_simulation_strategy import BatchedGCCBSimulationStrategy
from .simulation_strategy_interface import SimulationStrategyInterface


class GeneralizedCliffordsBUniformLossesSimulationStrategy(BatchedGCCBSimulationStrategy):
    def __init__(self, bs_permanent_calculator, transmissivity: float = 1.0, seed: int = 0,
                 chunk_bits=None) -> None:
        super().__init__(bs_permanent_calculator, transmissivity=transmissivity, seed=seed,
                         chunk_bits=chunk_bits)

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[np.ndarray]:
        """Samples as int64 arrays, as the reference returns them (:87-109)."""
        return [np.array(s, dtype=np.int64) for s in super().simulate(input_state, samples_number)]


class LossyNetworksGeneralizedCliffordsSimulationStrategy(SimulationStrategyInterface):
    def __init__(self, bs_permanent_calculator, seed: int = 0) -> None:
        bs_permanent_calculator.matrix = prepare_interferometer_matrix_in_expanded_space(
            bs_permanent_calculator.matrix)
        self._helper_strategy = BatchedGCCBSimulationStrategy(bs_permanent_calculator, seed=seed)

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        s = as_occupation(input_state)
        expanded = np.concatenate([s, np.zeros_like(s)])
        out = self._helper_strategy.simulate(expanded, samples_number)
        return [tuple(o[: len(s)]) for o in out]

    def set_new_matrix(self, matrix) -> None:
        self._helper_strategy.set_new_matrix(prepare_interferometer_matrix_in_expanded_space(matrix))
