"""Batched GCC-B sampling on the device (tb_gccb_sample).

Same distribution as GeneralizedCliffordsBSimulationStrategy
(theboss/simulation_strategies/generalized_cliffords_b_simulation_strategy.py:41-110) and,
with transmissivity < 1, as GeneralizedCliffordsBUniformLossesSimulationStrategy
(generalized_cliffords_b_uniform_losses_simulation_strategy.py:50-121); all samples'
sweeps, Laplace expansions and draws run in one replayed hipGraph. Randomness comes
from a counter-based Philox stream keyed by `seed` (or from forced draws).
"""
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .. import _device as D
from .. import ops
from ..states import as_occupation, mode_occupation_to_mode_assignment
from .simulation_strategy_interface import SimulationStrategyInterface


class BatchedGCCBSimulationStrategy(SimulationStrategyInterface):
    def __init__(self, bs_permanent_calculator, transmissivity: float = 1.0, seed: int = 0,
                 chunk_bits: Optional[Sequence[int]] = None) -> None:
        self._bs_permanent_calculator = bs_permanent_calculator
        self._transmissivity = float(transmissivity)
        self._seed = int(seed)
        self._calls = 0
        self.chunk_bits = chunk_bits
        self._cache = {}

    def set_new_matrix(self, new_matrix) -> None:
        self._bs_permanent_calculator.matrix = new_matrix

    def _device_inputs(self, input_state):
        U = np.asarray(self._bs_permanent_calculator.matrix, dtype=np.complex128)
        particles = np.asarray(mode_occupation_to_mode_assignment(input_state), dtype=np.int32)
        dev = D.device()
        return D.to_device_c128(U, dev), D.to_device_i32(particles, dev)

    def sample_tensor(self, input_state, samples_number: int, seed: Optional[int] = None,
                      forced_order=None, forced_uniforms=None, forced_steps=None, trace=False,
                      out=None):
        """samples_number x m int32 device tensor of output occupations (and the minors
        trace if requested)."""
        U, particles = self._device_inputs(input_state)
        if seed is None:
            seed = (self._seed * 0x9E3779B97F4A7C15 + self._calls) & (2 ** 64 - 1)
            self._calls += 1
        return ops.gccb_sample(U, particles, samples_number, seed, self._transmissivity,
                               forced_order, forced_uniforms, forced_steps, self.chunk_bits,
                               trace, out)

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        out, _ = self.sample_tensor(input_state, samples_number)
        return list(map(tuple, out.cpu().tolist()))
