"""Generalized Clifford & Clifford sampling, version B (GCC-B): the "plain" GPU path.

Restates theboss's GeneralizedCliffordsBSimulationStrategy
(theboss/simulation_strategies/generalized_cliffords_b_simulation_strategy.py:31-110):
per sample, per step a random remaining input particle joins the current input (:102-105),
one submatrices sweep gives the k minors (:73-77, here one GPU launch), the Laplace
expansion over all output modes gives the pmf (:80-92, on the host) and the next output
mode is drawn (:107-110). The submatrices calculator is injectable (the reference
hard-codes it at :73-77).
"""
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ..permanent_calculators.gpu_submatrices_calculators import GPUBSCCRyserSubmatricesPermanentCalculator
from ..states import as_occupation, mode_occupation_to_mode_assignment
from .simulation_strategy_interface import SimulationStrategyInterface


class GeneralizedCliffordsBSimulationStrategy(SimulationStrategyInterface):
    def __init__(self, bs_permanent_calculator, submatrices_calculator_class=None,
                 rng: Optional[np.random.Generator] = None) -> None:
        self._bs_permanent_calculator = bs_permanent_calculator
        self._submatrices_calculator_class = (
            submatrices_calculator_class or GPUBSCCRyserSubmatricesPermanentCalculator)
        self._rng = rng if rng is not None else np.random.default_rng()
        self.input_state: Sequence[int] = []
        self.number_of_input_photons = 0
        self.r_sample: List[int] = []
        self.pmf: np.ndarray = np.zeros(0)
        self._current_input = None
        self._working_input_state = None

    def set_new_matrix(self, new_matrix) -> None:
        self._bs_permanent_calculator.matrix = new_matrix

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        self.input_state = input_state
        self.number_of_input_photons = int(as_occupation(input_state).sum())
        particles = list(mode_occupation_to_mode_assignment(input_state))
        samples = []
        while len(samples) < samples_number:
            self._current_input = np.zeros(len(input_state), dtype=np.int64)
            self._working_input_state = particles.copy()
            self._fill_r_sample(self.number_of_input_photons)
            samples.append(tuple(self.r_sample))
        return samples

    def simulate_with_draws(self, input_state: Sequence[int], orders, uniforms,
                            steps=None) -> np.ndarray:
        """GCC-B with caller-supplied randomness: orders[s][k] is the index (into the
        mode-assignment particle list) of the particle added at step k, uniforms[s][k]
        the inverse-CDF uniform of step k, steps[s] the number of steps (default n)."""
        particles = list(mode_occupation_to_mode_assignment(input_state))
        n = len(particles)
        out = np.zeros((len(orders), len(input_state)), dtype=np.int32)
        for s in range(len(orders)):
            self._current_input = np.zeros(len(input_state), dtype=np.int64)
            self.r_sample = [0] * len(input_state)
            nsteps = n if steps is None else int(steps[s])
            for k in range(nsteps):
                self._current_input[particles[int(orders[s][k])]] += 1
                self._compute_pmf()
                self._pick(float(uniforms[s][k]))
            out[s] = self.r_sample
        return out

    def _fill_r_sample(self, particles_in_sample: int) -> None:
        self.r_sample = [0 for _ in self.input_state]
        while particles_in_sample > sum(self.r_sample):
            self._update_current_input()
            self._compute_pmf()
            self._pick(float(self._rng.random()))

    def _update_current_input(self) -> None:
        idx = int(self._rng.integers(0, len(self._working_input_state)))
        self._current_input[self._working_input_state.pop(idx)] += 1

    def _compute_pmf(self) -> None:
        """:69-92. The Laplace sum runs over the occupied input modes only (the reference's
        other terms are products with s_i = 0)."""
        U = np.asarray(self._bs_permanent_calculator.matrix, dtype=np.complex128)
        calc = self._submatrices_calculator_class(U, self._current_input, self.r_sample)
        as_array = getattr(calc, "minors_array", None)
        minors = as_array() if as_array is not None else \
            np.asarray(calc.compute_permanents(), dtype=np.complex128)
        occ = np.flatnonzero(self._current_input)
        amp = U[:, occ] @ (self._current_input[occ] * minors[occ])
        p = amp.real ** 2 + amp.imag ** 2
        self.pmf = p / p.sum()

    def _pick(self, u: float) -> None:
        """Inverse-CDF draw over the normalised pmf (numpy.random.choice semantics)."""
        cdf = np.cumsum(self.pmf)
        idx = int(np.searchsorted(cdf, u * cdf[-1], side="right"))
        idx = min(idx, len(self.pmf) - 1)
        self.r_sample[idx] += 1
