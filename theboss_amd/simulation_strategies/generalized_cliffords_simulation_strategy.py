"""Generalized Clifford & Clifford sampling, version A (GCC-A), with layer batching.

Restates theboss's GeneralizedCliffordsSimulationStrategy
(theboss/simulation_strategies/generalized_cliffords_simulation_strategy.py:20-266):
the output is built one particle at a time; layer k needs, for every candidate output
r + e_o (o < m) and every k-particle substate s of the input, the probability
|perm(U[r + e_o, s])|^2 / (prod s! k!) weighted by k!(n-k)!/n! prod C(n_i, k_i) (:136-247).
Layers are memoised by the r-prefix (:28-30, 132-133, 171). The reference computes each
permanent with a separate calculator call; here all m x C(n, k) permanents of a layer go
to the GPU in one batched launch of the calculator's compute_permanent_batch.
"""
from collections import defaultdict
from math import comb, factorial
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from ..states import as_occupation
from .simulation_strategy_interface import SimulationStrategyInterface


def input_substates(state: Sequence[int]) -> List[Tuple[int, ...]]:
    """All substates r <= s in the reference's order
    (generalized_cliffords_simulation_strategy.py:94-121: first mode outermost)."""
    if len(state) < 1:
        return [tuple()]
    first, rest = int(state[0]), list(state[1:])
    smaller = input_substates(rest)
    return [(i,) + sub for i in range(first + 1) for sub in smaller]


class GeneralizedCliffordsSimulationStrategy(SimulationStrategyInterface):
    def __init__(self, bs_permanent_calculator, rng: Optional[np.random.Generator] = None) -> None:
        self._bs_permanent_calculator = bs_permanent_calculator
        self._rng = rng if rng is not None else np.random.default_rng()
        self.input_state: Sequence[int] = []
        self.number_of_input_photons = 0
        self.r_sample: List[int] = []
        self.pmfs: Dict[Tuple[int, ...], np.ndarray] = {}
        self.possible_outputs: Dict[Tuple[int, ...], List[List[int]]] = {}
        self._labeled_states: Dict[int, List[Tuple[int, ...]]] = defaultdict(list)
        self.current_key: Tuple[int, ...] = tuple()
        self.current_sample_probability = 1.0
        self._layer_cache = {}

    def set_new_matrix(self, new_matrix) -> None:
        self._bs_permanent_calculator.matrix = new_matrix

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        self.input_state = [int(x) for x in as_occupation(input_state)]
        self.number_of_input_photons = sum(self.input_state)
        self._labeled_states = defaultdict(list)
        for s in input_substates(self.input_state):
            self._labeled_states[sum(s)].append(s)
        self._layer_cache = {}
        self.pmfs = {}
        samples = []
        while len(samples) < samples_number:
            self._fill_r_sample()
            samples.append(tuple(self.r_sample))
        return samples

    def _fill_r_sample(self) -> None:
        self.r_sample = [0 for _ in self.input_state]
        self.current_key = tuple(self.r_sample)
        self.current_sample_probability = 1.0
        while self.number_of_input_photons > sum(self.r_sample):
            if self.current_key not in self.pmfs:
                self._calculate_new_layer_of_pmfs()
            self._sample_from_latest_pmf()

    def _weights(self, substates) -> np.ndarray:
        """generalized_cliffords_simulation_strategy.py:173-205, normalised (:155)."""
        n = self.number_of_input_photons
        w = []
        for s in substates:
            kv = [self.input_state[i] - s[i] for i in range(len(s))]
            l = sum(kv)
            val = factorial(l) * factorial(n - l) / factorial(n)
            for i in range(len(self.input_state)):
                val *= comb(self.input_state[i], kv[i])
            w.append(val)
        w = np.asarray(w, dtype=np.float64)
        return w / w.sum()

    def _layer_constants(self, k: int):
        """Per particle number k: the substates as an array, their normalised weights and
        the prod s! k! normalisers (independent of the r-prefix, so computed once)."""
        if k not in self._layer_cache:
            substates = self._labeled_states[k]
            S = np.asarray(substates, dtype=np.int64).reshape(len(substates), -1)
            norm = np.array([float(np.prod([factorial(c) for c in s])) * factorial(sum(s))
                             for s in substates])
            self._layer_cache[k] = (S, self._weights(substates), norm)
        return self._layer_cache[k]

    def _calculate_new_layer_of_pmfs(self) -> None:
        k = sum(self.r_sample) + 1
        S, weights, norm = self._layer_constants(k)
        m = len(self.r_sample)
        outputs = np.tile(np.asarray(self.r_sample, dtype=np.int64), (m, 1)) + np.eye(m, dtype=np.int64)
        self.possible_outputs[self.current_key] = [list(map(int, o)) for o in outputs]
        ins = np.tile(S, (m, 1))                      # output-major: item = o * |S| + i
        outs = np.repeat(outputs, S.shape[0], axis=0)
        perms = self._bs_permanent_calculator.compute_permanent_batch_arrays(ins, outs)
        probs = (perms.real ** 2 + perms.imag ** 2).reshape(m, S.shape[0])
        pmf = (probs / norm[None, :]) @ weights
        self.pmfs[self.current_key] = pmf

    def _sample_from_latest_pmf(self) -> None:
        """generalized_cliffords_simulation_strategy.py:249-266 (unnormalised pmf)."""
        pmf = self.pmfs[self.current_key]
        value = self._rng.random() * pmf.sum()
        idx = int(np.searchsorted(np.cumsum(pmf), value, side="right"))
        idx = min(idx, len(pmf) - 1)
        self.current_sample_probability = float(pmf[idx])
        self.r_sample = list(self.possible_outputs[self.current_key][idx])
        self.current_key = tuple(self.r_sample)
