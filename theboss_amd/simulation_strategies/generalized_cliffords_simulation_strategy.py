"""Generalized Clifford & Clifford sampling, version A (GCC-A), with layer batching.

Restates theboss's GeneralizedCliffordsSimulationStrategy
(theboss/simulation_strategies/generalized_cliffords_simulation_strategy.py:20-266):
the output is built one particle at a time; layer k needs, for every candidate output
r + e_o (o < m) and every k-particle substate s of the input, the probability
|perm(U[r + e_o, s])|^2 / (prod s! k!) weighted by k!(n-k)!/n! prod C(n_i, k_i) (:136-247).
Layers are memoised by the r-prefix (:28-30, 132-133, 171). The reference draws the
samples one after the other and computes each permanent with a separate calculator call;
here the samples advance together, and all permanents of the layers a round needs
(m x C(n, k) per new prefix) go to the GPU in one batched launch.
"""
import itertools
from collections import defaultdict
from math import comb, factorial
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .. import _device as D
from ..states import as_occupation
from .simulation_strategy_interface import SimulationStrategyInterface


def input_substates(state: Sequence[int]) -> List[Tuple[int, ...]]:
    """All substates r <= s in the reference's order
    (generalized_cliffords_simulation_strategy.py:94-121: first mode outermost, which is
    the order of a Cartesian product over the modes)."""
    return list(itertools.product(*[range(int(c) + 1) for c in state]))


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
        self._device_weights = {}

    def set_new_matrix(self, new_matrix) -> None:
        self._bs_permanent_calculator.matrix = new_matrix

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        """All samples advance together, one particle per round: the layers a round needs
        (one per distinct r-prefix not yet memoised) go to the GPU in one launch per
        particle number. Each sample then draws from its prefix's pmf as the reference
        does (:249-266)."""
        self._initialize_simulation(input_state)
        return self._lock_step(max(int(samples_number), 0))

    def _initialize_simulation(self, input_state: Sequence[int]) -> None:
        """generalized_cliffords_simulation_strategy.py:37-60 (state, substates, memo)."""
        self.input_state = [int(x) for x in as_occupation(input_state)]
        self.number_of_input_photons = sum(self.input_state)
        self._labeled_states = defaultdict(list)
        for s in input_substates(self.input_state):
            self._labeled_states[sum(s)].append(s)
        self._layer_cache = {}
        self._device_weights = {}
        self.pmfs = {}

    def _particles_kept(self, samples: int) -> Optional[np.ndarray]:
        """Which samples take a particle in this round (None: all of them)."""
        return None

    def _draw(self, pmfs: np.ndarray) -> np.ndarray:
        """The reference's draw for a stack of pmfs (:249-266): the first index whose
        running sum (sequential, as cumsum) exceeds random() * sum(pmf)."""
        cum = np.cumsum(pmfs, axis=1)
        value = self._rng.random(len(pmfs)) * cum[:, -1]
        return np.minimum((cum <= value[:, None]).sum(axis=1), pmfs.shape[1] - 1)

    def _compute_missing(self, keys, counts) -> None:
        """Memoise the layers of the keys not seen yet, one launch per particle count
        (counts[i] = sum(keys[i])); each new key once, in sorted order."""
        by_count = defaultdict(set)
        for k, c in zip(keys, counts):
            if k not in self.pmfs:
                by_count[c].add(k)
        for _, group in sorted(by_count.items()):
            self._calculate_layers(sorted(group))

    def _lock_step(self, S: int) -> List[Tuple[int, ...]]:
        m = len(self.input_state)
        current = np.zeros((S, m), dtype=np.int64)
        keys = [tuple([0] * m)] * S
        for _ in range(self.number_of_input_photons if S else 0):
            kept = self._particles_kept(S)
            active = np.arange(S) if kept is None else np.nonzero(kept)[0]
            if active.size == 0:
                continue
            act_keys = [keys[i] for i in active]
            self._compute_missing(act_keys, current[active].sum(axis=1).tolist())
            idx = self._draw(np.stack([self.pmfs[k] for k in act_keys]))
            current[active, idx] += 1
            keys = list(map(tuple, current.tolist()))
        if S:
            self.r_sample = list(keys[-1])
            self.current_key = keys[-1]
        return keys

    def _fill_r_sample(self) -> None:
        self.r_sample = [0 for _ in self.input_state]
        self.current_key = tuple(self.r_sample)
        self.current_sample_probability = 1.0
        while self.number_of_input_photons > sum(self.r_sample):
            if self.current_key not in self.pmfs:
                self._calculate_new_layer_of_pmfs()
            self._sample_from_latest_pmf()

    def _weights(self, substates) -> np.ndarray:
        """generalized_cliffords_simulation_strategy.py:173-205, normalised (:155): for
        each substate s with remainder kv = input - s of l particles,
        l! (n - l)! / n! prod_i C(input_i, kv_i), the binomials multiplied in mode order as
        the reference does, a column of substates at a time."""
        n = self.number_of_input_photons
        inp = np.asarray(self.input_state, dtype=np.int64)
        S = np.asarray(substates, dtype=np.int64).reshape(len(substates), len(inp))
        kv = inp[None, :] - S
        base = np.array([factorial(l) * factorial(n - l) / factorial(n) for l in range(n + 1)])
        w = base[kv.sum(axis=1)]
        top = int(inp.max()) if inp.size else 0
        binom = np.array([[float(comb(a, b)) for b in range(top + 1)] for a in range(top + 1)])
        for i in range(len(inp)):
            w = w * binom[inp[i], kv[:, i]]
        return w / w.sum()

    def _norms(self, S: np.ndarray, k: int) -> np.ndarray:
        """prod_i s_i! k! per substate row of S (k particles each), as the float of the
        exact integer product times k! (the product is at most k!, exact in int64 for
        k <= 20; Python integers above)."""
        if k <= 20:
            fact = np.array([factorial(c) for c in range(k + 1)], dtype=np.int64)
            prods = np.prod(fact[S], axis=1).astype(np.float64)
        else:
            prods = np.array([float(np.prod([factorial(c) for c in row], dtype=object))
                              for row in S.tolist()])
        return prods * factorial(k)

    def _layer_constants(self, k: int):
        """Per particle number k: the substates as an array, their normalised weights and
        the prod s! k! normalisers (independent of the r-prefix, so computed once)."""
        if k not in self._layer_cache:
            substates = self._labeled_states[k]
            S = np.asarray(substates, dtype=np.int64).reshape(len(substates), -1)
            self._layer_cache[k] = (S, self._weights(substates), self._norms(S, k))
        return self._layer_cache[k]

    def _calculate_new_layer_of_pmfs(self) -> None:
        self._calculate_layers([self.current_key])

    def _calculate_layers(self, keys: List[Tuple[int, ...]]) -> None:
        """Memoise the pmfs of the given r-prefixes (all with the same particle count)."""
        outputs, pmfs = self._layer_pmfs(keys)
        for key, outs_k, pmf in zip(keys, outputs, pmfs):
            self.possible_outputs[key] = outs_k     # row o: the prefix plus one particle in o
            self.pmfs[key] = pmf

    def _layer_pmfs(self, keys: List[Tuple[int, ...]]):
        """(candidate outputs P x m x m, pmfs P x m) of the given r-prefixes (all with the
        same particle count) from one batched launch of m * |substates| permanents per
        prefix (:136-171)."""
        k = sum(keys[0]) + 1
        S, weights, norm = self._layer_constants(k)
        m = len(keys[0])
        eye = np.eye(m, dtype=np.int64)
        outputs = np.asarray(keys, dtype=np.int64)[:, None, :] + eye[None, :, :]   # P x m x m
        P, ns = len(keys), S.shape[0]
        calc = self._bs_permanent_calculator
        if hasattr(calc, "compute_permanent_layer"):
            perms = calc.compute_permanent_layer(S, np.asarray(keys, dtype=np.int64))
        elif hasattr(calc, "compute_permanent_outer"):
            perms = calc.compute_permanent_outer(S, outputs.reshape(P * m, m))
        else:
            perms = None
        if perms is not None:   # device tensor (P * m) x |S|: pmfs formed on the device
            T = D.torch()
            wn = self._device_weights.get((k, perms.device))
            if wn is None:   # weight over normaliser per substate, uploaded once per k
                wn = T.from_numpy(weights / norm).to(perms.device)
                self._device_weights[(k, perms.device)] = wn
            pmfs = (T.view_as_real(perms).square().sum(-1) @ wn).reshape(P, m).cpu().numpy()
        else:
            ins = np.tile(S, (P * m, 1))               # item = (prefix * m + o) * |S| + i
            outs = np.repeat(outputs.reshape(P * m, m), ns, axis=0)
            perms = calc.compute_permanent_batch_arrays(ins, outs)
            probs = (perms.real ** 2 + perms.imag ** 2).reshape(P, m, ns)
            pmfs = (probs / norm[None, None, :]) @ weights
        return outputs, pmfs

    def _sample_from_latest_pmf(self) -> None:
        """generalized_cliffords_simulation_strategy.py:249-266 (unnormalised pmf)."""
        pmf = self.pmfs[self.current_key]
        value = self._rng.random() * pmf.sum()
        idx = int(np.searchsorted(np.cumsum(pmf), value, side="right"))
        idx = min(idx, len(pmf) - 1)
        self.current_sample_probability = float(pmf[idx])
        self.r_sample = [int(x) for x in self.possible_outputs[self.current_key][idx]]
        self.current_key = tuple(self.r_sample)
