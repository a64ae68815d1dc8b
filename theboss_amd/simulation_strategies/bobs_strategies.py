"""Brod-Oszmaniec (BOBS) approximation strategies with the whole batch on the device.

Both reference strategies draw, for every sample, a lossy input and a fresh
interferometer (random phases and a QFT on the approximated modes), then run one GCC-B
sample on it, spread over a spawn process pool. Here the host prepares all samples'
inputs and interferometers with vectorised numpy, and one call of the device sampler in
per-sample mode (tb_gccb_sample with per_sample = 1) draws every sample.

* LossyStateApproximationSimulationStrategy restates
  theboss/simulation_strategies/lossy_state_approximated_simulation_strategy.py:35-362
  (uniform losses applied to the input state, GCC-B on the unitary);
* NonuniformLossesApproximationStrategy restates
  nonuniform_losses_approximation_strategy.py:65-347 (uniform losses extracted from the
  lossy interferometer by SVD, lossy-network GCC-B on the 2m-mode expansion).

The host randomness (loss draws, column permutations, phases, symmetrisation) comes from a
seeded numpy Generator; the per-sample matrices are formed on the device (batched complex
matrix products); the GCC-B draws come from the device Philox stream.
"""
import math
from typing import List, Sequence, Tuple

import numpy as np

from .. import _device as D
from .. import ops
from ..quantum_computations_utilities import compute_qft_matrix
from ..states import as_occupation, generate_lossy_n_particle_input_states
from .simulation_strategy_interface import SimulationStrategyInterface


def generate_qft_matrix_for_first_m_modes(m: int, all_modes_number: int) -> np.ndarray:
    """QFT on the first m modes, identity on the rest (boson_sampling_utilities.py:505-521)."""
    out = np.eye(all_modes_number, dtype=np.complex128)
    if m > 0:
        out[:m, :m] = compute_qft_matrix(m)
    return out


def _particle_lists(inputs: np.ndarray, n_max: int) -> Tuple[np.ndarray, np.ndarray]:
    """Mode assignment of each input row (sorted modes with multiplicity), padded with mode
    0 to n_max entries, and the particle counts: particle p of a row sits in the first
    mode whose running occupation exceeds p."""
    inputs = np.asarray(inputs, dtype=np.int64)
    S, m = inputs.shape
    counts = inputs.sum(axis=1)
    width = max(n_max, 1)
    # every particle's mode in mode order, then its position within its row
    modes = np.repeat(np.tile(np.arange(m, dtype=np.int32), S), inputs.reshape(-1))
    row = np.repeat(np.arange(S), counts)
    pos = np.arange(len(modes)) - np.repeat(np.cumsum(counts) - counts, counts)
    lists = np.zeros((S, width), dtype=np.int32)
    lists[row, pos] = modes
    return lists, counts.astype(np.int32)


def expand_batch(Us: np.ndarray) -> np.ndarray:
    """2m x 2m unitary dilations of a batch of m x m sub-unitaries, as
    prepare_interferometer_matrix_in_expanded_space (boson_sampling_utilities.py:287-342)
    does for one: [V 0; 0 I] [S L; L S] [W 0; 0 I] with U = V diag(sigma) W, S = diag(sigma),
    L = diag(sqrt(1 - sigma^2))."""
    v, sigma, w = np.linalg.svd(Us)
    S, m = Us.shape[0], Us.shape[1]
    loss = np.sqrt(np.clip(1.0 - sigma ** 2, 0.0, None))
    ev = np.zeros((S, 2 * m, 2 * m), dtype=np.complex128)
    ew = np.zeros_like(ev)
    es = np.zeros_like(ev)
    idx = np.arange(m)
    ev[:, :m, :m] = v
    ev[:, m + idx, m + idx] = 1
    ew[:, :m, :m] = w
    ew[:, m + idx, m + idx] = 1
    es[:, idx, idx] = sigma
    es[:, m + idx, m + idx] = sigma
    es[:, idx, m + idx] = loss
    es[:, m + idx, idx] = loss
    return ev @ es @ ew


def _sample_on_device(Us, inputs: np.ndarray, seed: int) -> np.ndarray:
    """One device call: sample s uses interferometer Us[s] (S x M x M complex128 device
    tensor) and input occupation inputs[s]."""
    T = D.torch()
    S = Us.shape[0]
    n_max = int(inputs.sum(axis=1).max()) if S else 0
    lists, counts = _particle_lists(inputs, n_max)
    dev = Us.device
    p_d = D.to_device_i32(lists, dev)
    steps = T.from_numpy(counts).to(dev)
    out, _ = ops.gccb_sample(Us.contiguous(), p_d, S, seed, forced_steps=steps)
    return out.cpu().numpy()


def _phases(rng, S: int, m: int, k: int) -> np.ndarray:
    """exp(2 pi i u) on the first k modes, 1 elsewhere
    (generate_random_phases_matrix_for_first_m_modes, boson_sampling_utilities.py:524-542)."""
    ph = np.ones((S, m), dtype=np.complex128)
    if k > 0:
        ph[:, :k] = np.exp(2j * np.pi * rng.random((S, k)))
    return ph


class LossyStateApproximationSimulationStrategy(SimulationStrategyInterface):
    """BOBS for uniform losses (lossy_state_approximated_simulation_strategy.py:35-362).

    The first hierarchy_level modes of the input are simulated exactly: their lossy part is
    drawn from all lossy sub-states with weight B(n, l; eta) / C(n, l) (:131-193). The
    particles of the other modes are approximated: l ~ B(n', eta) of them, all placed in the
    first approximated mode (:195-240). Each sample uses U with randomly permuted columns,
    times random phases and a QFT on the first m - hierarchy_level modes (:333-362).
    threads_number is accepted for interface compatibility; the device does the sampling.
    """

    def __init__(self, bs_permanent_calculator, uniform_transmissivity: float,
                 hierarchy_level: int, threads_number: int = -1, seed: int = 0) -> None:
        self._permanent_calculator = bs_permanent_calculator
        self._uniform_transmissivity = float(uniform_transmissivity)
        self._hierarchy_level = int(hierarchy_level)
        self._threads_number = threads_number
        self._rng = np.random.default_rng(seed)
        self._seed = int(seed)
        self._calls = 0
        m = len(np.asarray(bs_permanent_calculator.matrix))
        self._qft_matrix = generate_qft_matrix_for_first_m_modes(m - self._hierarchy_level, m)

    def set_new_matrix(self, new_matrix) -> None:
        self._permanent_calculator.matrix = new_matrix

    def _lossy_weights(self, n: int) -> List[float]:
        eta = self._uniform_transmissivity
        return [math.comb(n, l) * eta ** l * (1 - eta) ** (n - l) for l in range(n + 1)]

    def _input_parts(self, input_state: Sequence[int]):
        h = self._hierarchy_level
        exact = [int(x) for x in input_state[:h]]
        approx = [int(x) for x in input_state[h:]]
        parts = []
        for l in range(sum(exact) + 1):
            parts.extend(generate_lossy_n_particle_input_states(exact, l))
        n = sum(parts[-1]) if parts else 0
        possible = self._lossy_weights(n)
        weights = [possible[sum(p)] / math.comb(n, sum(p)) for p in parts]
        m = len(np.asarray(self._permanent_calculator.matrix))
        if not 0 <= h < m:  # exact simulation: no approximated part
            return parts, weights, [[]], [1.0]
        approx_parts = []
        for l in range(sum(approx) + 1):
            z = [0] * len(approx)
            z[0] = l
            approx_parts.append(z)
        return parts, weights, approx_parts, self._lossy_weights(sum(approx))

    def _choose(self, weights, size):
        w = np.asarray(weights, dtype=np.float64)
        return self._rng.choice(len(w), size=size, p=w / w.sum())

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        if samples_number < 1:
            return []
        input_state = [int(x) for x in as_occupation(input_state)]
        parts, weights, aparts, aweights = self._input_parts(input_state)
        S = samples_number
        pi = self._choose(weights, S)
        ai = self._choose(aweights, S)
        P = np.array(parts, dtype=np.int64).reshape(len(parts), -1)
        A = np.array(aparts, dtype=np.int64).reshape(len(aparts), -1)
        inputs = np.concatenate([P[pi], A[ai]], axis=1)
        T = D.torch()
        U = np.asarray(self._permanent_calculator.matrix, dtype=np.complex128)
        m = U.shape[0]
        k = m - self._hierarchy_level  # modes under phases and QFT
        perms = np.argsort(self._rng.random((S, m)), axis=1)
        phases = _phases(self._rng, S, m, k)
        # On the device: U[:, perm_s] diag(phases_s) QFT for every sample.
        dev = D.device()
        U_d = D.to_device_c128(U, dev)
        cols = U_d[:, T.from_numpy(perms.reshape(-1)).to(dev)].reshape(m, S, m).permute(1, 0, 2)
        scaled = (cols * D.to_device_c128(phases, dev)[:, None, :]).contiguous()
        Us = T.matmul(scaled, D.to_device_c128(self._qft_matrix, dev))
        seed = (self._seed * 0x9E3779B97F4A7C15 + self._calls) & (2 ** 64 - 1)
        self._calls += 1
        out = _sample_on_device(Us, inputs, seed)
        return list(map(tuple, out.tolist()))


class NonuniformLossesApproximationStrategy(SimulationStrategyInterface):
    """BOBS for non-uniform losses (nonuniform_losses_approximation_strategy.py:65-347).

    The largest uniform loss is extracted from the interferometer's singular values
    (:131-153); it is applied to the input mode by mode with binomial draws (:285-314).
    The particles of the first approximated_modes_number modes are gathered into one of
    those modes at random (:276-280), and each sample's interferometer is the loss-reduced
    matrix times random phases and a QFT on the approximated modes (:316-347), simulated
    as a lossy network (2m-mode dilation, outputs trimmed to m modes).

    Unlike the reference, which keeps adding to its approximated-particle count on every
    simulate call (:240-242), each call starts from the given input.
    """

    def __init__(self, bs_permanent_calculator, approximated_modes_number: int,
                 threads_number: int = -1, seed: int = 0) -> None:
        m = len(np.asarray(bs_permanent_calculator.matrix))
        self._approximated_modes_number = min(max(int(approximated_modes_number), 0), m)
        self._permanent_calculator = bs_permanent_calculator
        self._threads_number = threads_number
        self._rng = np.random.default_rng(seed)
        self._seed = int(seed)
        self._calls = 0
        self._extract_losses_from_the_interferometer(bs_permanent_calculator.matrix)

    def _extract_losses_from_the_interferometer(self, interferometer_matrix) -> None:
        u, s, v = np.linalg.svd(np.asarray(interferometer_matrix, dtype=np.complex128))
        transmissivities = s ** 2
        self._uniform_losses = float(np.min(1 - transmissivities))
        eta = transmissivities / (1 - self._uniform_losses)
        self._initial_matrix = u @ np.diag(np.sqrt(eta)) @ v
        # The SVD of initial_matrix @ P is (u, sqrt(eta), v @ P) for any unitary P, so the
        # per-sample dilations need no per-sample SVD: their first m columns are
        # [u diag(sqrt(eta)); diag(sqrt(1 - eta))] @ v @ P.
        self._dilation_left = np.concatenate(
            [u * np.sqrt(eta)[None, :], np.diag(np.sqrt(np.clip(1.0 - eta, 0.0, None)))], axis=0)
        self._right = v

    def _binomial_weights(self, max_particles: int):
        eta = 1 - self._uniform_losses
        return {n: [eta ** l * math.comb(n, l) * (1.0 - eta) ** (n - l) for l in range(n + 1)]
                for n in range(1, max_particles + 1)}

    def simulate(self, input_state: Sequence[int], samples_number: int = 1) -> List[Tuple[int, ...]]:
        if samples_number < 1:
            return []
        state = np.array([int(x) for x in as_occupation(input_state)], dtype=np.int64)
        k = self._approximated_modes_number
        S, m = samples_number, len(state)
        approx_n = int(state[:k].sum())
        base = state.copy()
        base[:k] = 0
        inputs = np.tile(base, (S, 1))
        if k > 0:  # symmetrisation: all approximated particles into one random mode
            inputs[np.arange(S), self._rng.integers(0, k, S)] = approx_n
        if not np.isclose(self._uniform_losses, 0):
            # per mode holding c particles: the first l whose cumulative binomial weight
            # exceeds a uniform draw (the reference's choice(), :31-62)
            weights = self._binomial_weights(int(max(state.max(), approx_n, 1)))
            u = self._rng.random((S, m))
            lossy = np.zeros_like(inputs)
            for c in np.unique(inputs[inputs > 0]):
                cdf = np.cumsum(weights[int(c)])
                hit = inputs == c
                lossy[hit] = np.minimum(np.searchsorted(cdf, u[hit], side="right"), c)
            inputs = lossy
        qft = generate_qft_matrix_for_first_m_modes(k, m)
        phases = _phases(self._rng, S, m, k)
        # On the device: the 2m-mode dilation of initial_matrix diag(phases_s) QFT. Only its
        # first m columns (the input modes) are ever read; the rest stay zero.
        T = D.torch()
        dev = D.device()
        right = (D.to_device_c128(self._right, dev)[None, :, :]
                 * D.to_device_c128(phases, dev)[:, None, :]).contiguous()
        right = T.matmul(right, D.to_device_c128(qft, dev))                    # S x m x m
        left = D.to_device_c128(self._dilation_left, dev)                      # 2m x m
        expanded = T.zeros((S, 2 * m, 2 * m), dtype=T.complex128, device=dev)
        expanded[:, :, :m] = T.matmul(left.expand(S, 2 * m, m).contiguous(), right)
        seed = (self._seed * 0x9E3779B97F4A7C15 + self._calls) & (2 ** 64 - 1)
        self._calls += 1
        padded = np.concatenate([inputs, np.zeros_like(inputs)], axis=1)
        out = _sample_on_device(expanded, padded, seed)
        return list(map(tuple, out[:, :m].tolist()))
