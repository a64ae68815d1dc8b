"""Distances between outcome distributions and sample-count bounds, as used with the
sample-based distribution calculator (theboss/quantum_computations_utilities.py:25-162).
Host-side numpy for lists and dicts; tvd_device for distributions already on the GPU.
"""
import math
from collections import defaultdict
from typing import Dict, Sequence, Tuple

import numpy as np


def count_total_variation_distance(distribution1: Sequence[float],
                                   distribution2: Sequence[float]) -> float:
    """Half the L1 distance of two equally long distributions (:57-78)."""
    assert len(distribution1) == len(distribution2), (
        f"Distributions must have equal lengths! Got: {len(distribution1)}"
        f"and {len(distribution2)}!")
    return float(np.sum(np.abs(np.asarray(distribution1, dtype=np.float64)
                               - np.asarray(distribution2, dtype=np.float64))) / 2)


def count_total_variation_distance_dicts(distribution_1: Dict[Tuple[int, ...], float],
                                         distribution_2: Dict[Tuple[int, ...], float]) -> float:
    """TVD over the union of keys; a missing key has probability 0 (:25-54)."""
    d1, d2 = defaultdict(float, distribution_1), defaultdict(float, distribution_2)
    keys = set(distribution_1) | set(distribution_2)
    return 0.5 * sum(abs(d1[k] - d2[k]) for k in keys)


def count_distance_between_matrices(matrix1: np.ndarray, matrix2: np.ndarray) -> float:
    """Frobenius norm of the difference (:81-91)."""
    return float(np.linalg.norm(np.asarray(matrix1) - np.asarray(matrix2)))


def count_tv_distance_error_bound_of_experiment_results(outcomes_number: int, samples_number: int,
                                                        error_probability: float) -> float:
    """sqrt((-log p + K log 2) / (2N)): with probability >= 1 - p, the TVD between N-sample
    frequencies over K outcomes and the true distribution is below this (:94-128; 2^K - 2
    approximated by 2^K as there)."""
    bound = -math.log(error_probability) + outcomes_number * math.log(2)
    return math.sqrt(bound / (2 * samples_number))


def compute_minimal_number_of_samples_for_desired_accuracy(outcomes_number: int,
                                                           error_probability: float,
                                                           expected_distance: float) -> int:
    """Smallest N for which the bound above reaches expected_distance (:146-155)."""
    n = -math.log(error_probability) + math.log(2) * outcomes_number
    return int(n / (2 * expected_distance ** 2)) + 1


def compute_qft_matrix(n: int) -> np.ndarray:
    """n x n quantum Fourier transform, entries exp(2 pi i jk / n) / sqrt(n) (:158-178)."""
    if n == 0:
        return np.asarray([])
    jk = np.outer(np.arange(n), np.arange(n))
    return np.exp(2j * np.pi * jk / n) / math.sqrt(n)


def tvd_device(p, q) -> float:
    """TVD of two distributions held as GPU tensors of equal length."""
    return float(0.5 * (p.to(q.dtype) - q).abs().sum().item())
