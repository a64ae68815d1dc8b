"""Host-side occupation-vector helpers (no permanent arithmetic here).

Restates the parts of theboss/boson_sampling_utilities/boson_sampling_utilities.py the
permanent path needs: mode assignment <-> mode occupation (:34-78), the effective
scattering matrix gather (:545-626) and the 2m loss expansion (:287-342).
"""
from typing import List, Sequence, Tuple

import math
import numpy as np


def as_occupation(state) -> np.ndarray:
    """Occupation vector as int64. Accepts lists, tuples, int or float arrays (the
    reference's tests pass float arrays, tests/test_bs_submatrices_permanent_calculators.py:48-50);
    rejects non-integer or negative entries."""
    a = np.asarray(state)
    if a.size == 0:
        return np.zeros(0, dtype=np.int64)
    f = np.asarray(a, dtype=np.float64).reshape(-1)
    r = np.round(f)
    if not np.all(r == f):
        raise ValueError(f"occupation numbers must be integers, got {state!r}")
    if np.any(r < 0):
        raise ValueError(f"occupation numbers must be non-negative, got {state!r}")
    return r.astype(np.int64)


def compact(state: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """(modes, multiplicities) of the occupied modes, in mode order."""
    modes = np.nonzero(state)[0]
    return modes.astype(np.int32), state[modes].astype(np.int32)


def compact_batch(states: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Row-wise compact for a B x m stack: occupied modes first (in mode order), padded
    with mode 0 of multiplicity 0. Returns (modes, multiplicities) as B x W int32, W >= 1."""
    states = np.asarray(states)
    nz = states > 0
    cnt = nz.sum(axis=1)
    W = max(1, int(cnt.max()) if cnt.size else 1)
    rows, cols = np.nonzero(nz)  # row-major: each row's occupied modes in mode order
    slot = np.arange(rows.size) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    modes = np.zeros((states.shape[0], W), dtype=np.int32)
    mult = np.zeros((states.shape[0], W), dtype=np.int32)
    modes[rows, slot] = cols
    mult[rows, slot] = states[rows, cols]
    return modes, mult


def layer_rows(prefixes: np.ndarray) -> np.ndarray:
    """The occupations prefix_p + e_o (o < m) of a GCC-A layer in compact form, built from
    the compacted prefixes: rows[0] modes, rows[1] multiplicities, row p * m + o of shape
    (2, P * m, W + 1) int32 with W the prefixes' compact width. Slots 0..W-1 hold the
    prefix's modes (one more particle where the mode is o), slot W holds o with
    multiplicity 1 when the prefix does not hold it, else 0. Slots are not in mode order."""
    Pf = np.asarray(prefixes, dtype=np.int64)
    P, m = Pf.shape
    if int(Pf.sum(axis=1).max(initial=0)) > 0:
        pm, pu = (np.asarray(a, dtype=np.int64) for a in compact_batch(Pf))
    else:
        pm, pu = np.zeros((P, 1), np.int64), np.zeros((P, 1), np.int64)
    W = pm.shape[1]
    o = np.arange(m)
    hit = (pm[:, None, :] == o[None, :, None]) & (pu[:, None, :] > 0)  # P x m x W
    rows = np.empty((2, P, m, W + 1), dtype=np.int32)
    rows[0, :, :, :W] = pm[:, None, :]
    rows[0, :, :, W] = o[None, :]
    rows[1, :, :, :W] = pu[:, None, :] + hit
    rows[1, :, :, W] = ~hit.any(axis=2)
    return rows.reshape(2, P * m, W + 1)


def effective_matrix(matrix, input_state, output_state) -> np.ndarray:
    """M[a][b] = U[out_mode(a)][in_mode(b)], rows repeated t_j times and columns s_i
    times (boson_sampling_utilities.py:595-626). Empty (0 x 0) when either side has no
    particles (:606-607)."""
    U = np.asarray(matrix, dtype=np.complex128)
    s, t = as_occupation(input_state), as_occupation(output_state)
    if s.sum() == 0 or t.sum() == 0:
        return np.zeros((0, 0), dtype=np.complex128)
    rows = np.repeat(np.arange(len(t)), t)
    cols = np.repeat(np.arange(len(s)), s)
    return np.ascontiguousarray(U[np.ix_(rows, cols)])


def mode_occupation_to_mode_assignment(mode_occupation: Sequence[int]) -> Tuple[int, ...]:
    """boson_sampling_utilities.py:60-78: occupation -> sorted particle modes."""
    occ = as_occupation(mode_occupation)
    return tuple(int(x) for x in np.repeat(np.arange(len(occ)), occ))


def mode_assignment_to_mode_occupation(modes_assignment: Sequence[int],
                                       observed_modes_number: int = 0) -> Tuple[int, ...]:
    """boson_sampling_utilities.py:34-57: particle modes -> occupation vector."""
    if observed_modes_number == 0:
        observed_modes_number = max(modes_assignment) + 1
    occ = np.zeros(observed_modes_number, dtype=np.int64)
    for mode in modes_assignment:
        occ[mode] += 1
    return tuple(int(x) for x in occ)


def prepare_interferometer_matrix_in_expanded_space(interferometer_matrix) -> np.ndarray:
    """Embed a (possibly lossy) m x m interferometer into a 2m x 2m unitary whose extra
    modes absorb lost particles (boson_sampling_utilities.py:287-342): with the SVD
    U = V diag(sigma) W, the expanded matrix is
        [V 0; 0 I] [diag(sigma) diag(sqrt(1-sigma^2)); diag(sqrt(1-sigma^2)) diag(sigma)] [W 0; 0 I].
    """
    U = np.asarray(interferometer_matrix, dtype=np.complex128)
    v, sigma, w = np.linalg.svd(U)
    m = U.shape[0]
    z = np.zeros_like(v)
    eye = np.eye(m)
    loss = np.sqrt(np.clip(1.0 - sigma ** 2, 0.0, None))
    ev = np.block([[v, z], [z, eye]])
    ew = np.block([[w, z], [z, eye]])
    es = np.block([[np.diag(sigma), np.diag(loss)], [np.diag(loss), np.diag(sigma)]])
    return ev @ es @ ew


def generate_possible_states(particles_number: int, modes_number: int,
                             losses: bool = False) -> List[Tuple[int, ...]]:
    """All m-mode states with n particles (or <= n with losses), each particle-number
    block in reverse lexicographic order (boson_sampling_utilities.py:81-152)."""
    if particles_number < 0 or modes_number < 1:
        return []
    if particles_number == 0:
        return [tuple([0] * modes_number)]
    start = 0 if losses else particles_number
    states: List[Tuple[int, ...]] = []
    for n in range(start, particles_number + 1):
        block: List[Tuple[int, ...]] = []

        def rec(prefix, left, modes_left):
            if modes_left == 1:
                block.append(tuple(prefix + [left]))
                return
            for c in range(left, -1, -1):
                rec(prefix + [c], left - c, modes_left - 1)

        rec([], n, modes_number)
        states.extend(sorted(block, reverse=True))
    return states


def outcome_rank(state: Sequence[int], particles_number: int, losses: bool = False) -> int:
    """Index of `state` in generate_possible_states(particles_number, m, losses), or -1 if
    it is not in that list. The same rank as the device histogram: within the block of l
    particles (reverse lexicographic), the states before s are those that agree on modes
    < i and hold more particles in mode i: C(r - s_i - 1 + m - i - 1, m - i - 1) of them
    for r particles left; block l starts at C(l - 1 + m, m) when losses are listed."""
    s = [int(x) for x in state]
    m, l = len(s), sum(s)
    if m < 1 or any(x < 0 for x in s) or l > particles_number or (not losses and l != particles_number):
        return -1
    rank = math.comb(l - 1 + m, m) if (losses and l > 0) else 0
    rem = l
    for i in range(m - 1):
        if rem - s[i] - 1 >= 0:
            rank += math.comb(rem - s[i] - 1 + m - i - 1, m - i - 1)
        rem -= s[i]
    return rank


def generate_lossy_n_particle_input_states(initial_state: Sequence[int],
                                           number_of_particles_left: int) -> List[Tuple[int, ...]]:
    """Distinct inputs left after losing particles, in the first-seen order of
    itertools.combinations over the particle list (boson_sampling_utilities.py:155-198)."""
    import itertools

    occ = as_occupation(initial_state)
    m = len(occ)
    if occ.sum() == 0:
        return [tuple(int(x) for x in occ)]
    particles = list(np.repeat(np.arange(m), occ))
    seen, out = set(), []
    for comb_ in itertools.combinations(range(len(particles)), number_of_particles_left):
        state = mode_assignment_to_mode_occupation([particles[e] for e in comb_], m)
        if state not in seen:
            seen.add(state)
            out.append(state)
    return out
