"""CPU oracle for theboss's permanent hot path (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker (or the timed CPU baseline). The product package
`theboss_amd` never imports it.

Two layers:
* ctypes bindings to oracle/liboracle.so (permanent_oracle.c), the C restatement of
  the reference's permanent calculators, submatrices calculators and GCC-B step;
* numpy restatements of the reference's small host-side helpers: seeded Haar
  unitaries, state enumeration, the lossy exact distributions used as a known-answer
  test, and the GCC-A layer pmf.

How each routine is pinned (tests/test_oracle.py):
* classic              -- equals an independent sum over permutations (itertools),
                          and the closed forms perm(J_n)=n!, perm(J_n-I_n)=!n,
                          perm(diag d)=prod d, perm(permutation matrix)=1.
* glynn (modes 0, 1)   -- equals classic for n<=8 on seeded Haar blocks; closed forms
                          at n up to 24 (rank-one u v^T gives n! prod u prod v,
                          J_n gives n!, derangements); perm(cA)=c^n perm(A); block
                          diagonal factorises; transpose and row/column permutation
                          invariance.
* glynn_py, ryser_py, cc_submatrices_py (timing restatements in pure Python)
                       -- equal the C glynn / ryser / cc submatrices on random patterns.
* ryser / chin-huh     -- equal classic on the nine occupation patterns of the
                          reference's tests/test_bs_permanent_calculators.py:88-122
                          and on random multiplicity patterns.
* cc submatrices       -- (Ryser restatement, the reference's algorithm) every entry
                          equals the classic permanent of U[t, s-e_i]
                          (the reference compares absolute values only,
                          tests/test_bs_submatrices_permanent_calculators.py:75-109;
                          the check here compares complex values), plus the Laplace
                          identity sum_i s_i U[m][i] minor_i = perm(U[t+e_m, s]).
                          The Ryser walk's alternating subset sums cancel strongly:
                          at k = 20 its own relative error reaches ~1e-10 (measured
                          against cc_submatrices_glynn), so large-k checks of the GPU
                          sweep use cc_submatrices_glynn / perm_minors (Glynn with
                          double-double accumulation, pinned like glynn above).
* minors               -- Laplace identities sum_j A_ij G_ij = perm(A) for every i and
                          classic minors for n <= 7.
* gccb pmf             -- equals |classic perm(U[r+e_m, s])|^2 normalised.
* exact distributions  -- the reference's deterministic known-answer vector
                          (tests/test_exact_distribution_calculator.py:75-142,
                          copied as data into tests/golden/exact_distribution_kat.json).
* guan sequence        -- visits every r <= s exactly once, one digit +-1 per step.
"""
import ctypes
import functools
import itertools
import math
import operator
import os
from typing import List, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle.build import build

            build()
        _lib = ctypes.CDLL(LIB_PATH)
        c128 = _C128
        _lib.orc_perm_classic_matrix.restype = c128
        _lib.orc_perm_classic.restype = c128
        _lib.orc_perm_glynn_matrix.restype = c128
        _lib.orc_perm_glynn.restype = c128
        _lib.orc_perm_guan.restype = c128
        _lib.orc_effective_matrix.restype = ctypes.c_int
        _lib.orc_guan_sequence.restype = ctypes.c_int
        _lib.orc_omp_max_threads.restype = ctypes.c_int
    return _lib


class _C128(ctypes.Structure):
    """C99 double _Complex as returned by value (two SSE doubles on x86-64)."""

    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def _p(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _c(v) -> complex:
    return complex(v.re, v.im)


def _cm(matrix) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128))


def _ci(state) -> np.ndarray:
    a = np.asarray(state)
    if a.size and not np.all(np.asarray(a, dtype=np.float64) == np.round(np.asarray(a, dtype=np.float64))):
        raise ValueError("occupation numbers must be integers")
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).round().astype(np.int32))


# ----------------------------------------------------------------------------------
# Seeded inputs
# ----------------------------------------------------------------------------------
def haar_unitary(m: int, seed: int) -> np.ndarray:
    """Haar-random unitary: QR of a seeded complex Gaussian matrix with the diagonal
    phase fix (Mezzadri, the construction cited by the reference's README [4])."""
    rng = np.random.default_rng(seed)
    z = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / math.sqrt(2.0)
    q, r = np.linalg.qr(z)
    d = np.diag(r)
    return np.ascontiguousarray(q * (d / np.abs(d)))


# ----------------------------------------------------------------------------------
# C oracle bindings
# ----------------------------------------------------------------------------------
def effective_matrix(U, s, t) -> np.ndarray:
    """boson_sampling_utilities.py:595-626 (C restatement)."""
    U = _cm(U)
    s, t = _ci(s), _ci(t)
    m = U.shape[0]
    rows, cols = int(t.sum()), int(s.sum())
    out = np.zeros((max(rows, 1), max(cols, 1)), dtype=np.complex128)
    c = ctypes.c_int(0)
    r = lib().orc_effective_matrix(_p(U), m, _p(s), _p(t), _p(out), ctypes.byref(c))
    if r == 0:
        return np.zeros((0, 0), dtype=np.complex128)
    return out[:r, : c.value].copy()


def perm_classic_matrix(A) -> complex:
    A = _cm(A)
    if A.size == 0:
        return 1.0 + 0j
    return _c(lib().orc_perm_classic_matrix(_p(A), A.shape[0], A.shape[1]))


def perm_classic(U, s, t) -> complex:
    U, s, t = _cm(U), _ci(s), _ci(t)
    return _c(lib().orc_perm_classic(_p(U), U.shape[0], _p(s), _p(t)))


def perm_glynn_matrix(A, mode: int = 1) -> complex:
    A = _cm(A)
    return _c(lib().orc_perm_glynn_matrix(_p(A), A.shape[0], int(mode)))


def perm_glynn(U, s, t, mode: int = 1) -> complex:
    U, s, t = _cm(U), _ci(s), _ci(t)
    return _c(lib().orc_perm_glynn(_p(U), U.shape[0], _p(s), _p(t), int(mode)))


def perm_glynn_batch(A, mode: int = 1, threads: int = 0) -> np.ndarray:
    A = _cm(A)
    b, n, _ = A.shape
    out = np.zeros(b, dtype=np.complex128)
    lib().orc_perm_glynn_batch(_p(A), n, b, int(mode), int(threads), _p(out))
    return out


def perm_ryser(U, s, t) -> complex:
    U, s, t = _cm(U), _ci(s), _ci(t)
    return _c(lib().orc_perm_guan(_p(U), U.shape[0], _p(s), _p(t), 0))


def perm_glynn_py(A) -> complex:
    """Pure-Python restatement of the reference's Glynn loop, kept for timing the
    reference's speed class (bench's Glynn CPU line): delta = 1, column sums
    sums_j = sum_i M[i][j], acc = prod(sums); per Gray flip i: mult = -mult,
    delta_i = -delta_i, sums_j += 2 delta_i M[i][j], acc += mult * reduce(mul, sums); then
    acc / 2^(n-1) (glynn_gray_permanent_calculator.py:41-84). Flip order: reflected Gray
    code, i = ctz(step). Python complex numbers throughout (the reference's numpy scalars
    are slower)."""
    M = np.asarray(A, dtype=np.complex128).tolist()
    n = len(M)
    if n == 0:
        return 1 + 0j
    delta = [1] * n
    mult = 1
    sums = [sum(delta[i] * M[i][j] for i in range(n)) for j in range(n)]
    acc = mult * functools.reduce(operator.mul, sums, 1)
    for step in range(1, 1 << (n - 1)):
        i = (step & -step).bit_length() - 1
        mult = -mult
        delta[i] = -delta[i]
        row = M[i]
        for j in range(n):
            sums[j] += 2 * delta[i] * row[j]
        acc += mult * functools.reduce(operator.mul, sums, 1)
    return complex(acc / 2 ** (n - 1))


def perm_ryser_py(U, s, t) -> complex:
    """Pure-Python restatement of the reference's Ryser walk, kept for timing the
    reference's speed class (bench's GCC-A CPU line): the Guan (reflected mixed-radix
    Gray) walk over r <= s across all m digits, zero-limit digits included
    (bs_permanent_calculator_base.py:123-209), the running binomial product Pi C(s_i, r_i)
    by ratios (:151-164), and sums_j += dr * U[j][i] with the addend
    (-1)^(|s| + |r|) Pi C(s, r) Pi_j sums_j^t_j (ryser_permanent_calculator.py:45-64).
    The digit step follows Knuth's reflected mixed-radix Gray code (the lowest digit that
    can move in its direction moves; the exhausted ones below it turn around)."""
    Ul = np.asarray(U, dtype=np.complex128).tolist()
    s = [int(x) for x in s]
    t = [int(x) for x in t]
    m = len(s)
    rows = [j for j in range(m) if t[j] > 0]
    r = [0] * m
    d = [1] * m
    sums = [0j] * m
    binom = 1.0
    sign = 1
    total = 0j
    while True:
        i = 0
        while i < m and not (0 <= r[i] + d[i] <= s[i]):
            d[i] = -d[i]
            i += 1
        if i == m:
            break
        old = r[i]
        r[i] = old + d[i]
        binom = binom * (s[i] - old) / (old + 1) if d[i] > 0 else binom * old / (s[i] - old + 1)
        sign = -sign
        for j in rows:
            sums[j] += d[i] * Ul[j][i]
        prod = 1 + 0j
        for j in rows:
            prod *= sums[j] ** t[j]
        total += sign * binom * prod
    return complex((-1) ** sum(s) * total)


def cc_submatrices_py(U, s, t) -> List[complex]:
    """Pure-Python restatement of the reference's C&C Ryser submatrices walk, kept for
    timing the reference's speed class on the GCC-B workload: the Guan digit scan over all
    m modes (bs_submatrices_permanent_calculator_base.py:99-148, 160-175), the sums over the
    occupied outputs with their powers (bs_cc_ryser_submatrices_permanent_calculator.py:80-94)
    and the per-step pass over all m minors (:96-119)."""
    Ul = np.asarray(U, dtype=np.complex128).tolist()
    s = [int(x) for x in s]
    t = [int(x) for x in t]
    m = len(s)
    if sum(s) == 1:
        return [complex(v) for v in s]
    perms = [0j] * m
    sums = {j: 0j for j in range(m) if t[j] != 0}
    mult = (-1) ** sum(t)
    binom = 1.0
    r = [0] * m
    u = [1] * m
    while r[-1] <= s[-1]:
        i = 0
        val = r[0] + u[0]
        while val > s[i] or val < 0:
            u[i] = -u[i]
            i += 1
            if i == m:
                return perms
            val = r[i] + u[i]
        last = r[i]
        r[i] = val
        mult = -mult
        binom *= (s[i] - last) / r[i] if r[i] > last else last / (s[i] - r[i])
        prod = 1 + 0j
        for j in sums:
            sums[j] += (r[i] - last) * Ul[j][i]
            prod *= sums[j] ** t[j]
        prod *= mult
        for q in range(m):
            if s[q] == 0 or r[q] == s[q]:
                continue
            perms[q] += prod * (binom / (s[q] / (s[q] - r[q])))
    return perms


def perm_chin_huh(U, s, t) -> complex:
    U, s, t = _cm(U), _ci(s), _ci(t)
    return _c(lib().orc_perm_guan(_p(U), U.shape[0], _p(s), _p(t), 1))


def cc_submatrices(U, s, t, formula: int = 0) -> np.ndarray:
    """All minors perm(U[t, s - e_i]) (one per input mode; 0 where s_i == 0)."""
    U, s, t = _cm(U), _ci(s), _ci(t)
    out = np.zeros(U.shape[0], dtype=np.complex128)
    lib().orc_cc_submatrices(_p(U), U.shape[0], _p(s), _p(t), int(formula), _p(out))
    return out


def perm_minors(A, threads: int = 0) -> np.ndarray:
    """G[i][j] = perm(A without row i and column j)."""
    A = _cm(A)
    n = A.shape[0]
    G = np.zeros((n, n), dtype=np.complex128)
    lib().orc_perm_minors(_p(A), n, int(threads), _p(G))
    return G


def gccb_pmf(U, cur_in, r_sample) -> Tuple[np.ndarray, np.ndarray]:
    U = _cm(U)
    m = U.shape[0]
    pmf = np.zeros(m, dtype=np.float64)
    minors = np.zeros(m, dtype=np.complex128)
    cur_in, r_sample = _ci(cur_in), _ci(r_sample)
    lib().orc_gccb_pmf(_p(U), m, _p(cur_in), _p(r_sample), _p(pmf), _p(minors))
    return pmf, minors


def gccb_sample_batch(U, order, uniforms, steps=None, threads: int = 0) -> np.ndarray:
    """GCC-B samples with caller-supplied randomness (see permanent_oracle.c)."""
    U = _cm(U)
    m = U.shape[0]
    order = np.ascontiguousarray(order, dtype=np.int32)
    uniforms = np.ascontiguousarray(uniforms, dtype=np.float64)
    samples, n = order.shape
    out = np.zeros((samples, m), dtype=np.int32)
    st = None if steps is None else np.ascontiguousarray(steps, dtype=np.int32)
    lib().orc_gccb_sample_batch(
        _p(U), m, n, samples, _p(order), _p(uniforms),
        _p(st) if st is not None else ctypes.c_void_p(0), int(threads), _p(out))
    return out


def guan_sequence(s, max_codes: int = 1 << 20) -> np.ndarray:
    s = _ci(s)
    m = s.shape[0]
    total = int(np.prod(s.astype(np.int64) + 1))
    cap = min(max_codes, total + 1)
    out = np.zeros((cap, m), dtype=np.int32)
    cnt = lib().orc_guan_sequence(m, _p(s), _p(out), cap)
    return out[:cnt].copy()


def omp_max_threads() -> int:
    return int(lib().orc_omp_max_threads())


# ----------------------------------------------------------------------------------
# Textbook brute force (independent of the C code)
# ----------------------------------------------------------------------------------
def perm_by_permutations(A) -> complex:
    """perm(A) = sum over permutations sigma of prod_i A[i, sigma(i)]."""
    A = np.asarray(A, dtype=np.complex128)
    n = A.shape[0]
    if n == 0:
        return 1.0 + 0j
    total = 0j
    for sigma in itertools.permutations(range(n)):
        p = 1 + 0j
        for i in range(n):
            p *= A[i, sigma[i]]
        total += p
    return total


def derangements(n: int) -> int:
    d = [1, 0]
    for k in range(2, n + 1):
        d.append((k - 1) * (d[-1] + d[-2]))
    return d[n]


# ----------------------------------------------------------------------------------
# State enumeration and exact lossy distributions (numpy restatements)
# ----------------------------------------------------------------------------------
def mode_occupation_to_mode_assignment(occ: Sequence[int]) -> Tuple[int, ...]:
    """boson_sampling_utilities.py:60-78."""
    out = []
    for i, c in enumerate(occ):
        out.extend([i] * int(c))
    return tuple(out)


def mode_assignment_to_mode_occupation(assign: Sequence[int], modes: int = 0) -> Tuple[int, ...]:
    """boson_sampling_utilities.py:34-57."""
    if modes == 0:
        modes = max(assign) + 1
    occ = [0] * modes
    for a in assign:
        occ[a] += 1
    return tuple(occ)


def _n_particle_states(n: int, m: int) -> List[Tuple[int, ...]]:
    """boson_sampling_utilities.py:116-152: all n-particle m-mode states, sorted in
    reverse lexicographic order."""
    states = []

    def rec(prefix, left, modes_left):
        if modes_left == 1:
            states.append(tuple(prefix + [left]))
            return
        for c in range(left, -1, -1):
            rec(prefix + [c], left - c, modes_left - 1)

    rec([], n, m)
    return sorted(states, reverse=True)


def generate_possible_states(n: int, m: int, losses: bool = False) -> List[Tuple[int, ...]]:
    """boson_sampling_utilities.py:81-113."""
    if n < 0 or m < 1:
        return []
    if n == 0:
        return [tuple([0] * m)]
    start = 0 if losses else n
    out = []
    for k in range(start, n + 1):
        out.extend(_n_particle_states(k, m))
    return out


def generate_lossy_n_particle_input_states(initial_state, left: int):
    """boson_sampling_utilities.py:155-198: distinct lossy inputs, in first-seen order
    over itertools.combinations of the particle list."""
    m = len(initial_state)
    x0 = []
    for i in range(m):
        x0.extend([i] * int(initial_state[i]))
    if sum(initial_state) == 0:
        return [tuple(initial_state)]
    seen = []
    for comb in itertools.combinations(range(len(x0)), left):
        occ = mode_assignment_to_mode_occupation([x0[e] for e in comb], m)
        if occ not in seen:
            seen.append(occ)
    return seen


def bosonic_space_dimension(n: int, m: int, losses: bool = False) -> int:
    """boson_sampling_utilities.py:201-232."""
    dim = math.comb(n + m - 1, n)
    if not losses:
        return dim
    for k in range(n):
        dim += math.comb(k + m - 1, k)
    return dim


def fixed_losses_probability(U, initial_state, outcome, perm=perm_chin_huh) -> float:
    """bs_distribution_calculator_with_fixed_losses.py:82-151: probability of `outcome`
    when sum(outcome) particles survive."""
    left = int(sum(outcome))
    n0 = int(sum(initial_state))
    total = 0.0
    last = None
    for lossy in generate_lossy_n_particle_input_states(initial_state, left):
        mult = 1
        for i in range(len(lossy)):
            mult *= math.comb(int(initial_state[i]), int(initial_state[i]) - int(lossy[i]))
        sub = abs(perm(U, lossy, outcome)) ** 2
        for c in lossy:
            sub /= math.factorial(int(c))
        total += sub * mult
        last = lossy
    total /= math.factorial(int(sum(last)))
    total /= math.comb(n0, left)
    total *= math.factorial(left)
    for c in outcome:
        total /= math.factorial(int(c))
    return total


def uniform_losses_distribution(U, initial_state, eta: float, perm=perm_chin_huh) -> List[float]:
    """bs_exact_distribution_with_uniform_losses.py:36-105: binomially weighted
    fixed-loss probabilities over all lossy outcomes."""
    n0 = int(sum(initial_state))
    m = len(initial_state)
    weights = [eta ** l * math.comb(n0, l) * (1 - eta) ** (n0 - l) for l in range(n0 + 1)]
    out = []
    for outcome in generate_possible_states(n0, m, losses=True):
        l = int(sum(outcome))
        if l == 0:
            out.append(weights[0])
            continue
        out.append(fixed_losses_probability(U, initial_state, outcome, perm) * weights[l])
    return out


def tvd_error_bound(outcomes_number: int, samples_number: int, error_probability: float) -> float:
    """quantum_computations_utilities.py:95-127."""
    b = -math.log(error_probability) + outcomes_number * math.log(2)
    return math.sqrt(b / (2 * samples_number))


def sample_frequencies(samples, outcomes, samples_number: int) -> np.ndarray:
    """Outcome frequencies as the reference counts them
    (bs_sample_based_distribution_calculator.py:92-106): each sample adds one to the first
    outcome equal to it; samples matching none are dropped; division by samples_number."""
    probs = np.zeros(len(outcomes), dtype=np.float64)
    outs = [np.asarray(o) for o in outcomes]
    for sample in samples:
        sample = np.asarray(sample)
        for j, o in enumerate(outs):
            if o.shape == sample.shape and np.all(sample == o):
                probs[j] += 1
                break
    return probs / samples_number


def total_variation_distance(d1, d2) -> float:
    """quantum_computations_utilities.py:57-78 (loop form)."""
    assert len(d1) == len(d2)
    t = 0.0
    for a, b in zip(d1, d2):
        t += abs(a - b)
    return t / 2


# ----------------------------------------------------------------------------------
# GCC-A layer pmf (generalized_cliffords_simulation_strategy.py:136-247)
# ----------------------------------------------------------------------------------
def _input_substates(state: Sequence[int]) -> List[Tuple[int, ...]]:
    """generalized_cliffords_simulation_strategy.py:94-121 (same order)."""
    if len(state) < 1:
        return [tuple()]
    first, rest = int(state[0]), list(state[1:])
    smaller = _input_substates(rest)
    out = []
    for i in range(first + 1):
        for sub in smaller:
            out.append((i,) + sub)
    return out


def gcca_layer_pmf(U, input_state, r_sample, perm=perm_ryser) -> np.ndarray:
    """Unnormalised pmf over the m candidate outputs r + e_o of the next GCC-A layer."""
    n = int(sum(input_state))
    k = int(sum(r_sample)) + 1
    subs = [s for s in _input_substates(list(input_state)) if sum(s) == k]
    weights = []
    for s in subs:
        kv = [int(input_state[i]) - s[i] for i in range(len(s))]
        l = sum(kv)
        w = math.factorial(l) * math.factorial(n - l) / math.factorial(n)
        for i in range(len(input_state)):
            w *= math.comb(int(input_state[i]), kv[i])
        weights.append(w)
    weights = np.asarray(weights) / sum(weights)
    m = len(input_state)
    pmf = np.zeros(m)
    for o in range(m):
        out = list(r_sample)
        out[o] += 1
        acc = 0.0
        for w, s in zip(weights, subs):
            p = abs(perm(U, s, out)) ** 2
            for c in s:
                p /= math.factorial(c)
            p /= math.factorial(sum(s))
            acc += p * w
        pmf[o] = acc
    return pmf


def cc_submatrices_glynn(U, s, t, threads: int = 0) -> np.ndarray:
    """Oracle-grade minors for large k: each perm(U[t, s - e_i]) by Glynn with re-seeded
    column sums and double-double accumulation (mode 1). Independent of the Ryser walk,
    whose alternating subset sums lose ~1e-10 relative accuracy by k = 20. The minors are
    independent permanents, computed in parallel (threads 0 = all host threads)."""
    s = _ci(s)
    out = np.zeros(len(s), dtype=np.complex128)
    idx = np.nonzero(s)[0]
    mats = []
    for i in idx:
        sm = s.copy()
        sm[i] -= 1
        mats.append(effective_matrix(U, sm, t))
    if not mats or mats[0].size == 0:  # one particle: the empty effective matrix gives 1
        for i in idx:
            out[i] = 1.0
        return out
    out[idx] = perm_glynn_batch(np.stack(mats), 1, threads)
    return out


def gcca_layer_pmf_batch(U, input_state, r_samples, threads: int = 0) -> np.ndarray:
    """gcca_layer_pmf for many layers at once, by the C restatement
    (orc_gcca_layer_pmf_batch, OpenMP over layers): rows of r_samples -> rows of pmfs."""
    U = _cm(U)
    m = U.shape[0]
    inp = _ci(input_state)
    rs = np.ascontiguousarray(np.asarray(r_samples, dtype=np.int32).reshape(-1, m))
    out = np.zeros(rs.shape, dtype=np.float64)
    lib().orc_gcca_layer_pmf_batch(_p(U), m, _p(inp), _p(rs), rs.shape[0], int(threads),
                                   _p(out))
    return out
