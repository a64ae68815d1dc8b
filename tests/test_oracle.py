"""Pins the CPU oracle before it is trusted as the checker of the HIP kernels.

Each test checks an oracle routine against something that does not share its
formula: the reference's own known-answer vector, textbook brute force, closed forms
or invariants chosen so that a dropped term, a wrong sign or a transposed operand
fails at least one of them.
"""
import json
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))

# The nine occupation patterns of the reference's
# tests/test_bs_permanent_calculators.py:88-122 (m = 4).
REFERENCE_PATTERNS = [
    ([1, 1, 1, 1], [1, 1, 1, 1]),
    ([1, 1, 1, 0], [1, 0, 1, 1]),
    ([2, 1, 0, 0], [0, 1, 1, 1]),
    ([1, 1, 1, 0], [2, 1, 0, 0]),
    ([2, 1, 1, 0], [1, 1, 0, 2]),
    ([0, 0, 4, 0], [1, 1, 0, 2]),
    ([1, 1, 0, 2], [0, 1, 3, 0]),
    ([0, 2, 0, 2], [0, 0, 4, 0]),
    ([0, 0, 0, 0], [0, 0, 0, 0]),
]


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_haar_is_unitary_and_seeded(oracle):
    U = oracle.haar_unitary(12, 7)
    assert np.allclose(U.conj().T @ U, np.eye(12), atol=1e-13)
    assert np.array_equal(U, oracle.haar_unitary(12, 7))
    assert not np.array_equal(U, oracle.haar_unitary(12, 8))


def test_effective_matrix_definition(oracle):
    U = oracle.haar_unitary(5, 1)
    s, t = [2, 0, 1, 0, 1], [0, 3, 0, 1, 0]
    M = oracle.effective_matrix(U, s, t)
    rows = np.repeat(np.arange(5), t)
    cols = np.repeat(np.arange(5), s)
    assert np.array_equal(M, U[np.ix_(rows, cols)])
    assert oracle.effective_matrix(U, [0] * 5, t).shape == (0, 0)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7])
def test_classic_equals_permutation_sum(oracle, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    assert rel(oracle.perm_classic_matrix(A), oracle.perm_by_permutations(A)) < 1e-12


@pytest.mark.parametrize("n", [1, 4, 8])
def test_classic_closed_forms(oracle, n):
    J = np.ones((n, n), dtype=complex)
    assert rel(oracle.perm_classic_matrix(J), math.factorial(n)) < 1e-13
    if n > 1:
        assert rel(oracle.perm_classic_matrix(J - np.eye(n)), oracle.derangements(n)) < 1e-13
    d = np.arange(1, n + 1) * (1 + 0.5j)
    assert rel(oracle.perm_classic_matrix(np.diag(d)), np.prod(d)) < 1e-13
    P = np.eye(n)[np.random.default_rng(0).permutation(n)]
    assert rel(oracle.perm_classic_matrix(P), 1.0) < 1e-15


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("n", [1, 2, 3, 6, 8])
def test_glynn_equals_classic(oracle, n, mode):
    A = oracle.haar_unitary(2 * n, 100 + n)[:n, :n]
    assert rel(oracle.perm_glynn_matrix(A, mode), oracle.perm_classic_matrix(A)) < 1e-11


@pytest.mark.parametrize("n", [10, 16, 21])
def test_glynn_closed_forms_large(oracle, n):
    rng = np.random.default_rng(n)
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    # Rank one: every Glynn term is prod_j (sum_i delta_i u_i) v_j; heavy cancellation.
    expect = math.factorial(n) * np.prod(u) * np.prod(v)
    assert rel(oracle.perm_glynn_matrix(np.outer(u, v), 1), expect) < 1e-10
    J = np.ones((n, n), dtype=complex)
    assert rel(oracle.perm_glynn_matrix(J, 1), math.factorial(n)) < 1e-12
    assert rel(oracle.perm_glynn_matrix(J - np.eye(n), 1), oracle.derangements(n)) < 1e-12


def _glynn_clongdouble(A):
    """Independent 80-bit restatement of the Glynn walk: one sequential reflected-Gray walk
    in numpy clongdouble (no chunks, no re-seeding), plain summation."""
    M = np.asarray(A).astype(np.clongdouble)
    n = M.shape[0]
    delta = np.ones(n, dtype=np.longdouble)
    h = M.sum(axis=0)
    acc = np.prod(h)
    for t in range(1, 1 << (n - 1)):
        i = (t & -t).bit_length() - 1
        delta[i] = -delta[i]
        h = h + 2 * delta[i] * M[i]
        acc = acc + (-1 if t & 1 else 1) * np.prod(h)
    return complex(acc / np.longdouble(2) ** (n - 1))


def test_glynn_80bit_mode(oracle):
    """Mode 2 (80-bit walk, the reference grade of the full-size golden values): brute force
    up to n = 8, closed forms (perm J = n!, perm(J - I) = !n, rank one n! prod u prod v)
    at n = 12..20, and the independent clongdouble walk above at n = 10..13; mode 1 is
    within its FP64 error of mode 2."""
    rng = np.random.default_rng(11)
    for n in range(1, 9):
        A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        assert rel(oracle.perm_glynn_batch(A[None], 2)[0], oracle.perm_classic_matrix(A)) < 1e-14
    for n in (12, 16, 20):
        u, v = np.exp(2j * np.pi * rng.random(n)), np.exp(2j * np.pi * rng.random(n))
        J = np.ones((n, n), dtype=complex)
        got = oracle.perm_glynn_batch(np.stack([np.outer(u, v), J, J - np.eye(n)]), 2)
        assert rel(got[0], math.factorial(n) * np.prod(u) * np.prod(v)) < 1e-15
        assert got[1] == math.factorial(n) and got[2] == oracle.derangements(n)
    for n in (10, 13):
        A = oracle.haar_unitary(2 * n, 700 + n)[:n, :n]
        ld = _glynn_clongdouble(A)
        assert rel(oracle.perm_glynn_batch(A[None], 2)[0], ld) < 1e-15
        assert rel(oracle.perm_glynn_matrix(A, 1), ld) < 1e-12


def test_glynn_invariances(oracle):
    n = 9
    rng = np.random.default_rng(3)
    A = oracle.haar_unitary(18, 3)[:n, :n]
    p = oracle.perm_glynn_matrix(A, 1)
    assert rel(oracle.perm_glynn_matrix(A.T, 1), p) < 1e-12
    rp, cp = rng.permutation(n), rng.permutation(n)
    assert rel(oracle.perm_glynn_matrix(A[rp][:, cp], 1), p) < 1e-12
    c = 0.7 - 0.4j
    assert rel(oracle.perm_glynn_matrix(c * A, 1), c ** n * p) < 1e-12
    B = oracle.haar_unitary(8, 4)[:4, :4]
    blk = np.zeros((n + 4, n + 4), dtype=complex)
    blk[:n, :n], blk[n:, n:] = A, B
    assert rel(oracle.perm_glynn_matrix(blk, 1), p * oracle.perm_classic_matrix(B)) < 1e-11


@pytest.mark.parametrize("s,t", REFERENCE_PATTERNS)
def test_guan_calculators_reference_patterns(oracle, s, t):
    U = oracle.haar_unitary(4, 11)
    expect = oracle.perm_classic(U, s, t)
    assert abs(oracle.perm_ryser(U, s, t) - expect) < 1e-12 * max(1, abs(expect))
    assert abs(oracle.perm_chin_huh(U, s, t) - expect) < 1e-12 * max(1, abs(expect))
    if sum(s):
        assert abs(oracle.perm_glynn(U, s, t, 0) - expect) < 1e-12 * max(1, abs(expect))
    else:
        assert oracle.perm_glynn(U, s, t, 0) == 1  # empty effective matrix -> 1


def test_guan_calculators_random_multiplicities(oracle):
    rng = np.random.default_rng(5)
    for trial in range(20):
        m = int(rng.integers(3, 7))
        n = int(rng.integers(1, 7))
        U = oracle.haar_unitary(m, 200 + trial)
        s = np.bincount(rng.integers(0, m, n), minlength=m)
        t = np.bincount(rng.integers(0, m, n), minlength=m)
        expect = oracle.perm_classic(U, s, t)
        assert abs(oracle.perm_ryser(U, s, t) - expect) < 1e-12
        assert abs(oracle.perm_chin_huh(U, s, t) - expect) < 1e-12
        assert abs(oracle.perm_ryser_py(U, s, t) - expect) < 1e-12


def test_python_cc_walk_equals_c_walk(oracle):
    """The pure-Python C&C walk (timing baseline) equals the C restatement, collision
    inputs and zero-limit digits included."""
    rng = np.random.default_rng(8)
    for trial in range(12):
        m = int(rng.integers(3, 8))
        k = int(rng.integers(2, 7))
        U = oracle.haar_unitary(m, 400 + trial)
        s = np.bincount(rng.integers(0, m, k), minlength=m)
        t = np.bincount(rng.integers(0, m, k - 1), minlength=m)
        got = np.array(oracle.cc_submatrices_py(U, s, t))
        assert np.max(np.abs(got - oracle.cc_submatrices(U, s, t))) < 1e-12, (s, t)


def test_python_glynn_walk_equals_c_walk_and_closed_forms(oracle):
    """The pure-Python Glynn walk (timing baseline) equals the C walk in the reference's
    arithmetic on seeded Haar blocks, and the closed forms perm(J_n) = n! and
    perm(J_n - I_n) = !n (derangements)."""
    for n in range(1, 11):
        A = oracle.haar_unitary(2 * n, 500 + n)[:n, :n]
        want = oracle.perm_glynn_matrix(A, 0)
        assert abs(oracle.perm_glynn_py(A) - want) <= 1e-12 * max(1.0, abs(want)), n
    for n in (4, 7, 9):
        J = np.ones((n, n))
        assert abs(oracle.perm_glynn_py(J) - math.factorial(n)) < 1e-9 * math.factorial(n)
        assert abs(oracle.perm_glynn_py(J - np.eye(n)) - oracle.derangements(n)) < 1e-9 * math.factorial(n)
    assert oracle.perm_glynn_py(np.zeros((0, 0))) == 1


def test_python_ryser_walk_on_gcca_shapes(oracle):
    """The pure-Python Ryser walk (timing baseline) on GCC-A shapes: a 0/1 substate of the
    input against a bunched output, over all m digits, equals the C walk."""
    rng = np.random.default_rng(6)
    m = 12
    U = oracle.haar_unitary(m, 300)
    for k in (1, 3, 6):
        s = np.zeros(m, int)
        s[rng.choice(6, k, replace=False)] = 1
        t = np.bincount(rng.integers(0, m, k), minlength=m)
        assert abs(oracle.perm_ryser_py(U, s, t) - oracle.perm_ryser(U, s, t)) < 1e-12


# Cases of the reference's tests/test_bs_submatrices_permanent_calculators.py:41-73.
SUBMATRIX_CASES = [
    ([1, 1, 1, 1], [0, 1, 1, 1]),
    ([1, 3, 0, 0], [0, 1, 1, 1]),
    ([1, 0, 0, 0], [0, 0, 0, 0]),
]


@pytest.mark.parametrize("formula", [0, 1])
@pytest.mark.parametrize("s,t", SUBMATRIX_CASES)
def test_cc_submatrices_against_classic_minors(oracle, s, t, formula):
    U = oracle.haar_unitary(4, 21)
    got = oracle.cc_submatrices(U, s, t, formula)
    for i in range(4):
        if s[i] == 0:
            assert got[i] == 0
            continue
        sm = list(s)
        sm[i] -= 1
        expect = oracle.perm_classic(U, sm, t)
        assert abs(got[i] - expect) < 1e-12, (i, got[i], expect)


def test_cc_submatrices_random_and_laplace(oracle):
    rng = np.random.default_rng(9)
    for trial in range(15):
        m = int(rng.integers(3, 8))
        k = int(rng.integers(2, 7))
        U = oracle.haar_unitary(m, 300 + trial)
        s = np.bincount(rng.integers(0, m, k), minlength=m)
        t = np.bincount(rng.integers(0, m, k - 1), minlength=m)
        minors = oracle.cc_submatrices(U, s, t, 0)
        for i in np.nonzero(s)[0]:
            sm = s.copy()
            sm[i] -= 1
            assert abs(minors[i] - oracle.perm_classic(U, sm, t)) < 1e-12
        # Laplace expansion along the added output row e_o.
        for o in range(m):
            to = t.copy()
            to[o] += 1
            lap = sum(s[i] * minors[i] * U[o, i] for i in range(m))
            assert abs(lap - oracle.perm_classic(U, s, to)) < 1e-12


def test_minors_laplace_and_classic(oracle):
    for n, seed in [(2, 1), (5, 2), (7, 3), (12, 4)]:
        A = oracle.haar_unitary(2 * n, seed)[:n, :n]
        G = oracle.perm_minors(A)
        p = oracle.perm_glynn_matrix(A, 1)
        for i in range(n):
            assert rel(np.sum(A[i] * G[i]), p) < 1e-10
            assert rel(np.sum(A[:, i] * G[:, i]), p) < 1e-10
        if n <= 7:
            for i in range(n):
                for j in range(n):
                    sub = np.delete(np.delete(A, i, 0), j, 1)
                    assert abs(G[i, j] - oracle.perm_classic_matrix(sub)) < 1e-13


def test_gccb_pmf_matches_direct_permanents(oracle):
    m = 6
    U = oracle.haar_unitary(m, 42)
    cur = [1, 1, 0, 1, 0, 0]
    r = [0, 2, 0, 0, 0, 0]
    pmf, _ = oracle.gccb_pmf(U, cur, r)
    direct = []
    for o in range(m):
        t = list(r)
        t[o] += 1
        direct.append(abs(oracle.perm_classic(U, cur, t)) ** 2)
    direct = np.asarray(direct) / sum(direct)
    assert np.allclose(pmf, direct, atol=1e-14, rtol=1e-12)
    assert abs(pmf.sum() - 1) < 1e-14


def test_exact_distribution_known_answer(oracle):
    """The reference's only deterministic known-answer test."""
    with open(os.path.join(HERE, "golden", "exact_distribution_kat.json")) as f:
        kat = json.load(f)
    U = np.asarray(kat["interferometer_real"], dtype=np.complex128)
    dist = oracle.uniform_losses_distribution(U, kat["initial_state"], kat["uniform_transmissivity"])
    assert len(dist) == len(kat["expected_uniform_losses_distribution"])
    assert np.allclose(dist, kat["expected_uniform_losses_distribution"])
    assert abs(sum(dist) - 1) < kat["probability_sum_tolerance"]
    # Same with Ryser and the classic calculator.
    for perm in (oracle.perm_ryser, oracle.perm_classic):
        d2 = oracle.uniform_losses_distribution(U, kat["initial_state"], 0.8, perm)
        assert np.allclose(d2, kat["expected_uniform_losses_distribution"])


def test_exact_distribution_sums_to_one_haar(oracle):
    U = oracle.haar_unitary(4, 77)
    for init in ([1, 1, 1, 0], [2, 1, 0, 0]):
        d = oracle.uniform_losses_distribution(U, init, 0.6, oracle.perm_ryser)
        assert abs(sum(d) - 1) < 1e-12


@pytest.mark.parametrize("s", [[1, 1, 1], [2, 0, 3], [0, 0, 0], [1, 2, 1, 3]])
def test_guan_sequence_is_mixed_radix_gray(oracle, s):
    seq = oracle.guan_sequence(s)
    total = int(np.prod(np.asarray(s) + 1))
    assert len(seq) == total
    assert len({tuple(r) for r in seq}) == total
    assert np.all(seq >= 0) and np.all(seq <= np.asarray(s))
    d = np.abs(np.diff(seq, axis=0))
    assert np.all(d.sum(axis=1) == 1)


def test_gcca_layer_pmf_normalises_to_fixed_loss_marginal(oracle):
    """Summed over the first layer, GCC-A weights reproduce the one-particle marginal
    sum_i s_i |U[o][i]|^2 / n (closed form for k = 1)."""
    U = oracle.haar_unitary(5, 5)
    s = [1, 2, 0, 1, 0]
    pmf = oracle.gcca_layer_pmf(U, s, [0] * 5)
    marg = np.array([sum(s[i] * abs(U[o, i]) ** 2 for i in range(5)) for o in range(5)]) / sum(s)
    assert np.allclose(pmf / pmf.sum(), marg, atol=1e-13)


def test_sample_frequencies_reference_semantics(oracle):
    """First matching outcome takes the count (a repeated outcome gets 0), unmatched
    samples are dropped, division by samples_number (known answer)."""
    samples = [(1, 0), (0, 1), (1, 0), (2, 0)]
    outcomes = [(1, 0), (0, 1), (1, 0)]
    assert np.allclose(oracle.sample_frequencies(samples, outcomes, 4), [0.5, 0.25, 0.0])
    # distinct outcomes: equal to an independent Counter-based frequency
    import collections

    rng = np.random.default_rng(3)
    outs = [tuple(o) for o in oracle.generate_possible_states(3, 4)]
    idx = rng.integers(0, len(outs), 500)
    smp = [outs[i] for i in idx]
    cnt = collections.Counter(smp)
    assert np.allclose(oracle.sample_frequencies(smp, outs, 500), [cnt[o] / 500 for o in outs])


def test_total_variation_distance_closed_forms(oracle):
    assert oracle.total_variation_distance([0.2, 0.8], [0.2, 0.8]) == 0
    assert oracle.total_variation_distance([1, 0, 0], [0, 0.5, 0.5]) == 1
    assert abs(oracle.total_variation_distance([0.5, 0.5], [0.9, 0.1]) - 0.4) < 1e-15


def test_gcca_layer_chain_reproduces_exact_output_distribution(oracle):
    """Composing the GCC-A layers (normalised pmfs, one particle per layer) over every
    sequence of picks gives the exact boson-sampling output distribution
    |perm(U[t, s])|^2 / (prod s! prod t!), computed here by brute-force permutations: the
    generalized C&C algorithm's defining property. Pins the layer weights and the
    normalisation of gcca_layer_pmf (Python) and of its C restatement."""
    import itertools

    m = 4
    U = oracle.haar_unitary(m, 41)
    for s in ([1, 1, 1, 0], [2, 0, 1, 0]):
        n = sum(s)
        exact = {}
        for t in oracle.generate_possible_states(n, m):
            A = oracle.effective_matrix(U, s, t)
            p = abs(oracle.perm_by_permutations(A)) ** 2
            for v in list(s) + list(t):
                p /= math.factorial(v)
            exact[tuple(t)] = p
        chain = {t: 0.0 for t in exact}
        for picks in itertools.product(range(m), repeat=n):
            r, prob = [0] * m, 1.0
            for o in picks:
                pmf = oracle.gcca_layer_pmf(U, s, r)
                pmf_c = oracle.gcca_layer_pmf_batch(U, s, [r])[0]
                assert np.allclose(pmf_c, pmf, rtol=1e-13, atol=1e-16)
                prob *= pmf[o] / pmf.sum()
                r[o] += 1
            chain[tuple(r)] += prob
        for t in exact:
            assert abs(chain[t] - exact[t]) <= 1e-12, (s, t)


def test_gcca_layer_pmf_c_equals_python_restatement(oracle):
    """The C layer pmf (OpenMP over layers) equals the Python restatement on inputs with
    mode multiplicities and on the m = 20, n = 10 workload's shapes."""
    U = oracle.haar_unitary(6, 44)
    inp = [2, 1, 0, 1, 1, 0]
    rng = np.random.default_rng(4)
    layers = []
    for k in range(1, 6):
        r = np.zeros(6, int)
        for o in rng.integers(0, 6, k - 1):
            r[o] += 1
        layers.append(r)
    got = oracle.gcca_layer_pmf_batch(U, inp, layers)
    for r, g in zip(layers, got):
        assert np.allclose(g, oracle.gcca_layer_pmf(U, inp, list(r)), rtol=1e-12, atol=1e-16)
    U20 = oracle.haar_unitary(20, 1)
    inp20 = [1] * 10 + [0] * 10
    layers20 = [np.zeros(20, int), np.eye(20, dtype=int)[3] + np.eye(20, dtype=int)[17]]
    got20 = oracle.gcca_layer_pmf_batch(U20, inp20, layers20)
    for r, g in zip(layers20, got20):
        assert np.allclose(g, oracle.gcca_layer_pmf(U20, inp20, list(r)), rtol=1e-12, atol=1e-17)


def _gccb_reference_loop(oracle, U, order, uniforms, steps, perm):
    """GCC-B with given randomness, composed here from the step formula of
    generalized_cliffords_b_simulation_strategy.py:69-110: particle order[s][k] joins the
    input, pmf[o] proportional to |perm(U[r + e_o, input])|^2, pick = first mode whose
    running sum exceeds the uniform (numpy.random.choice)."""
    m = U.shape[0]
    out = np.zeros((len(order), m), dtype=np.int32)
    for s in range(len(order)):
        cur, r = np.zeros(m, int), np.zeros(m, int)
        for k in range(int(steps[s])):
            cur[order[s][k]] += 1
            amp = np.array([abs(perm(U, cur, r + np.eye(m, dtype=int)[o])) ** 2 for o in range(m)])
            cdf = np.cumsum(amp / amp.sum())
            pick = int(np.searchsorted(cdf, uniforms[s][k], side="right"))
            r[min(pick, m - 1)] += 1
        out[s] = r
    return out


def test_gccb_sample_batch_equals_composed_step_formula(oracle):
    """oracle.gccb_sample_batch (C, OpenMP over samples) against the step formula composed
    in Python from brute-force classic permanents, with forced orders, uniforms and step
    counts (0 and partial steps included)."""
    m, n, S = 5, 3, 40
    U = oracle.haar_unitary(m, 71)
    rng = np.random.default_rng(71)
    order = np.stack([rng.permutation(n) for _ in range(S)]).astype(np.int32)
    unif = rng.random((S, n))
    steps = rng.integers(0, n + 1, S).astype(np.int32)

    def classic(U_, s_, t_):
        return oracle.perm_by_permutations(oracle.effective_matrix(U_, s_, t_))

    want = _gccb_reference_loop(oracle, U, order, unif, steps, classic)
    assert np.array_equal(oracle.gccb_sample_batch(U, order, unif, steps), want)
    full = _gccb_reference_loop(oracle, U, order, unif, np.full(S, n), classic)
    assert np.array_equal(oracle.gccb_sample_batch(U, order, unif), full)
    assert np.all(full.sum(axis=1) == n)


def test_tvd_error_bound_hand_computed(oracle):
    """quantum_computations_utilities.py:95-127: sqrt((-ln p + N ln 2) / (2 S)), checked
    against values worked out by hand (ln 10 = 2.302585093, ln 2 = 0.693147181)."""
    # N = 10, S = 1000, p = 1e-4: (9.210340372 + 6.931471806) / 2000 = 0.008070906089
    assert abs(oracle.tvd_error_bound(10, 1000, 1e-4) - 0.0898382218) < 1e-9
    # N = 1, S = 2, p = 0.5: (0.693147181 + 0.693147181) / 4 = 0.3465735903
    assert abs(oracle.tvd_error_bound(1, 2, 0.5) - 0.5887050113) < 1e-9
    # N = 0, S = 50, p = 1: 0
    assert oracle.tvd_error_bound(0, 50, 1.0) == 0.0
