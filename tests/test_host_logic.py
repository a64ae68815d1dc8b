"""CPU tests of the host-side logic (no GPU, no kernel calls)."""
import math
import pickle
from copy import deepcopy

import numpy as np
import pytest


def test_as_occupation_accepts_reference_input_forms():
    from theboss_amd.states import as_occupation

    assert list(as_occupation([1, 0, 2])) == [1, 0, 2]
    assert list(as_occupation(np.array([1.0, 3.0, 0.0]))) == [1, 3, 0]
    assert list(as_occupation((0, 1))) == [0, 1]
    with pytest.raises(ValueError):
        as_occupation([0.5, 1])
    with pytest.raises(ValueError):
        as_occupation([-1, 1])


def test_compact_batch_order_and_padding():
    from theboss_amd.states import compact_batch

    S = np.array([[0, 2, 0, 1], [0, 0, 0, 0], [3, 0, 0, 0]])
    modes, mult = compact_batch(S)
    assert modes.shape == (3, 2)
    assert list(modes[0]) == [1, 3] and list(mult[0]) == [2, 1]
    assert list(mult[1]) == [0, 0]
    assert modes[2][0] == 0 and list(mult[2]) == [3, 0]


def test_device_compaction_equals_compact_batch():
    """The scatter-based compaction of the all-pairs launch (run here on CPU tensors)
    gives compact_batch's arrays: modes in mode order, padded with mode 0 of multiplicity
    0, width the largest occupied-mode count; including all-empty rows."""
    import torch

    from theboss_amd.permanent_calculators.gpu_permanent_calculators import _compact_on_device
    from theboss_amd.states import compact_batch

    rng = np.random.default_rng(0)
    cases = [np.array([[0, 2, 0, 1], [0, 0, 0, 0], [3, 0, 0, 0]]), np.zeros((2, 3), dtype=np.int64)]
    for _ in range(100):
        B, m, k = (int(x) for x in (rng.integers(1, 40), rng.integers(1, 22), rng.integers(1, 11)))
        cases.append(np.array([np.bincount(rng.integers(0, m, k), minlength=m) for _ in range(B)]))
    for S in cases:
        modes, mult = compact_batch(S)
        dm, du = _compact_on_device(torch, S, "cpu")
        assert np.array_equal(modes, dm.numpy()) and np.array_equal(mult, du.numpy())


def test_effective_matrix_matches_oracle(oracle):
    from theboss_amd.states import effective_matrix

    U = oracle.haar_unitary(5, 2)
    s, t = [0, 2, 1, 0, 0], [1, 0, 0, 2, 0]
    assert np.array_equal(effective_matrix(U, s, t), oracle.effective_matrix(U, s, t))
    assert effective_matrix(U, [0] * 5, t).shape == (0, 0)


def test_mode_assignment_round_trip():
    from theboss_amd.states import mode_assignment_to_mode_occupation, mode_occupation_to_mode_assignment

    occ = (0, 2, 1, 0, 3)
    assign = mode_occupation_to_mode_assignment(occ)
    assert assign == (1, 1, 2, 4, 4, 4)
    assert mode_assignment_to_mode_occupation(assign, 5) == occ


def test_expanded_space_first_columns_embed_lossy_matrix(oracle):
    from theboss_amd.states import prepare_interferometer_matrix_in_expanded_space

    m = 5
    U = oracle.haar_unitary(m, 4) @ np.diag(np.sqrt([0.9, 0.5, 0.7, 1.0, 0.3]))
    E = prepare_interferometer_matrix_in_expanded_space(U)
    assert E.shape == (2 * m, 2 * m)
    assert np.allclose(E[:m, :m], U, atol=1e-12)
    cols = E[:, :m]
    assert np.allclose(cols.conj().T @ cols, np.eye(m), atol=1e-12)


def test_input_substates_match_reference_order(oracle):
    from theboss_amd.simulation_strategies.generalized_cliffords_simulation_strategy import input_substates

    for s in ([1, 2, 0], [2, 1, 1, 0], [1, 1, 1, 1]):
        assert input_substates(s) == oracle._input_substates(s)


def test_clements_layout_counts():
    from theboss_amd.interferometer import clements_layout

    assert len(clements_layout(16)) == 120
    for m in (2, 3, 6, 9):
        units = clements_layout(m)
        assert len(units) == m * (m - 1) // 2
        assert all(0 <= k < m - 1 for _, k in units)


def test_clements_unitary_equals_product_of_unit_matrices():
    """The column-batched mesh equals the product of the embedded 2x2 unit matrices
    [[e^{i phi} cos theta, -sin theta], [e^{i phi} sin theta, cos theta]] applied on the
    left in layout order, built with numpy."""
    import torch

    from theboss_amd.interferometer import clements_layout, clements_unitary

    for m in (2, 5, 8):
        nu = m * (m - 1) // 2
        rng = np.random.default_rng(m)
        th, ph = rng.random(nu) * 3, rng.random(nu) * 6
        want = np.eye(m, dtype=np.complex128)
        for u, (_, k) in enumerate(clements_layout(m)):
            M = np.eye(m, dtype=np.complex128)
            e = np.exp(1j * ph[u])
            M[k, k], M[k, k + 1] = e * np.cos(th[u]), -np.sin(th[u])
            M[k + 1, k], M[k + 1, k + 1] = e * np.sin(th[u]), np.cos(th[u])
            want = M @ want
        got = clements_unitary(torch.from_numpy(th), torch.from_numpy(ph), m).numpy()
        assert np.max(np.abs(got - want)) < 1e-13, m


def test_clements_unitary_cpu_is_unitary():
    import torch

    from theboss_amd.interferometer import clements_unitary

    g = torch.Generator().manual_seed(0)
    th = torch.rand(28, generator=g, dtype=torch.float64)
    ph = torch.rand(28, generator=g, dtype=torch.float64)
    U = clements_unitary(th, ph, 8)
    assert torch.allclose(U.conj().T @ U, torch.eye(8, dtype=torch.complex128), atol=1e-12)


def test_calculators_deepcopy_and_pickle_without_gpu(oracle):
    """The strategy factory deep-copies calculators and spawn pools pickle them
    (simulation_strategy_factory.py:58; nonuniform_losses_approximation_strategy.py:254-257):
    the GPU calculators hold only host data."""
    from theboss_amd.permanent_calculators import BSPermanentCalculatorFactory, PermanentCalculatorType

    U = oracle.haar_unitary(4, 1)
    for t in PermanentCalculatorType:
        calc = BSPermanentCalculatorFactory(U, [1, 1, 0, 0], [0, 1, 1, 0], t).generate_calculator()
        c2 = deepcopy(calc)
        c3 = pickle.loads(pickle.dumps(calc))
        assert np.array_equal(c2.matrix, U) and np.array_equal(c3.matrix, U)


def test_factory_enum_values_and_default(oracle):
    from theboss_amd.permanent_calculators import (
        BSPermanentCalculatorFactory, GPUChinHuhPermanentCalculator, GPURyserPermanentCalculator,
        PermanentCalculatorType)

    assert [int(t) for t in PermanentCalculatorType][:4] == [1, 2, 3, 4]
    f = BSPermanentCalculatorFactory(None, None, None)
    assert isinstance(f.generate_calculator(), GPURyserPermanentCalculator)
    f2 = BSPermanentCalculatorFactory(None, None, None, calculator_type=99)
    assert isinstance(f2.generate_calculator(), GPUChinHuhPermanentCalculator)


def test_calculator_stores_references():
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator

    U = np.eye(3, dtype=complex)
    calc = GPURyserPermanentCalculator(U)
    calc.matrix *= 0.5  # the reference's tests mutate the matrix in place
    assert calc.matrix is U and U[0, 0] == 0.5
    assert calc.input_state == [] and calc.output_state == []


def test_shape_check_raises_attribute_error_before_gpu():
    from theboss_amd.permanent_calculators import GPUGlynnGrayPermanentCalculator

    calc = GPUGlynnGrayPermanentCalculator(np.eye(3, dtype=complex), [1, 1, 1], [1, 1])
    with pytest.raises(AttributeError):
        calc.compute_permanent()


def test_bench_flop_model():
    import bench

    assert bench.sweep_flops(20) == 2 ** 19 * 188
    # one sample: sum_k [2^(k-1)(10k-12) + 8 m k]
    expect = sum(2 ** (k - 1) * (10 * k - 12) + 8 * 400 * k for k in range(1, 21))
    assert bench.sample_flops(20, 400) == expect
    import bench_workloads as W

    # gradient: groups of 8 columns (8n + 114 per step), the last n mod 8 <= 4 columns in
    # one group of 4 (8n + 42), else an overlapping group of 8
    assert W.grad_executed_flops(20) == 2 ** 19 * (2 * 274 + 202)
    assert W.grad_executed_flops(24) == 2 ** 23 * 3 * 306
    assert W.grad_executed_flops(21) == 2 ** 20 * 3 * 282
    assert W.grad_executed_flops(17) == 2 ** 16 * (2 * 250 + 178)


def test_state_enumeration_matches_oracle(oracle):
    from theboss_amd.states import generate_lossy_n_particle_input_states, generate_possible_states

    for n, m, losses in [(2, 4, False), (3, 5, True), (0, 3, False), (4, 2, True)]:
        assert generate_possible_states(n, m, losses) == oracle.generate_possible_states(n, m, losses)
    for init, left in [([1, 1, 1, 0, 0], 1), ([2, 1, 0], 2), ([1, 2, 1], 3)]:
        assert generate_lossy_n_particle_input_states(init, left) == \
            oracle.generate_lossy_n_particle_input_states(init, left)


@pytest.mark.parametrize("n,m,losses", [(0, 3, False), (0, 2, True), (3, 4, False), (4, 5, True),
                                        (5, 1, False), (6, 6, True), (2, 7, False), (7, 3, True)])
def test_outcome_rank_is_the_enumeration_index(oracle, n, m, losses):
    """The device histogram's rank (restated on the host) indexes the reference's
    enumeration order; states outside it rank -1."""
    from theboss_amd.states import outcome_rank

    states = oracle.generate_possible_states(n, m, losses=losses)
    assert [outcome_rank(s, n, losses) for s in states] == list(range(len(states)))
    if m > 1:
        assert outcome_rank([n + 1] + [0] * (m - 1), n, losses) == -1
        if not losses and n > 0:
            assert outcome_rank([n - 1] + [0] * (m - 1), n, losses) == -1


def test_quantum_computations_utilities(oracle):
    import math

    from theboss_amd import quantum_computations_utilities as q

    rng = np.random.default_rng(5)
    a, b = rng.random(30), rng.random(30)
    a, b = a / a.sum(), b / b.sum()
    assert abs(q.count_total_variation_distance(a, b) - oracle.total_variation_distance(a, b)) < 1e-15
    keys = [(i, 30 - i) for i in range(30)]
    da = dict(zip(keys, a))
    db = dict(zip(keys[:20], b[:20]))  # missing keys count as 0
    assert abs(q.count_total_variation_distance_dicts(da, db)
               - oracle.total_variation_distance(a, np.r_[b[:20], np.zeros(10)])) < 1e-15
    for K, N, p in ((10, 1000, 1e-3), (126, 20000, 1e-4)):
        assert q.count_tv_distance_error_bound_of_experiment_results(K, N, p) == oracle.tvd_error_bound(K, N, p)
        n_min = q.compute_minimal_number_of_samples_for_desired_accuracy(K, p, 0.05)
        assert q.count_tv_distance_error_bound_of_experiment_results(K, n_min, p) <= 0.05
        assert q.count_tv_distance_error_bound_of_experiment_results(K, n_min - 2, p) > 0.05
    F = q.compute_qft_matrix(7)
    assert np.allclose(F @ F.conj().T, np.eye(7))
    assert abs(F[1, 1] - np.exp(2j * math.pi / 7) / math.sqrt(7)) < 1e-15
    assert q.compute_qft_matrix(0).size == 0


def test_bobs_dilation_without_per_sample_svd(oracle):
    """The non-uniform BOBS strategy forms each sample's 2m-mode dilation from one SVD:
    its first m columns [u sqrt(eta); diag(sqrt(1 - eta))] v P must carry the sample's
    interferometer W P in the first m rows and the same loss-row Gram matrix as the
    per-sample SVD dilation of the reference (prepare_interferometer_matrix_in_expanded_space),
    which fixes the marginal output distribution."""
    from theboss_amd.simulation_strategies.bobs_strategies import (
        NonuniformLossesApproximationStrategy, expand_batch, generate_qft_matrix_for_first_m_modes)

    class Calc:
        matrix = np.sqrt(np.linspace(0.3, 0.9, 6))[:, None] * oracle.haar_unitary(6, 9)

    st = NonuniformLossesApproximationStrategy(Calc(), 3, seed=1)
    rng = np.random.default_rng(2)
    for _ in range(5):
        P = np.diag(np.r_[np.exp(2j * np.pi * rng.random(3)), np.ones(3)]) @ generate_qft_matrix_for_first_m_modes(3, 6)
        first = st._dilation_left @ (st._right @ P)
        ref = expand_batch((st._initial_matrix @ P)[None])[0][:, :6]
        assert np.allclose(first[:6], st._initial_matrix @ P, atol=1e-13)
        assert np.allclose(first[:6], ref[:6], atol=1e-13)
        assert np.allclose(first[6:].conj().T @ first[6:], ref[6:].conj().T @ ref[6:], atol=1e-13)
        assert np.allclose(first.conj().T @ first, np.eye(6), atol=1e-13)  # isometry


def test_gcca_lock_step_draws_equal_sequential_reference_loop(oracle):
    """GCC-A with its permanents from the oracle (a CPU stub calculator): the samples, drawn
    for all samples at once per particle, equal those of the reference's per-sample loop
    (generalized_cliffords_simulation_strategy.py:249-266: running sum until it exceeds
    random() * sum(pmf)) fed the same uniforms, over pmfs from the oracle."""
    from theboss_amd.simulation_strategies import GeneralizedCliffordsSimulationStrategy

    m, S = 5, 40
    U = oracle.haar_unitary(m, 31)
    inp = [1, 2, 0, 1, 0]
    n = sum(inp)

    class Stub:
        def compute_permanent_batch_arrays(self, ins, outs):
            return np.array([oracle.perm_ryser(U, s, t) for s, t in zip(ins, outs)])

    got = GeneralizedCliffordsSimulationStrategy(Stub(), rng=np.random.default_rng(3)).simulate(inp, S)
    u = np.random.default_rng(3).random((n, S))  # one uniform per sample and particle
    for s in range(S):
        r = [0] * m
        for l in range(n):
            pmf = oracle.gcca_layer_pmf(U, inp, r)
            value, acc, idx = u[l, s] * sum(pmf), 0.0, 0
            for p in pmf:
                acc += p
                if acc > value:
                    break
                idx += 1
            r[idx] += 1
        assert tuple(r) == got[s], s


class _OracleStub:
    """A CPU calculator over the oracle's Ryser, for strategy logic without a GPU."""

    def __init__(self, oracle, U):
        self.matrix, self._oracle = U, oracle

    def compute_permanent_batch_arrays(self, ins, outs):
        return np.array([self._oracle.perm_ryser(self.matrix, s, t) for s, t in zip(ins, outs)])


def test_gcca_uniform_losses_distribution_equals_exact_uniform_loss_distribution(oracle):
    """The distribution the uniform-loss GCC-A fills from its layers
    (generalized_cliffords_uniform_losses_simulation_strategy.py:90-107, 154-173) is the
    exact uniform-loss distribution of the oracle, in generate_possible_states order."""
    from theboss_amd.simulation_strategies import GeneralizedCliffordsUniformLossesSimulationStrategy

    m, eta = 4, 0.6
    U = oracle.haar_unitary(m, 32)
    for init in ([1, 1, 1, 0], [2, 1, 0, 0]):
        st = GeneralizedCliffordsUniformLossesSimulationStrategy(_OracleStub(oracle, U), eta,
                                                                 rng=np.random.default_rng(4))
        dist = st.compute_distribution_up_to_accuracy(init)
        exact = oracle.uniform_losses_distribution(U, init, eta)
        assert np.allclose(dist, exact, atol=1e-12, rtol=0)
        unweighted = st.compute_unweighted_distribution_up_to_accuracy(init)
        assert unweighted[0] == 0 and abs(sum(unweighted) - sum(init)) < 1e-9


def test_gcca_uniform_losses_draws_follow_reference_rules(oracle):
    """Each round draws one survival uniform per sample (a particle is skipped when
    u >= eta, :142-152), then one draw uniform per kept sample; replayed sample by sample
    with the reference's running-sum walk over the oracle's pmfs, the samples agree."""
    from theboss_amd.simulation_strategies import GeneralizedCliffordsUniformLossesSimulationStrategy

    m, S, eta = 4, 30, 0.5
    U = oracle.haar_unitary(m, 33)
    inp = [1, 1, 1, 0]
    got = GeneralizedCliffordsUniformLossesSimulationStrategy(
        _OracleStub(oracle, U), eta, rng=np.random.default_rng(5)).simulate(inp, S)
    rng = np.random.default_rng(5)
    rows = [[0] * m for _ in range(S)]
    for _ in range(sum(inp)):
        kept = np.nonzero(rng.random(S) < eta)[0]
        if kept.size == 0:
            continue
        u = rng.random(kept.size)
        for s, us in zip(kept, u):
            pmf = oracle.gcca_layer_pmf(U, inp, rows[s])
            value, acc, idx = us * sum(pmf), 0.0, 0
            for p in pmf:
                acc += p
                if acc > value:
                    break
                idx += 1
            rows[s][idx] += 1
    assert [list(g) for g in got] == rows
    assert all(g.dtype == np.int64 for g in got)


def test_mode_assignment_gcca_matches_gcca_draws(oracle):
    """The mode-assignment GCC-A walks the same pmfs (a layer depends only on the prefix's
    occupation), so with the same uniforms it returns the same occupations; its memo is keyed
    by ordered mode tuples (mode_assignment_generalized_cliffords_simulation_strategy.py:53-96)."""
    from theboss_amd.simulation_strategies import (
        GeneralizedCliffordsSimulationStrategy, ModeAssignmentGeneralizedCliffordsSimulationStrategyInterface)

    m = 5
    U = oracle.haar_unitary(m, 34)
    inp = [1, 0, 2, 1, 0]
    a = GeneralizedCliffordsSimulationStrategy(_OracleStub(oracle, U), rng=np.random.default_rng(6))
    b = ModeAssignmentGeneralizedCliffordsSimulationStrategyInterface(_OracleStub(oracle, U),
                                                                      rng=np.random.default_rng(6))
    assert a.simulate(inp, 25) == b.simulate(inp, 25)
    for key, outs in b.possible_outputs.items():
        assert outs == [key + (i,) for i in range(m)]
        occ = tuple(int(x) for x in np.bincount(np.asarray(key, dtype=np.int64), minlength=m))
        assert np.allclose(b.pmfs[key], oracle.gcca_layer_pmf(U, inp, list(occ)), rtol=1e-12, atol=0)


def test_strategy_factory_types_and_copies(oracle):
    """Enum values of simulation_strategy_factory.py:36-45; the calculator is deep-copied
    on construction and on set, each GCC strategy gets its own copy, BOBS approximates
    m - hierarchy_level modes, UNIFORM_LOSSES_BOBS shares the factory's calculator, and the
    strategies without a counterpart here raise NotImplementedError."""
    from theboss_amd.distribution_calculators import BosonSamplingExperimentConfiguration
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import (
        GeneralizedCliffordsSimulationStrategy, GeneralizedCliffordsUniformLossesSimulationStrategy,
        LossyNetworksGeneralizedCliffordsSimulationStrategy, LossyStateApproximationSimulationStrategy,
        NonuniformLossesApproximationStrategy, SimulationStrategyFactory, StrategyType)

    assert [int(t) for t in StrategyType] == list(range(1, 10))
    assert [t.name for t in StrategyType] == [
        "FIXED_LOSS", "UNIFORM_LOSS", "CLIFFORD_R", "GCC", "LOSSY_NET_GCC",
        "LOSSLESS_MODES_STRATEGY", "UNIFORM_LOSSES_GCC", "BOBS", "UNIFORM_LOSSES_BOBS"]
    U = oracle.haar_unitary(5, 35)
    cfg = BosonSamplingExperimentConfiguration(U, [1, 1, 1, 1, 0], 4, 5, 2, 2, 0.8, hierarchy_level=2)
    calc = GPURyserPermanentCalculator(U)
    f = SimulationStrategyFactory(cfg, calc)
    assert f.bs_permanent_calculator is not calc
    assert np.array_equal(f.bs_permanent_calculator.matrix, U)
    f.strategy_type = StrategyType.GCC
    s = f.generate_strategy()
    assert isinstance(s, GeneralizedCliffordsSimulationStrategy)
    assert s._bs_permanent_calculator is not f.bs_permanent_calculator
    f.strategy_type = StrategyType.UNIFORM_LOSSES_GCC
    s = f.generate_strategy()
    assert isinstance(s, GeneralizedCliffordsUniformLossesSimulationStrategy) and s._transmissivity == 0.8
    f.strategy_type = StrategyType.LOSSY_NET_GCC
    assert isinstance(f.generate_strategy(), LossyNetworksGeneralizedCliffordsSimulationStrategy)
    assert f.bs_permanent_calculator.matrix.shape == (5, 5)  # the copy was expanded, not the factory's
    f.strategy_type = StrategyType.BOBS
    s = f.generate_strategy()
    assert isinstance(s, NonuniformLossesApproximationStrategy) and s._approximated_modes_number == 3
    f.strategy_type = StrategyType.UNIFORM_LOSSES_BOBS
    f.available_threads_number = 3
    s = f.generate_strategy()
    assert isinstance(s, LossyStateApproximationSimulationStrategy)
    assert s._permanent_calculator is f.bs_permanent_calculator and s._threads_number == 3
    for t in (StrategyType.FIXED_LOSS, StrategyType.UNIFORM_LOSS, StrategyType.CLIFFORD_R,
              StrategyType.LOSSLESS_MODES_STRATEGY):
        f.strategy_type = t
        with pytest.raises(NotImplementedError):
            f.generate_strategy()
    other = GPURyserPermanentCalculator(U * 0.5)
    f.bs_permanent_calculator = other
    assert f.bs_permanent_calculator is not other


def test_layer_rows_are_prefix_plus_one_particle_in_each_mode():
    """states.layer_rows, the compact output rows of a GCC-A layer (prefix + e_o for every
    mode o), rebuilt as dense occupations: equal to the prefixes plus the identity, for
    empty prefixes, prefixes with collisions and mixed widths; at most one slot per mode
    carries particles."""
    from theboss_amd.states import layer_rows

    rng = np.random.default_rng(12)
    m = 11
    for k in (1, 2, 4, 7):
        P = 30
        Pf = np.array([np.bincount(rng.integers(0, m, k - 1), minlength=m) for _ in range(P)]) \
            if k > 1 else np.zeros((P, m), dtype=np.int64)
        rows = layer_rows(Pf)
        assert rows.dtype == np.int32 and rows.shape[:2] == (2, P * m)
        dense = np.zeros((P * m, m), dtype=np.int64)
        np.add.at(dense, (np.repeat(np.arange(P * m), rows.shape[2]), rows[0].ravel()),
                  rows[1].ravel())
        want = (Pf[:, None, :] + np.eye(m, dtype=np.int64)[None]).reshape(-1, m)
        assert np.array_equal(dense, want), k
        for r in range(P * m):
            used = rows[0, r][rows[1, r] > 0]
            assert len(used) == len(set(used.tolist()))
