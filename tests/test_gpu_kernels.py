"""Parity of the gfx950 kernels with the oracle (run on an MI355X).

Bars: FP64 paths at relative error <= 1e-10 against the oracle; integer outputs of the
sampler bit-exact against the oracle sampler driven by the same forced draws.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-10  # FP64 relative-error bar for permanents


@pytest.fixture(scope="module")
def T():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from theboss_amd import _lib

    _lib.load()
    return torch


@pytest.fixture(scope="module")
def ops(T):
    from theboss_amd import ops as o

    return o


def dev_c(T, a):
    return T.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


def dev_i(T, a):
    return T.from_numpy(np.ascontiguousarray(a, dtype=np.int32)).cuda()


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300)


# ------------------------------------------------------------------------------ Glynn
@pytest.mark.parametrize("n", list(range(0, 16)))
def test_glynn_fp64_small_n_vs_oracle(T, ops, oracle, n):
    B = 5
    As = np.stack([oracle.haar_unitary(max(2 * n, 1), 10 * n + b)[:n, :n] for b in range(B)]) \
        if n else np.zeros((B, 0, 0), complex)
    got = ops.glynn_batch(dev_c(T, As)).cpu().numpy()
    if n == 0:
        assert np.all(got == 1)
        return
    ref = np.array([oracle.perm_glynn_matrix(A, 1) for A in As])
    for g, r in zip(got, ref):
        assert abs(g - r) <= REL * abs(r), (n, g, r)


@pytest.mark.parametrize("n", [16, 20, 22])
def test_glynn_fp64_medium_vs_oracle(T, ops, oracle, n):
    As = np.stack([oracle.haar_unitary(2 * n, 1000 + n + b)[:n, :n] for b in range(3)])
    got = ops.glynn_batch(dev_c(T, As)).cpu().numpy()
    ref = oracle.perm_glynn_batch(As, 1)
    for g, r in zip(got, ref):
        assert abs(g - r) <= REL * abs(r), (n, g, r)


@pytest.mark.parametrize("n", [24, 28])
def test_glynn_fp64_closed_forms(T, ops, oracle, n):
    import math

    rng = np.random.default_rng(n)
    u = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    v = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    mats = np.stack([np.outer(u, v), np.ones((n, n)), np.ones((n, n)) - np.eye(n)])
    got = ops.glynn_batch(dev_c(T, mats)).cpu().numpy()
    expect = [math.factorial(n) * np.prod(u) * np.prod(v), math.factorial(n),
              oracle.derangements(n)]
    for g, e in zip(got, expect):
        assert abs(g - e) <= REL * abs(e), (n, g, e)


@pytest.mark.parametrize("n", [14, 18, 23])
def test_glynn_paired_walk_and_plain_fallback_in_one_batch(T, ops, oracle, n):
    """FP64 sizes 14..39 run the paired walk (row 0 folded into two products per pair of
    Gray codes) on every matrix with a row whose entries all exceed 2^-16 of the largest
    entry, and the plain walk on the others, in the same call. One batch holds: a Haar
    block; the same with a zero in every row (plain walk); the best row moved to the fixed
    last row and to the middle (both swapped into row 0); one row clear of tiny entries
    that every other row has (that row is picked); every row with an entry at 2^-17 of the
    largest (plain walk); a block scaled by 2^-30 (the result by 2^-30n, through the
    exponent kept beside the row-0 product). Each against the double-double oracle."""
    rng = np.random.default_rng(300 + n)
    H = [oracle.haar_unitary(2 * n, 3000 + 10 * n + k)[:n, :n] for k in range(6)]
    mats = [H[0]]
    Z = H[1].copy()
    Z[np.arange(n), rng.integers(0, n, n)] = 0.0
    mats.append(Z)
    for k, dest in ((2, n - 1), (3, n // 2)):
        A = H[k].copy()
        best = int(np.argmax(np.min(np.abs(A), axis=1)))
        A[[best, dest]] = A[[dest, best]]
        mats.append(A)
    A = H[4].copy()
    keep = int(rng.integers(0, n))
    tiny = 2.0 ** -20 * np.max(np.abs(A))
    for i in range(n):
        if i != keep:
            A[i, rng.integers(0, n)] = tiny
    mats.append(A)
    A = H[5].copy()
    A[np.arange(n), rng.integers(0, n, n)] = 2.0 ** -17 * np.max(np.abs(A))
    mats.append(A)
    mats.append(H[0] * 2.0 ** -30)
    As = np.stack(mats)
    got = ops.glynn_batch(dev_c(T, As)).cpu().numpy()
    ref = oracle.perm_glynn_batch(As, 1)
    for k, (g, r) in enumerate(zip(got, ref)):
        assert abs(g - r) <= REL * abs(r), (n, k, g, r)
    assert abs(ref[-1] - ref[0] * 2.0 ** (-30 * n)) <= 1e-12 * abs(ref[-1])


@pytest.mark.parametrize("n", [14, 18, 23])
def test_glynn_fp32_paired_walk_and_plain_fallback_in_one_batch(T, ops, oracle, n):
    """FP32 column sums at 14 <= n <= 23 run the paired walk too (quotients by the pair row
    formed in FP64 from the FP64 matrix, then FP32 sums and products) and leave matrices
    without a row clear of small entries to the plain FP32 walk. The same mix of blocks as
    the FP64 test, minus the 2^-30 scaling (the plain FP32 walk keeps no exponent), each
    within 3e-4 relative of the oracle (over 64 Haar blocks the largest measured error at
    these n is 8.1e-5 with the paired walk and 4.8e-4 with the plain one, both at n = 23;
    these blocks sit far below them)."""
    from theboss_amd import _lib

    rng = np.random.default_rng(400 + n)
    H = [oracle.haar_unitary(2 * n, 4000 + 10 * n + k)[:n, :n] for k in range(6)]
    mats = [H[0]]
    Z = H[1].copy()
    Z[np.arange(n), rng.integers(0, n, n)] = 0.0
    mats.append(Z)
    for k, dest in ((2, n - 1), (3, n // 2)):
        A = H[k].copy()
        best = int(np.argmax(np.min(np.abs(A), axis=1)))
        A[[best, dest]] = A[[dest, best]]
        mats.append(A)
    A = H[4].copy()
    keep = int(rng.integers(0, n))
    for i in range(n):
        if i != keep:
            A[i, rng.integers(0, n)] = 2.0 ** -20 * np.max(np.abs(A))
    mats.append(A)
    A = H[5].copy()
    A[np.arange(n), rng.integers(0, n, n)] = 2.0 ** -17 * np.max(np.abs(A))
    mats.append(A)
    As = np.stack(mats)
    got = ops.glynn_batch(dev_c(T, As), _lib.TB_PREC_FP32_COLSUM).cpu().numpy()
    f64 = ops.glynn_batch(dev_c(T, As), _lib.TB_PREC_FP64).cpu().numpy()
    ref = oracle.perm_glynn_batch(As, 1)
    err = np.abs(got - ref) / np.abs(ref)
    print(f"fp32 paired/plain n={n}: rel err per block {np.array2string(err, precision=2)}")
    assert err.max() <= 3e-4, (n, err)
    assert not np.array_equal(got, f64)  # FP32 kernels ran, not the FP64 ones


def test_glynn_fp32_batch_whose_plan_counts_four_waves(T, ops, oracle):
    """n = 14, 200 matrices: the paired FP32 kernel (4 waves per SIMD) plans smaller
    chunks, i.e. more partials per matrix, than a 2- or 3-wave plan would, and the
    workspace must hold them: the partials sit before the FP32 copy of the matrices that
    the plain FP32 walk reads for the matrices the paired kernel leaves (every 7th here,
    a zero in every row). Through both entry points, against the oracle."""
    from theboss_amd import _device as D
    from theboss_amd import _lib

    n, B = 14, 200
    rng = np.random.default_rng(1400)
    mats = []
    for b in range(B):
        A = oracle.haar_unitary(2 * n, 14000 + b)[:n, :n]
        if b % 7 == 3:
            A = A.copy()
            A[np.arange(n), rng.integers(0, n, n)] = 0.0
        mats.append(A)
    As = np.stack(mats)
    ref = oracle.perm_glynn_batch(As, 1)
    Ad = dev_c(T, As)
    got = ops.glynn_batch(Ad, _lib.TB_PREC_FP32_COLSUM).cpu().numpy()
    err = np.abs(got - ref) / np.abs(ref)
    plain = np.arange(B) % 7 == 3
    assert err[~plain].max() <= 3e-4, (int(err.argmax()), err.max())
    # the plain walk's blocks (zeros, more cancellation): a bound that an overwritten
    # FP32 copy of the matrix would miss by orders of magnitude
    assert err[plain].max() <= 1e-2, (int(err.argmax()), err.max())
    out = T.empty(B, dtype=T.complex128, device=Ad.device)
    _lib.check(_lib.load().tb_glynn_batch_noalloc(D.ptr(Ad), n, B, n * n,
                                                  _lib.TB_PREC_FP32_COLSUM, D.ptr(out),
                                                  D.stream_ptr(Ad.device)))
    T.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), got)


@pytest.mark.parametrize("n", [11, 13, 14, 16, 19, 21, 22, 23, 26])
def test_glynn_chunk_plans_across_batch_sizes(T, ops, oracle, n):
    """The chunk plan depends on n, the batch and the residency the launch counts (2 waves
    per SIMD for FP64, 3-4 for the paired FP32 kernel); every plan must fit the workspace
    and give the same permanents. Batches of 1, 5, 37 and 200 (fewer at large n) with
    every 5th matrix left to the plain walk (a zero in every row): FP64 against the same
    matrices one per call (a different plan) to 1e-12, FP32 against FP64 (3e-4 where the
    paired kernel ran, 1e-2 on the plain walk's blocks), and the no-alloc entry point
    bit-identical to the workspace one."""
    from theboss_amd import _device as D
    from theboss_amd import _lib

    rng = np.random.default_rng(7000 + n)
    for B in ((1, 5, 37, 200) if n <= 21 else (1, 5, 37)):
        mats = []
        for b in range(B):
            A = oracle.haar_unitary(2 * n, 70000 + 1000 * n + b)[:n, :n]
            if b % 5 == 4:
                A = A.copy()
                A[np.arange(n), rng.integers(0, n, n)] = 0.0
            mats.append(A)
        As = np.stack(mats)
        Ad = dev_c(T, As)
        f64 = ops.glynn_batch(Ad, _lib.TB_PREC_FP64).cpu().numpy()
        for b in sorted({0, B // 2, B - 1}):
            one = ops.glynn_batch(dev_c(T, As[b:b + 1]), _lib.TB_PREC_FP64).cpu().numpy()[0]
            assert abs(one - f64[b]) <= 1e-12 * abs(f64[b]), (n, B, b)
        for prec in (_lib.TB_PREC_FP64, _lib.TB_PREC_FP32_COLSUM):
            got = ops.glynn_batch(Ad, prec).cpu().numpy()
            out = T.empty(B, dtype=T.complex128, device=Ad.device)
            _lib.check(_lib.load().tb_glynn_batch_noalloc(D.ptr(Ad), n, B, n * n, prec,
                                                          D.ptr(out), D.stream_ptr(Ad.device)))
            T.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy(), got), (n, B, prec)
            if prec == _lib.TB_PREC_FP32_COLSUM and n <= 23:
                err = np.abs(got - f64) / np.abs(f64)
                plain = np.arange(B) % 5 == 4
                assert err[~plain].max() <= 3e-4, (n, B, err.max())
                if plain.any():
                    assert err[plain].max() <= 1e-2, (n, B, err.max())


@pytest.mark.parametrize("n", [39, 40, 41])
def test_glynn_fp64_rank_one_large(T, ops, oracle, n):
    """Near the largest supported size (2^38..2^40 Gray steps; n = 39 is the largest with
    the paired walk, n = 40 the largest with broadcast rows, n = 41 keeps row 0 in SGPRs):
    perm(u v^T) = n! prod u prod v with unit-modulus entries, whose Glynn terms stay far
    below the result."""
    import math

    rng = np.random.default_rng(n)
    u, v = np.exp(2j * np.pi * rng.random(n)), np.exp(2j * np.pi * rng.random(n))
    got = ops.glynn_batch(dev_c(T, np.outer(u, v)[None])).cpu().numpy()[0]
    expect = math.factorial(n) * np.prod(u) * np.prod(v)
    assert abs(got - expect) <= REL * abs(expect), (got, expect)


def test_glynn_deterministic(T, ops, oracle):
    A = oracle.haar_unitary(40, 5)[:20, :20]
    a = ops.glynn_batch(dev_c(T, A[None])).cpu().numpy()
    b = ops.glynn_batch(dev_c(T, A[None])).cpu().numpy()
    assert np.array_equal(a, b)


# ------------------------------------------------------------------------ Guan kernels
PATTERNS = [
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


@pytest.mark.parametrize("ctype", ["CLASSIC", "GLYNN", "CHIN_HUH", "RYSER"])
def test_calculators_reference_patterns(T, oracle, ctype):
    from theboss_amd.permanent_calculators import BSPermanentCalculatorFactory, PermanentCalculatorType

    U = oracle.haar_unitary(4, 11)
    calc = BSPermanentCalculatorFactory(U, None, None, PermanentCalculatorType[ctype]).generate_calculator()
    for s, t in PATTERNS:
        calc.input_state, calc.output_state = s, t
        got = calc.compute_permanent()
        expect = oracle.perm_classic(U, s, t)
        assert abs(got - expect) <= 1e-12 * max(1.0, abs(expect)), (ctype, s, t, got, expect)


def test_guan_random_multiplicities_batch(T, oracle):
    from theboss_amd.permanent_calculators import GPUChinHuhPermanentCalculator, GPURyserPermanentCalculator

    rng = np.random.default_rng(5)
    m = 7
    U = oracle.haar_unitary(m, 99)
    ins, outs = [], []
    for _ in range(40):
        n = int(rng.integers(1, 8))
        ins.append(np.bincount(rng.integers(0, m, n), minlength=m))
        outs.append(np.bincount(rng.integers(0, m, n), minlength=m))
    ref = np.array([oracle.perm_classic(U, s, t) for s, t in zip(ins, outs)])
    for cls in (GPURyserPermanentCalculator, GPUChinHuhPermanentCalculator):
        got = cls(U).compute_permanent_batch(ins, outs)
        assert np.max(np.abs(got - ref)) <= 1e-12, cls.__name__


def test_guan_outer_batch_vs_oracle(T, oracle):
    """All (output, input) pairs of two lists of 4-particle occupations, with collisions
    on either side so that both the direct and the transposed launches run."""
    from theboss_amd.permanent_calculators import GPUChinHuhPermanentCalculator, GPURyserPermanentCalculator

    rng = np.random.default_rng(6)
    m, k = 6, 4
    U = oracle.haar_unitary(m, 98)
    S = np.array([np.bincount(rng.integers(0, m, k), minlength=m) for _ in range(9)] + [[4, 0, 0, 0, 0, 0]])
    Tm = np.array([np.bincount(rng.integers(0, m, k), minlength=m) for _ in range(7)] + [[1, 1, 1, 1, 0, 0]])
    ref = np.array([[oracle.perm_classic(U, s, t) for s in S] for t in Tm])
    for cls in (GPURyserPermanentCalculator, GPUChinHuhPermanentCalculator):
        got = cls(U).compute_permanent_outer(S, Tm).cpu().numpy()
        assert got.shape == (len(Tm), len(S))
        assert np.max(np.abs(got - ref)) <= 1e-12, cls.__name__
    with pytest.raises(AttributeError):
        GPURyserPermanentCalculator(U).compute_permanent_outer(S, np.array([[1, 0, 0, 0, 0, 0]]))


def test_guan_outer_batch_uniform_side_and_cached_copies(T, oracle):
    """The all-pairs path whose S side has one Guan code count (the 0/1 substates of a
    collision-free input, as GCC-A passes them): whole T rows per orientation, T rows with
    and without collisions so both launches run. Same values as the pairwise batch (the
    same launches, per orientation, over the same pairs) and the oracle; an in-place edit
    of the matrix between calls is seen (the device copies are keyed by content)."""
    from itertools import combinations

    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator

    rng = np.random.default_rng(16)
    m, k = 8, 4
    U = oracle.haar_unitary(m, 99)
    S = np.zeros((0, m), dtype=np.int64)
    for c in combinations(range(6), k):
        row = np.zeros(m, dtype=np.int64)
        row[list(c)] = 1
        S = np.vstack([S, row])
    Tm = np.array([np.bincount(rng.choice(m, k, replace=False), minlength=m) for _ in range(6)]
                  + [np.bincount(rng.integers(0, m, k), minlength=m) for _ in range(6)]
                  + [[2, 2, 0, 0, 0, 0, 0, 0]])
    codes_t = np.prod(Tm + 1, axis=1)
    assert (codes_t < 2 ** k).any() and (codes_t == 2 ** k).any()
    calc = GPURyserPermanentCalculator(U)
    got = calc.compute_permanent_outer(S, Tm).cpu().numpy()
    pairs = calc.compute_permanent_batch_arrays(np.tile(S, (len(Tm), 1)), np.repeat(Tm, len(S), axis=0))
    assert np.array_equal(got, pairs.reshape(len(Tm), len(S)))
    ref = np.array([[oracle.perm_classic(U, s, t) for s in S] for t in Tm])
    assert np.max(np.abs(got - ref)) <= 1e-12
    # T rows of one orientation only: all collision-free (none transposed), all with a
    # collision (all transposed); one launch over the rows in order
    for rows in (np.nonzero(codes_t == 2 ** k)[0], np.nonzero(codes_t < 2 ** k)[0]):
        one = calc.compute_permanent_outer(S, Tm[rows]).cpu().numpy()
        assert np.array_equal(one, got[rows])
    calc.matrix *= 0.5  # in place: same object, new content
    again = calc.compute_permanent_outer(S, Tm).cpu().numpy()
    assert np.max(np.abs(again - ref * 0.5 ** k)) <= 1e-12


def test_guan_outer_eight_items_per_wave_and_large_item_fallback(T, oracle):
    """All-pairs launches of rows with at most 8 slots and >= 16384 pairs put eight pairs in
    a wave, 8 lanes each; a wave holding a pair of more than 512 Guan codes walks its pairs
    one after another with all 64 lanes. Rows of 12 particles on 6 modes (2^6..4^6 codes
    per row) mixed with rows of every multiplicity >= 1, so both paths run: against the
    pairwise batch entry over the same pairs and the oracle's Ryser on a few pairs. Bar
    1e-8 of the pair's value (floored at 1e-3 of the largest): at 12 particles with
    multiplicities up to 5 the Ryser terms cancel (sum of |terms| over |perm| is 2e4..2.5e5
    on typical pairs here), so other lane splits and the oracle's sequential walk round
    differently by up to ~1e-9; a wrong walk, sign or lane mapping errs by O(1)."""
    from theboss_amd import _device as D
    from theboss_amd import _lib as L
    from theboss_amd import ops
    from theboss_amd.states import compact_batch

    rng = np.random.default_rng(88)
    m, k = 12, 12
    U = oracle.haar_unitary(m, 12)
    S = np.array([np.bincount(rng.integers(0, 6, k), minlength=m) for _ in range(30)]
                 + [np.bincount(np.r_[np.arange(6), rng.integers(0, 6, k - 6)], minlength=m)
                    for _ in range(10)])
    Tm = np.array([np.bincount(6 + rng.integers(0, 6, k), minlength=m) for _ in range(500)])
    codes_s, codes_t = np.prod(S + 1.0, axis=1), np.prod(Tm + 1.0, axis=1)
    walk = np.minimum(codes_t[:, None], codes_s[None, :])
    assert (walk > 512).any() and (walk <= 512).any()
    sm, su = compact_batch(S)
    tm, tu = compact_batch(Tm)
    assert max(sm.shape[1], tm.shape[1]) <= 8 and len(S) * len(Tm) >= 16384
    Ud, Utd = D.to_device_c128(U), D.to_device_c128(np.ascontiguousarray(U.T))

    def dev(a):
        return D.to_device_i32(np.ascontiguousarray(a))

    got = ops.guan_perm_outer(Ud, dev(tm), dev(tu), dev(sm), dev(su), L.TB_FORMULA_RYSER)
    got = got.cpu().numpy()
    j, i = np.divmod(np.arange(len(Tm) * len(S)), len(S))
    tr = codes_t[j] < codes_s[i]
    want = np.empty(len(j), dtype=np.complex128)
    for flag in (False, True):
        idx = np.nonzero(tr == flag)[0]
        if flag:
            r = ops.guan_perm_batch(Utd, dev(sm[i[idx]]), dev(su[i[idx]]), dev(tm[j[idx]]),
                                    dev(tu[j[idx]]), L.TB_FORMULA_RYSER)
        else:
            r = ops.guan_perm_batch(Ud, dev(tm[j[idx]]), dev(tu[j[idx]]), dev(sm[i[idx]]),
                                    dev(su[i[idx]]), L.TB_FORMULA_RYSER)
        want[idx] = r.cpu().numpy()
    err = np.abs(got.reshape(-1) - want) / np.maximum(np.abs(want), 1e-3 * np.max(np.abs(want)))
    assert err.max() <= 1e-8, err.max()
    for jj, ii in zip(rng.integers(0, len(Tm), 4), rng.integers(0, len(S), 4)):
        ref = oracle.perm_ryser(U, S[ii], Tm[jj])
        assert abs(got[jj, ii] - ref) <= 1e-8 * max(abs(ref), 1e-3 * np.max(np.abs(want)))


def test_guan_layer_rows_from_prefixes_equal_outer(T, oracle):
    """compute_permanent_layer (GCC-A's output rows prefix + e_o built in compact form from
    the prefixes on the host) against compute_permanent_outer on the same rows given
    densely: k = 1 (empty prefixes), prefixes with collisions, substates with
    multiplicities (both orientations), both formulas, to rounding (the layer rows' slots
    are not in mode order); spot values against brute force."""
    from theboss_amd.permanent_calculators import GPUChinHuhPermanentCalculator, GPURyserPermanentCalculator

    rng = np.random.default_rng(33)
    m = 9
    U = oracle.haar_unitary(m, 55)
    for k in (1, 3, 5):
        S = np.array([np.bincount(rng.integers(0, m, k), minlength=m) for _ in range(12)])
        Pf = np.array([np.bincount(rng.integers(0, m, k - 1), minlength=m) for _ in range(7)])
        rows = (Pf[:, None, :] + np.eye(m, dtype=np.int64)[None]).reshape(-1, m)
        for cls in (GPURyserPermanentCalculator, GPUChinHuhPermanentCalculator):
            calc = cls(U)
            got = calc.compute_permanent_layer(S, Pf).cpu().numpy()
            want = calc.compute_permanent_outer(S, rows).cpu().numpy()
            assert got.shape == (len(rows), len(S))
            assert np.max(np.abs(got - want)) <= 1e-13 * max(np.max(np.abs(want)), 1e-300), (k, cls)
        for r, i in zip(rng.integers(0, len(rows), 6), rng.integers(0, len(S), 6)):
            assert abs(got[r, i] - oracle.perm_classic(U, S[i], rows[r])) <= 1e-12
    with pytest.raises(AttributeError):
        GPURyserPermanentCalculator(U).compute_permanent_layer(S, Pf[:, :-1])


def test_guan_perm_outer_entry_equals_pairwise_batch(T, oracle):
    """tb_guan_perm_outer against tb_guan_perm_batch over the same pairs, each pair run in
    the orientation the all-pairs kernel picks (the row with fewer Guan codes enumerated;
    U transposed when that is the output row), both formulas: a packed launch (20,000
    pairs, four per wave, both orientations mixed inside waves) and a small one (one pair
    per wave), with the output side padded to more slots than the input side. Spot pairs
    against brute force."""
    from theboss_amd import _device as D
    from theboss_amd import _lib as L
    from theboss_amd import ops
    from theboss_amd.states import compact_batch

    rng = np.random.default_rng(21)
    m, k = 12, 5
    U = oracle.haar_unitary(m, 77)
    S = np.array([np.bincount(rng.integers(0, m, k), minlength=m) for _ in range(30)]
                 + [np.bincount(rng.choice(m, k, replace=False), minlength=m) for _ in range(10)])
    Tm = np.array([np.bincount(rng.integers(0, m, k), minlength=m) for _ in range(500)])
    sm, su = compact_batch(S)
    tm, tu = compact_batch(Tm)
    pad = 9 - tm.shape[1]
    tm, tu = np.pad(tm, ((0, 0), (0, pad))), np.pad(tu, ((0, 0), (0, pad)))
    Ud, Utd = D.to_device_c128(U), D.to_device_c128(np.ascontiguousarray(U.T))

    def dev(a):
        return D.to_device_i32(np.ascontiguousarray(a))

    codes_s, codes_t = np.prod(S + 1.0, axis=1), np.prod(Tm + 1.0, axis=1)
    for n_out in (500, 7):
        for formula in (L.TB_FORMULA_RYSER, L.TB_FORMULA_CHIN_HUH):
            got = ops.guan_perm_outer(Ud, dev(tm[:n_out]), dev(tu[:n_out]), dev(sm), dev(su),
                                      formula).cpu().numpy()
            assert got.shape == (n_out, len(S))
            j, i = np.divmod(np.arange(n_out * len(S)), len(S))
            tr = codes_t[j] < codes_s[i]
            assert tr.any() and (~tr).any()
            want = np.empty(n_out * len(S), dtype=np.complex128)
            for flag in (False, True):
                idx = np.nonzero(tr == flag)[0]
                if flag:  # the output row enumerated: perm(U^T[s, t])
                    r = ops.guan_perm_batch(Utd, dev(sm[i[idx]]), dev(su[i[idx]]),
                                            dev(tm[j[idx]]), dev(tu[j[idx]]), formula)
                else:
                    r = ops.guan_perm_batch(Ud, dev(tm[j[idx]]), dev(tu[j[idx]]),
                                            dev(sm[i[idx]]), dev(su[i[idx]]), formula)
                want[idx] = r.cpu().numpy()
            err = np.max(np.abs(got.reshape(-1) - want)) / np.max(np.abs(want))
            assert err <= 1e-13, (n_out, formula, err)
    for jj, ii in zip(rng.integers(0, 7, 12), rng.integers(0, len(S), 12)):
        assert abs(got[jj, ii] - oracle.perm_classic(U, S[ii], Tm[jj])) <= 1e-12


def _occupations(rng, m, ks):
    return np.array([np.bincount(rng.integers(0, m, int(k)), minlength=m) for k in ks])


@pytest.mark.parametrize("formula", [0, 1])
def test_guan_packed_waves_vs_one_item_per_wave(T, ops, oracle, formula):
    """A launch of >= 16384 items runs four items per wave: 16 lanes per item, or all 64
    lanes in turn for the wave's items when one of them has more than 1024 Guan codes;
    smaller launches run one item per wave. The same items in one launch of 16387 (not a
    multiple of 4) and in launches of 4000 agree to rounding (the rank ranges differ), and
    a strided subset (to 1e-12) plus every large item (to the relative FP64 bar) agree
    with the oracle. Two batches: 1..7
    particles (at most 7 rows, the 8-row kernel), then the same with 40 scattered
    16-particle items of 2048 codes (the 16-row kernel, both wave paths)."""
    from theboss_amd.states import compact_batch

    rng = np.random.default_rng(17 + formula)
    m, B = 12, 16387
    U = oracle.haar_unitary(m, 77)
    Ud = dev_c(T, U)
    ks = rng.integers(1, 8, B)
    S, Tm = _occupations(rng, m, ks), _occupations(rng, m, ks)
    for with_big in (False, True):
        S2, T2 = S.copy(), Tm.copy()
        big = np.sort(rng.choice(B, 40, replace=False)) if with_big else np.zeros(0, int)
        for b in big:
            S2[b] = 0
            S2[b, :6] = [3, 3, 3, 3, 3, 1]
            T2[b] = np.bincount(rng.integers(0, m, 16), minlength=m)
        im, iu = compact_batch(S2)
        om, ou = compact_batch(T2)
        assert (om.shape[1] <= 8) != with_big
        args = [dev_i(T, a) for a in (om, ou, im, iu)]
        full = ops.guan_perm_batch(Ud, *args, formula=formula).cpu().numpy()
        parts = np.concatenate([ops.guan_perm_batch(Ud, *[a[i:i + 4000] for a in args],
                                                    formula=formula).cpu().numpy()
                                for i in range(0, B, 4000)])
        assert np.max(np.abs(full - parts)) <= 1e-12, with_big
        small = np.setdiff1d(np.arange(0, B, 97), big)
        ref = np.array([oracle.perm_classic(U, S2[b], T2[b]) for b in small])
        assert np.max(np.abs(full[small] - ref)) <= 1e-12, with_big
        if with_big:  # 2048-term sums with binomial weights up to 3^5: the FP64 bar
            ref = np.array([oracle.perm_glynn(U, S2[b], T2[b]) for b in big])
            assert np.max(np.abs(full[big] - ref) / np.abs(ref)) <= REL


def test_cc_minors_collision_inputs_packed_waves(T, ops, oracle):
    """The C&C minors of collision inputs (mixed-radix kernel, minors mode) in one launch
    of 16387 items of 12 particles (codes from 2^5 to 2^12: packed waves and waves that
    walk their items in turn) against launches of 4000 and, on a strided subset, the
    oracle, to the relative FP64 bar; slots of an item beyond its occupied modes hold 0."""
    from theboss_amd.states import compact_batch

    rng = np.random.default_rng(23)
    m, B, k = 12, 16387, 12
    U = oracle.haar_unitary(m, 78)
    S = _occupations(rng, m, np.full(B, k))
    Tm = _occupations(rng, m, np.full(B, k - 1))
    codes = np.prod(S + 1, axis=1)
    assert codes.min() <= 1024 < codes.max()
    im, iu = compact_batch(S)
    om, ou = compact_batch(Tm)
    args = [dev_i(T, a) for a in (om, ou, im, iu)]
    Ud = dev_c(T, U)
    full = ops.cc_submatrices_batch(Ud, *args, k, False).cpu().numpy()
    parts = np.concatenate([ops.cc_submatrices_batch(Ud, *[a[i:i + 4000] for a in args], k,
                                                     False).cpu().numpy()
                            for i in range(0, B, 4000)])
    assert rel(full, parts) <= REL  # 16 against 64 lanes per item: another summation order
    assert np.all(full[iu == 0] == 0)
    for b in range(0, B, 211):
        ref = oracle.cc_submatrices(U, S[b], Tm[b])
        live = iu[b] > 0
        assert rel(full[b][live], ref[im[b][live]]) <= REL, b


def test_shape_mismatch_raises_attribute_error(T, oracle):
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator

    calc = GPURyserPermanentCalculator(oracle.haar_unitary(4, 1), [1, 1, 0, 0], [1, 1, 0])
    with pytest.raises(AttributeError):
        calc.compute_permanent()


# ---------------------------------------------------------------- C&C submatrices
@pytest.mark.parametrize("s,t", [
    ([1, 1, 1, 1], [0, 1, 1, 1]),
    ([1, 3, 0, 0], [0, 1, 1, 1]),
    ([1, 0, 0, 0], [0, 0, 0, 0]),
])
def test_submatrices_reference_cases(T, oracle, s, t):
    from theboss_amd.permanent_calculators import GPUBSCCRyserSubmatricesPermanentCalculator

    U = oracle.haar_unitary(4, 21)
    got = np.asarray(GPUBSCCRyserSubmatricesPermanentCalculator(U, np.array(s, float), t).compute_permanents())
    ref = oracle.cc_submatrices(U, s, t)
    assert got.shape == (4,)
    assert np.max(np.abs(got - ref)) <= 1e-12


@pytest.mark.parametrize("k", [2, 3, 5, 8, 12, 16, 20])
def test_submatrices_binary_sweep_vs_oracle(T, oracle, k):
    from theboss_amd.permanent_calculators import GPUBSCCRyserSubmatricesPermanentCalculator

    rng = np.random.default_rng(k)
    m = 3 * k
    U = oracle.haar_unitary(m, 500 + k)
    s = np.zeros(m, int)
    s[rng.choice(m, k, replace=False)] = 1
    t = np.bincount(rng.integers(0, m, k - 1), minlength=m)  # outputs may collide
    got = np.asarray(GPUBSCCRyserSubmatricesPermanentCalculator(U, s, t).compute_permanents())
    # Oracle-grade minors (Glynn, double-double); the Ryser restatement of the reference
    # (cc_submatrices) carries ~1e-10 error of its own at k = 20, so it is held to 1e-9.
    ref = oracle.cc_submatrices_glynn(U, s, t) if k >= 12 else oracle.cc_submatrices(U, s, t)
    assert rel(got, ref) <= REL, (k, rel(got, ref))
    occ = np.nonzero(s)[0]
    assert np.max(np.abs(got[occ] - ref[occ]) / np.abs(ref[occ])) <= REL  # every minor
    assert rel(got, oracle.cc_submatrices(U, s, t)) <= (1e-9 if k >= 16 else REL)


def test_submatrices_collision_inputs_vs_oracle(T, oracle):
    from theboss_amd.permanent_calculators import GPUBSCCRyserSubmatricesPermanentCalculator

    rng = np.random.default_rng(3)
    for trial in range(10):
        m = 6
        k = int(rng.integers(2, 8))
        U = oracle.haar_unitary(m, 700 + trial)
        s = np.bincount(rng.integers(0, 3, k), minlength=m)
        t = np.bincount(rng.integers(0, m, k - 1), minlength=m)
        got = np.asarray(GPUBSCCRyserSubmatricesPermanentCalculator(U, s, t).compute_permanents())
        ref = oracle.cc_submatrices(U, s, t)
        assert np.max(np.abs(got - ref)) <= 1e-12, (s, t)


@pytest.mark.parametrize("k", [12, 16, 21, 22, 26, 28])
def test_sweep_paired_and_plain_items_in_one_batch(T, ops, oracle, k):
    """Sweeps of 8 <= k <= 31 whose chunks hold at least 2^4 codes (here: 7 items, chunks
    of 2^5) run the paired walk (particle 1 folded into two products per pair of Gray
    codes) on every item with a particle whose column entries all exceed
    2^-16 of the item's largest entry, the plain walk on the others, in one launch. The
    batch: a Haar item; zeros in every particle's column (plain); the best particle moved
    to particle 0, 1 and k-1 (swapped into particle 1 and back for the output); every
    column with an entry at 2^-17 of the largest (plain); an item scaled by 2^-30 (minors by
    2^-30(k-1)). Every minor against oracle permanents of the particle-deleted matrices
    (k <= 26; at k = 28 against the FP64 Glynn kernel, whose n = 28 values are pinned to
    the 80-bit golden fixtures)."""
    rng = np.random.default_rng(400 + k)
    R = k - 1
    base = [oracle.haar_unitary(2 * k, 4000 + 10 * k + i)[:k, :R] for i in range(6)]
    items = [base[0]]
    Z = base[1].copy()
    Z[np.arange(k), rng.integers(0, R, k)] = 0.0
    items.append(Z)
    for i, dest in ((2, 0), (3, 1), (4, k - 1)):
        V = base[i].copy()
        best = int(np.argmax(np.min(np.abs(V), axis=1)))
        V[[best, dest]] = V[[dest, best]]
        items.append(V)
    V = base[5].copy()
    V[np.arange(k), rng.integers(0, R, k)] = 2.0 ** -17 * np.max(np.abs(V))
    items.append(V)
    items.append(base[0] * 2.0 ** -30)
    Vts = np.stack(items)  # item x particle x row
    got = ops.sweep_minors_batch(dev_c(T, Vts), k).cpu().numpy()
    for b, V in enumerate(Vts):
        subs = np.stack([np.delete(V, nu, axis=0).T for nu in range(k)])
        ref = oracle.perm_glynn_batch(subs, 1) if k <= 26 else ops.glynn_batch(dev_c(T, subs)).cpu().numpy()
        assert rel(got[b], ref) <= REL, (k, b, rel(got[b], ref))
    assert rel(got[-1], got[0] * 2.0 ** (-30 * R)) <= 1e-12


def test_sweep_every_k_laplace_vs_glynn(T, ops, oracle):
    """Every sweep instantiation k = 2..31 (broadcast columns; the lane offsets, group
    counts and static bits differ per k) and the SGPR-row ones up to k = 34: the Laplace
    identity against the Glynn kernel, and against the oracle up to k = 20."""
    for k in range(2, 35):
        A = oracle.haar_unitary(2 * k, 950 + k)[:k, :k]
        rows = sorted({0, k // 2, k - 1})
        Vts = np.stack([np.delete(A, o, axis=0).T for o in rows])
        minors = ops.sweep_minors_batch(dev_c(T, Vts), k).cpu().numpy()
        perm = ops.glynn_batch(dev_c(T, A[None])).cpu().numpy()[0]
        if k <= 20:
            ref = oracle.perm_glynn_matrix(A, 1)
            assert abs(perm - ref) <= REL * abs(ref), k
        for row, o in enumerate(rows):
            lap = np.sum(A[o, :] * minors[row])
            assert abs(lap - perm) <= REL * abs(perm), (k, o, lap, perm)


def test_glynn_every_broadcast_n(T, ops, oracle):
    """Every FP64 wide-kernel instantiation with broadcast rows: n = 11..26 against the
    oracle, n = 27..39 by the rank-one closed form (n = 40 has its own test)."""
    import math

    for n in range(11, 27):
        As = np.stack([oracle.haar_unitary(2 * n, 4000 + 10 * n + b)[:n, :n] for b in range(2)])
        got = ops.glynn_batch(dev_c(T, As)).cpu().numpy()
        ref = oracle.perm_glynn_batch(As, 1)
        assert np.max(np.abs(got - ref) / np.abs(ref)) <= REL, n
    for n in range(27, 40):
        rng = np.random.default_rng(n)
        u, v = np.exp(2j * np.pi * rng.random(n)), np.exp(2j * np.pi * rng.random(n))
        got = ops.glynn_batch(dev_c(T, np.outer(u, v)[None])).cpu().numpy()[0]
        expect = math.factorial(n) * np.prod(u) * np.prod(v)
        assert abs(got - expect) <= REL * abs(expect), n


@pytest.mark.parametrize("k", [22, 26, 32])
def test_sweep_large_k_laplace_vs_glynn(T, ops, oracle, k):
    """Broadcast columns up to k = 31, particle 1's first values in SGPRs above. For
    every output row o of a random k x k matrix A: sum_nu A[o][nu] minor_nu(A without
    row o) = perm(A), with perm(A) from the Glynn kernel (and from the oracle at k = 22)."""
    A = oracle.haar_unitary(2 * k, 900 + k)[:k, :k]
    Vts = np.stack([np.delete(A, o, axis=0).T for o in (0, k // 2, k - 1)])  # particle-major
    minors = ops.sweep_minors_batch(dev_c(T, Vts), k).cpu().numpy()
    perm = ops.glynn_batch(dev_c(T, A[None])).cpu().numpy()[0]
    if k == 22:
        ref = oracle.perm_glynn_matrix(A, 1)
        assert abs(perm - ref) <= REL * abs(ref)
    for row, o in enumerate((0, k // 2, k - 1)):
        lap = np.sum(A[o, :] * minors[row])
        assert abs(lap - perm) <= REL * abs(perm), (k, o, lap, perm)


def test_sweep_rank_one_at_max_k(T, ops):
    """The largest sweep, k = TB_SWEEP_MAX_K = 40: for V[nu][j] = a_nu b_j,
    minor_nu = (k-1)! prod_{mu != nu} a_mu prod_j b_j."""
    import math

    from theboss_amd import _lib

    k = _lib.TB_SWEEP_MAX_K
    rng = np.random.default_rng(k)
    a, b = np.exp(2j * np.pi * rng.random(k)), np.exp(2j * np.pi * rng.random(k - 1))
    minors = ops.sweep_minors_batch(dev_c(T, np.outer(a, b)[None]), k).cpu().numpy()[0]
    expect = math.factorial(k - 1) * np.prod(a) / a * np.prod(b)
    assert np.max(np.abs(minors - expect) / np.abs(expect)) <= REL


def test_sweep_k20_laplace_identity(T, ops, oracle):
    """Size-independent property at k = 20: sum_nu U[o][nu] minor_nu = perm(U[t+e_o, s])."""
    k = 20
    U = oracle.haar_unitary(40, 4)
    A = U[:k, :k]  # rows = outputs 0..19, columns = inputs 0..19
    Vt = A[: k - 1, :].T.copy()  # particle nu -> column of the first k-1 output rows
    minors = ops.sweep_minors_batch(dev_c(T, Vt[None]), k).cpu().numpy()[0]
    lap = np.sum(A[k - 1, :] * minors)
    ref = oracle.perm_glynn_matrix(A, 1)
    assert abs(lap - ref) <= REL * abs(ref)


# ------------------------------------------------------------------------- Gradient
@pytest.mark.parametrize("n", list(range(1, 14)))
def test_grad_vs_oracle_minors(T, ops, oracle, n):
    """Every n below the single-sweep kernel (n sweeps of the C&C kernel)."""
    As = np.stack([oracle.haar_unitary(2 * n, 40 + n + b)[:n, :n] for b in range(3)])
    perm, G = ops.perm_grad_batch(dev_c(T, As))
    perm, G = perm.cpu().numpy(), G.cpu().numpy()
    for b in range(3):
        ref_p = oracle.perm_classic_matrix(As[b]) if n <= 9 else oracle.perm_glynn_matrix(As[b], 1)
        assert abs(perm[b] - ref_p) <= 1e-12
        assert np.max(np.abs(G[b] - oracle.perm_minors(As[b]))) <= 1e-12


@pytest.mark.parametrize("n,B", [(14, 3), (15, 1), (17, 100), (19, 400), (20, 3), (21, 1), (22, 200)])
def test_grad_single_sweep_vs_oracle_minors(T, ops, oracle, n, B):
    """The one-sweep-per-column-group gradient (n = 14..26): every minor against the
    oracle's double-double Glynn minors; column groups of 8 that overlap (n % 8 != 0),
    one-item and many-item plans (chunk sizes from 2^2 to 2^11)."""
    As = np.stack([oracle.haar_unitary(2 * n, 500 + 7 * n + b)[:n, :n] for b in range(B)])
    perm, G = ops.perm_grad_batch(dev_c(T, As))
    perm, G = perm.cpu().numpy(), G.cpu().numpy()
    for b in sorted({0, B // 2, B - 1}):
        Gref = oracle.perm_minors(As[b])
        assert rel(G[b], Gref) <= REL, (n, b)
        p = oracle.perm_glynn_matrix(As[b], 1)
        assert abs(perm[b] - p) <= REL * abs(p), (n, b)


@pytest.mark.parametrize("n", [16, 18, 23, 24, 25, 26, 27])
def test_grad_single_sweep_laplace_large(T, ops, oracle, n):
    """Sizes outside the all-minors oracle test (n >= 23 is too slow for all n^2 oracle
    minors): Laplace expansion along every row and every column equals the Glynn
    kernel's permanent, and three minors against the oracle. n = 27 is past the
    single-sweep kernel (one C&C sweep per row instead)."""
    As = np.stack([oracle.haar_unitary(2 * n, 900 + n + b)[:n, :n] for b in range(2)])
    Ad = dev_c(T, As)
    perm, G = ops.perm_grad_batch(Ad)
    G = G.cpu().numpy()
    pg = ops.glynn_batch(Ad).cpu().numpy()
    for b in range(2):
        rows = np.sum(As[b] * G[b], axis=1)
        cols = np.sum(As[b] * G[b], axis=0)
        assert np.max(np.abs(rows - pg[b])) <= REL * abs(pg[b]), (n, b)
        assert np.max(np.abs(cols - pg[b])) <= REL * abs(pg[b]), (n, b)
    for (i, j) in ((0, 0), (n // 2, n - 1), (n - 1, 3)):
        minor = np.delete(np.delete(As[1], i, axis=0), j, axis=1)
        ref = oracle.perm_glynn_matrix(minor, 1)
        assert abs(G[1, i, j] - ref) <= REL * np.max(np.abs(G[1])), (n, i, j)


def test_grad_n12_laplace_and_oracle(T, ops, oracle):
    n = 12
    A = oracle.haar_unitary(24, 8)[:n, :n]
    perm, G = ops.perm_grad_batch(dev_c(T, A[None]))
    G = G.cpu().numpy()[0]
    ref = oracle.perm_minors(A)
    assert rel(G, ref) <= REL
    p = oracle.perm_glynn_matrix(A, 1)
    for i in range(n):
        assert abs(np.sum(A[i] * G[i]) - p) <= REL * abs(p)


# -------------------------------------------------------------------------- Sampler
def _forced(rng, S, n):
    order = np.stack([rng.permutation(n) for _ in range(S)]).astype(np.int32)
    unif = rng.random((S, n))
    return order, unif


@pytest.mark.parametrize("m,n,S", [(6, 3, 64), (10, 5, 200), (24, 8, 64)])
def test_batched_sampler_forced_draws_match_oracle(T, ops, oracle, m, n, S):
    rng = np.random.default_rng(m + n)
    U = oracle.haar_unitary(m, m * n)
    particles = np.arange(n, dtype=np.int32)
    order, unif = _forced(rng, S, n)
    out, trace = ops.gccb_sample(dev_c(T, U), dev_i(T, particles), S, 1, 1.0, dev_i(T, order),
                                 T.from_numpy(unif).cuda(), None, trace=True)
    expect = oracle.gccb_sample_batch(U, order, unif)
    assert np.array_equal(out.cpu().numpy(), expect)
    # Per-step minors against the oracle's C&C sweep on the same prefix.
    tr = trace.cpu().numpy()
    for s in range(min(S, 8)):
        cur = np.zeros(m, int)
        r = np.zeros(m, int)
        hist = []
        for k in range(1, n + 1):
            cur[particles[order[s, k - 1]]] += 1
            ref = oracle.cc_submatrices(U, cur, r)
            got = tr[s, k - 1, :k]
            want = np.array([ref[particles[order[s, nu]]] for nu in range(k)])
            assert rel(got, want) <= REL
            pmf, _ = oracle.gccb_pmf(U, cur, r)
            r[int(np.searchsorted(np.cumsum(pmf), unif[s, k - 1], side="right"))] += 1
            hist.append(k)


@pytest.mark.parametrize("m,n,S", [(5, 1, 65), (30, 12, 33), (30, 13, 100), (40, 14, 1),
                                   (24, 2, 257)])
def test_batched_sampler_edges_match_oracle(T, ops, oracle, m, n, S):
    """Edges of the launch structure: one photon; n = 12 (every step in the fused launch);
    n = 13, 14 (the fused steps then separate sweep and step launches); a single sample;
    sample counts that are not multiples of the 64-sample init blocks."""
    rng = np.random.default_rng(100 * n + S)
    U = oracle.haar_unitary(m, 7 * m + n)
    order, unif = _forced(rng, S, n)
    out, _ = ops.gccb_sample(dev_c(T, U), dev_i(T, np.arange(n)), S, 1, 1.0, dev_i(T, order),
                             T.from_numpy(unif).cuda(), None)
    out = out.cpu().numpy()
    assert np.array_equal(out, oracle.gccb_sample_batch(U, order, unif))
    assert np.all(out.sum(axis=1) == n)


def test_batched_sampler_uniform_loss_forced_steps(T, ops, oracle):
    m, n, S = 12, 6, 128
    rng = np.random.default_rng(1)
    U = oracle.haar_unitary(m, 12)
    particles = np.arange(n, dtype=np.int32)
    order, unif = _forced(rng, S, n)
    steps = rng.integers(0, n + 1, S).astype(np.int32)
    out, _ = ops.gccb_sample(dev_c(T, U), dev_i(T, particles), S, 1, 0.5, dev_i(T, order),
                             T.from_numpy(unif).cuda(), dev_i(T, steps))
    expect = oracle.gccb_sample_batch(U, order, unif, steps)
    assert np.array_equal(out.cpu().numpy(), expect)
    assert np.array_equal(out.cpu().numpy().sum(1), steps)


def test_batched_sampler_philox_statistics(T, ops, oracle):
    """Device RNG path: samples follow the exact distribution (TVD bound of the
    reference's tests, quantum_computations_utilities.py:95-127)."""
    m, n, S = 4, 2, 20000
    U = oracle.haar_unitary(m, 31)
    particles = np.array([0, 1], dtype=np.int32)
    out, _ = ops.gccb_sample(dev_c(T, U), dev_i(T, particles), S, 1234)
    out = out.cpu().numpy()
    states = oracle.generate_possible_states(n, m)
    freq = np.zeros(len(states))
    index = {s: i for i, s in enumerate(states)}
    for row in out:
        freq[index[tuple(int(x) for x in row)]] += 1
    freq /= S
    exact = np.array([oracle.fixed_losses_probability(U, [1, 1, 0, 0], st, oracle.perm_ryser)
                      for st in states])
    assert abs(exact.sum() - 1) < 1e-12
    tvd = 0.5 * np.abs(freq - exact).sum()
    assert tvd < oracle.tvd_error_bound(len(states), S, 1e-4)


def test_plain_strategy_equals_batched_with_same_draws(T, ops, oracle):
    from theboss_amd.permanent_calculators import GPURyserPermanentCalculator
    from theboss_amd.simulation_strategies import GeneralizedCliffordsBSimulationStrategy

    m, n, S = 16, 6, 12
    rng = np.random.default_rng(2)
    U = oracle.haar_unitary(m, 77)
    inp = [1] * n + [0] * (m - n)
    order, unif = _forced(rng, S, n)
    plain = GeneralizedCliffordsBSimulationStrategy(GPURyserPermanentCalculator(U)).simulate_with_draws(
        inp, order, unif)
    out, _ = ops.gccb_sample(dev_c(T, U), dev_i(T, np.arange(n)), S, 1, 1.0, dev_i(T, order),
                             T.from_numpy(unif).cuda(), None)
    assert np.array_equal(out.cpu().numpy(), plain)


def test_batched_sampler_minors_equal_separate_sweep(T, ops, oracle):
    """Steps 1..12 run fused in one launch per sample, later steps as sweep + step
    launches: the traced minors of steps 13.. equal, bit for bit, tb_sweep_minors_batch on
    the same columns (same kernel, chunks, wave sums and summation order); those of the
    fused steps equal it to 1e-12 relative (at k = 11, 12 the separate launch runs the
    paired walk, whose rounding differs from the fused plain walk)."""
    m, n, S = 400, 16, 24
    rng = np.random.default_rng(16)
    U = oracle.haar_unitary(m, 4016)
    order, unif = _forced(rng, S, n)
    out, trace = ops.gccb_sample(dev_c(T, U), dev_i(T, np.arange(n)), S, 1, 1.0, dev_i(T, order),
                                 T.from_numpy(unif).cuda(), None, trace=True)
    out, tr = out.cpu().numpy(), trace.cpu().numpy()
    assert np.array_equal(out, oracle.gccb_sample_batch(U, order, unif))
    # The picks, from the device's own minors (Laplace expansion and inverse CDF).
    picks = np.zeros((S, n), dtype=np.int64)
    for s in range(S):
        for k in range(1, n + 1):
            amp = U[:, order[s, :k]] @ tr[s, k - 1, :k]
            cum = np.cumsum(np.abs(amp) ** 2)
            picks[s, k - 1] = min(int(np.searchsorted(cum, unif[s, k - 1] * cum[-1], side="right")),
                                  m - 1)
    assert np.array_equal(np.stack([np.bincount(p, minlength=m) for p in picks]), out)
    for k in range(2, n + 1):
        Vt = np.stack([U[np.ix_(picks[s, :k - 1], order[s, :k])].T for s in range(S)])
        got = ops.sweep_minors_batch(dev_c(T, Vt), k).cpu().numpy()
        if k > 12:
            assert np.array_equal(got, tr[:, k - 1, :k]), k
        else:
            want = tr[:, k - 1, :k]
            assert np.max(np.abs(got - want) / np.max(np.abs(want), axis=1, keepdims=True)) <= 1e-12, k


def test_errors_surface_as_exceptions(T, ops):
    from theboss_amd._lib import TbError

    A = T.zeros((1, 49, 49), dtype=T.complex128, device="cuda")
    with pytest.raises(TbError):
        ops.glynn_batch(A)


def test_batched_sampler_bench_shape_minors_vs_oracle(T, ops, oracle):
    """m = 400, n = 20 (the benchmark workload) with forced draws: the samples equal the
    oracle sampler's, and the traced minors of the last steps equal the oracle-grade
    minors of the same (input, output) prefix at <= 1e-10 relative error, as do the step
    amplitudes perm(U[r + e_o, s]) obtained by the Laplace expansion."""
    m, n, S = 400, 20, 3
    rng = np.random.default_rng(400)
    U = oracle.haar_unitary(m, 20240611)
    order, unif = _forced(rng, S, n)
    out, trace = ops.gccb_sample(dev_c(T, U), dev_i(T, np.arange(n)), S, 1, 1.0, dev_i(T, order),
                                 T.from_numpy(unif).cuda(), None, trace=True)
    out, tr = out.cpu().numpy(), trace.cpu().numpy()
    expect = oracle.gccb_sample_batch(U, order, unif)
    assert np.array_equal(out, expect)
    s_idx = 0
    cur = np.zeros(m, int)
    r = np.zeros(m, int)
    for k in range(1, n + 1):
        cur[order[s_idx, k - 1]] += 1
        if k in (n - 1, n):
            ref = oracle.cc_submatrices_glynn(U, cur, r)
            got = tr[s_idx, k - 1, :k]
            want = np.array([ref[order[s_idx, nu]] for nu in range(k)])
            assert np.max(np.abs(got - want) / np.abs(want)) <= REL, k
            o = int(np.argmax(np.abs(U[:, :n] @ (cur[:n] * ref[:n]))))
            amp = np.sum(U[o, order[s_idx, :k]] * got)
            to = r.copy()
            to[o] += 1
            ref_amp = oracle.perm_glynn(U, cur, to, 1)
            assert abs(amp - ref_amp) <= REL * abs(ref_amp), k
        pmf, _ = oracle.gccb_pmf(U, cur, r)
        r[int(np.searchsorted(np.cumsum(pmf), unif[s_idx, k - 1], side="right"))] += 1
