"""The autotuner and the tuning table on an MI355X: any chunk size a tuned table can
hold gives the same permanents (to rounding) and the same samples (bit-exact with forced
draws) as the built-in plans, and the tuner fills the table from its candidates."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REL = 1e-12  # same sums in a different order


@pytest.fixture(scope="module")
def T():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from theboss_amd import _lib

    _lib.load()
    return torch


@pytest.fixture()
def tuning(T):
    from theboss_amd import autotune

    autotune.reset()
    yield autotune
    autotune.reset()


def dev_c(T, a):
    return T.from_numpy(np.ascontiguousarray(a, dtype=np.complex128)).cuda()


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


@pytest.mark.parametrize("k", [12, 20])
def test_sweep_same_minors_at_every_chunk_size(T, oracle, tuning, k):
    from theboss_amd import ops

    V = np.stack([oracle.haar_unitary(2 * k, 60 + b)[: k - 1, :k].T for b in range(64)])
    Vd = dev_c(T, V)
    base = ops.sweep_minors_batch(Vd, k).cpu().numpy()
    for c in (3, 6, 9, 13):
        tuning.set_entry("sweep", k, c, 64)
        got = ops.sweep_minors_batch(Vd, k).cpu().numpy()
        assert rel(got, base) <= REL, (k, c)


def test_glynn_same_permanents_at_every_chunk_size(T, oracle, tuning):
    from theboss_amd import ops

    n = 24
    As = np.stack([oracle.haar_unitary(2 * n, 70 + b)[:n, :n] for b in range(4)])
    A = dev_c(T, As)
    # |perm| ~ 1e-10 against O(1) terms at n = 24: the summation order moves the result
    # by ~1e-12 relative, so each chunk size is held to the oracle bar instead.
    ref = oracle.perm_glynn_batch(As, 1)
    for c in (0, 2, 5, 9, 12):
        tuning.set_entry("glynn", n, c, 4)
        got = ops.glynn_batch(A).cpu().numpy()
        assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-10, c


def test_grad_same_minors_at_every_chunk_size(T, oracle, tuning):
    from theboss_amd import ops

    n = 20
    A = dev_c(T, np.stack([oracle.haar_unitary(2 * n, 80 + b)[:n, :n] for b in range(8)]))
    p0, g0 = (x.cpu().numpy() for x in ops.perm_grad_batch(A))
    for c in (2, 3, 6, 9, 11):
        tuning.set_entry("grad", n, c, 8)
        p, g = (x.cpu().numpy() for x in ops.perm_grad_batch(A))
        assert rel(g, g0) <= REL, c
        # perm is the Laplace sum along row 0 of these minors, which cancels: its relative
        # spread over chunk sizes exceeds the minors' (1.0e-12 measured at c = 11)
        assert np.max(np.abs(p - p0) / np.abs(p0)) <= 10 * REL, c


def test_sampler_bit_exact_under_tuned_chunks(T, oracle, tuning):
    """Forced draws: the samples are integers, so a retuned (and re-captured) sampler
    graph must reproduce them exactly."""
    from theboss_amd import ops

    m, n, S = 40, 16, 256
    U = dev_c(T, oracle.haar_unitary(m, 90))
    parts = T.arange(n, dtype=T.int32, device="cuda")
    rng = np.random.default_rng(9)
    order = T.from_numpy(np.stack([rng.permutation(n) for _ in range(S)]).astype(np.int32)).cuda()
    unif = T.from_numpy(rng.random((S, n))).cuda()
    base = ops.gccb_sample(U, parts, S, 1, forced_order=order, forced_uniforms=unif)[0].cpu().numpy()
    for k in range(8, n + 1):
        tuning.set_entry("sweep", k, 4 + (k % 5), S)
    got = ops.gccb_sample(U, parts, S, 1, forced_order=order, forced_uniforms=unif)[0].cpu().numpy()
    assert np.array_equal(got, base)


def test_autotuner_fills_table_from_candidates(T, tuning):
    res = tuning.tune_sweep([12, 16], items=256, reps=2)
    assert set(res) == {12, 16}
    for k, r in res.items():
        assert r["chunk_bits"] in r["seconds"]
        assert tuning.get_entry("sweep", k, 256) == r["chunk_bits"]
        fastest = min(r["seconds"].values())
        # the built-in choice stands unless a candidate beats it by the margin
        assert r["seconds"][r["chunk_bits"]] <= fastest / (1.0 - tuning.MARGIN)
    g = tuning.tune_grad([16], batch=8, reps=1)
    assert tuning.get_entry("grad", 16, 8) == g[16]["chunk_bits"]
    bucket = tuning.batch_bucket(256)
    assert tuning.table()["sweep"] == {f"{k}/{bucket}": r["chunk_bits"]
                                       for k, r in res.items() if r["chunk_bits"]}
