"""GPU brute force against the oracle: keys and distances bit-identical to
the crate's BruteForceIndex (AVX2 kernels, ties broken by key); the FP64
oracle kernel against the CPU FP64 restatement."""
import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import same_bits

pytestmark = pytest.mark.gpu


def test_reference_known_answers():
    """src/brute.rs:218-262"""
    idx = H.BruteForceIndex(1)
    idx.insert_batch([5, 3, 9, 1], np.array([5.0, 3.0, 9.0, 1.0]))
    hits = idx.search([4.0], 3)
    assert [h.key for h in hits] == [3, 5, 1] and hits[0].distance == 1.0
    assert idx.search([4.0], 0) == []
    assert len(idx.search([4.0], 100)) == 4
    assert len(idx.search([4.0], 1 << 60)) == 4
    assert len(idx.search_filtered([4.0], 1 << 60, lambda k: k > 4)) == 2
    assert [h.key for h in idx.search_filtered([4.0], 2, lambda k: k > 4)] == [5, 9]
    b = idx.search_batch(np.array([0.0, 10.0]), 1)
    assert (b[0][0].key, b[1][0].key) == (1, 9)
    t = H.BruteForceIndex(1)
    t.insert_batch([7, 2, 4], np.array([1.0, -1.0, 1.0]))
    assert [h.key for h in t.search([0.0], 3)] == [2, 4, 7]
    e = H.BruteForceIndex(2, H.Metric.L2)
    e.insert(10, [0.0, 0.0])
    e.insert(20, [3.0, 4.0])
    h = e.search([2.5, 4.0], 1)
    assert h[0].key == 20 and h[0].distance == 0.25


@pytest.mark.parametrize("metric", [O.L2, O.IP, O.COSINE])
@pytest.mark.parametrize("dim", [128, 96, 100, 37, 3, 960])
def test_brute_matches_oracle_bit_exact(metric, dim):
    print(f"case metric={metric} dim={dim}", flush=True)
    b, q = O.dataset_gaussian(1500, 24, dim, dim)
    keys = (np.arange(1500, dtype=np.uint64) * 7 + 3)[::-1].copy()
    idx = H.BruteForceIndex(dim, H.Metric(metric))
    idx.insert_batch(keys, b)
    prepared = np.stack([O.prepare(metric, dim, r)[1] for r in b])
    for k in (1, 10, 100):
        gk, gd, gc = idx.search_arrays(q, k)
        for i in range(len(q)):
            ok, od = O.brute_search(metric, keys, prepared, q[i], k)
            assert gc[i] == len(ok)
            assert np.array_equal(gk[i, :gc[i]], ok), (k, i)
            assert same_bits(gd[i, :gc[i]], od), (k, i)


def test_ties_break_by_key_at_scale():
    """Duplicate rows give equal distances; order must follow the keys."""
    rng = np.random.default_rng(3)
    base = rng.standard_normal((50, 16)).astype(np.float32)
    rows = np.concatenate([base] * 8)
    keys = rng.permutation(400).astype(np.uint64) + 1000
    idx = H.BruteForceIndex(16)
    idx.insert_batch(keys, rows)
    q = rng.standard_normal((10, 16)).astype(np.float32)
    gk, gd, gc = idx.search_arrays(q, 37)
    for i in range(10):
        ok, od = O.brute_search(O.L2, keys, rows, q[i], 37)
        assert np.array_equal(gk[i], ok) and same_bits(gd[i], od)


def test_f64_oracle_kernel_matches_cpu_f64():
    b, q = O.dataset_gaussian(3000, 16, 64, 5)
    keys = np.arange(3000, dtype=np.uint64)
    idx = H.BruteForceIndex(64)
    idx.insert_batch(keys, b)
    gk, gd, gc = idx.search_arrays(q, 10, f64=True)
    for i in range(len(q)):
        ok, od = O.brute_search_f64(O.L2, keys, b, q[i], 10)
        assert np.array_equal(gk[i], ok)
        np.testing.assert_allclose(gd[i], od, rtol=1e-12)


def test_invalid_query_errors():
    idx = H.BruteForceIndex(2, H.Metric.Cosine)
    idx.insert_batch([1, 2], np.array([1.0, 0.0, 0.0, 1.0]))
    with pytest.raises(H.ZeroVector):
        idx.search([0.0, 0.0], 1)
    with pytest.raises(H.DuplicateKey):
        idx.insert_batch([5, 5], np.array([1.0, 0.0, 0.0, 1.0]))
    assert len(idx) == 2
    with pytest.raises(H.InvalidDimension):
        H.BruteForceIndex(0)


# ---------------------------------------------------------------- MFMA path
@pytest.mark.parametrize("metric", [O.L2, O.IP, O.COSINE])
@pytest.mark.parametrize("dim", [128, 100, 64, 37, 3])
def test_mfma_shortlist_matches_oracle(metric, dim):
    """The MFMA shortlist (approximate GEMM filter, exact re-rank, proof or
    exact re-run per query) returns exactly the oracle's lists."""
    print(f"case mfma metric={metric} dim={dim}", flush=True)
    b, q = O.dataset_gaussian(4096, 70, dim, 11 + dim)
    keys = (np.arange(4096, dtype=np.uint64) * 5 + 1)[::-1].copy()
    idx = H.BruteForceIndex(dim, H.Metric(metric))
    idx.insert_batch(keys, b)
    idx.set_mfma(2)
    prepared = np.stack([O.prepare(metric, dim, r)[1] for r in b])
    for k in (1, 10, 64):
        gk, gd, gc = idx.search_arrays(q, k)
        used, redo = idx.mfma_stats()
        assert used == len(q), (k, used)
        ok, od, oc = O.brute_search_batch(metric, keys, prepared, q, k)
        assert np.array_equal(gc, oc)
        for i in range(len(q)):
            assert np.array_equal(gk[i, :gc[i]], ok[i, :oc[i]]), (k, i)
            assert same_bits(gd[i, :gc[i]], od[i, :oc[i]].astype(np.float32)), (k, i)


def test_mfma_auto_on_clustered_rows():
    """Automatic mode takes the MFMA path at n >= 32768 and proves almost
    every query exact without the streaming re-run."""
    print("case mfma auto clustered 40000x128", flush=True)
    b, q = O.dataset_clustered(40000, 100, 128, 50, 9)
    keys = np.arange(40000, dtype=np.uint64) + 7
    idx = H.BruteForceIndex(128)
    idx.insert_batch(keys, b)
    gk, gd, gc = idx.search_arrays(q, 10)
    used, redo = idx.mfma_stats()
    assert used == 100 and redo <= 10, (used, redo)
    ok, od, oc = O.brute_search_batch(O.L2, keys, b, q, 10)
    assert np.array_equal(gk, ok)
    assert same_bits(gd, od.astype(np.float32))


def test_mfma_duplicate_rows_rerun_exactly():
    """Many equal distances defeat the exactness proof; those queries are
    re-run by the streaming kernel and the ties still break by key."""
    rng = np.random.default_rng(4)
    base = rng.standard_normal((8, 32)).astype(np.float32)
    rows = np.concatenate([base] * 512)
    keys = rng.permutation(rows.shape[0]).astype(np.uint64) + 100
    idx = H.BruteForceIndex(32)
    idx.insert_batch(keys, rows)
    idx.set_mfma(2)
    q = rng.standard_normal((40, 32)).astype(np.float32)
    gk, gd, gc = idx.search_arrays(q, 50)
    used, redo = idx.mfma_stats()
    assert used == 40
    ok, od, oc = O.brute_search_batch(O.L2, keys, rows, q, 50)
    assert np.array_equal(gk, ok)
    assert same_bits(gd, od.astype(np.float32))


def test_mfma_not_used_for_large_k_filters_and_wide_rows():
    b, q = O.dataset_gaussian(2000, 40, 64, 3)
    keys = np.arange(2000, dtype=np.uint64)
    idx = H.BruteForceIndex(64)
    idx.insert_batch(keys, b)
    idx.set_mfma(2)
    idx.search_arrays(q, 65)
    assert idx.mfma_stats() == (0, 0)
    gk, gd, gc = idx.search_arrays(q, 10, filter=lambda key: key % 3 == 0)
    assert idx.mfma_stats() == (0, 0)
    assert all(int(x) % 3 == 0 for x in gk.ravel())
    idx.set_mfma(0)
    e = idx.search_arrays(q, 10)
    assert idx.mfma_stats() == (0, 0)
    idx.set_mfma(2)
    m = idx.search_arrays(q, 10)
    assert idx.mfma_stats()[0] == 40
    assert np.array_equal(e[0], m[0]) and same_bits(e[1], m[1])
    w = H.BruteForceIndex(160)
    bw, qw = O.dataset_gaussian(500, 40, 160, 2)
    w.insert_batch(np.arange(500, dtype=np.uint64), bw)
    w.set_mfma(2)
    w.search_arrays(qw, 10)
    assert w.mfma_stats() == (0, 0)


def test_mfma_invalid_query_in_batch():
    b, q = O.dataset_gaussian(1000, 64, 16, 8)
    idx = H.BruteForceIndex(16)
    idx.insert_batch(np.arange(1000, dtype=np.uint64), b)
    idx.set_mfma(2)
    q[17, 3] = np.nan
    with pytest.raises(H.NonFiniteVector):
        idx.search_arrays(q, 5)
    q[17, 3] = 0.0
    gk, gd, gc = idx.search_arrays(q, 5)
    ok, od, oc = O.brute_search_batch(O.L2, np.arange(1000, dtype=np.uint64), b, q, 5)
    assert np.array_equal(gk, ok)
