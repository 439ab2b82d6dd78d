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
