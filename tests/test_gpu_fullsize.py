"""The benchmark's own workload at full size: 1,000,000 x 128 clustered rows
(1,000 centres uniform in [-1,1]^128, sigma 0.1), L2, M=16, efC=200, 10,000
queries, k=10, ef=64 (the search_batch path of src/hnsw/mod.rs:520-536).

A sample of the queries is compared bit for bit with the oracle on the same
graph; the whole batch is checked through properties that do not depend on
the size: result order, counts, one pass against two, batch-order
invariance, int8/bf16 re-rank distances, and recall against the exact GPU
brute force.
"""
import os

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import oracle_graph, same_bits

pytestmark = pytest.mark.gpu

N, NQ, DIM, K, EF = 1_000_000, 10_000, 128, 10, 64
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def c2():
    print("case full-size build 1M x 128", flush=True)
    data = H.synthetic.Dataset.uniform_clusters(N, NQ, DIM, 1000, 0.1, 0.12, 42, threads=THREADS)
    idx = H.HnswIndex(DIM, H.Metric.L2, H.HnswParams(16))
    keys = np.arange(N, dtype=np.uint64)
    for s in range(0, N, 250_000):
        idx.insert_batch(keys[s:s + 250_000], data.base()[s:s + 250_000], threads=THREADS)
    return data, idx


def _sorted_by_distance_then_key(k, d, c):
    for i in range(k.shape[0]):
        n = int(c[i])
        dd, kk = d[i, :n].astype(np.float64), k[i, :n]
        for j in range(n - 1):
            if not (dd[j] < dd[j + 1] or (dd[j] == dd[j + 1] and kk[j] < kk[j + 1])):
                return False
    return True


def test_full_size_sample_matches_oracle(c2):
    data, idx = c2
    q = data.queries()
    gk, gd, gc = idx.search_arrays(q, K, EF)
    g = oracle_graph(idx)
    print("case oracle sample 400", flush=True)
    ok, od, oc, _ = g.search_batch(q[:400], K, EF, threads=THREADS)
    assert np.array_equal(gc[:400], oc) and np.array_equal(gk[:400], ok)
    assert same_bits(gd[:400], od)


def test_full_size_properties(c2):
    data, idx = c2
    q = data.queries()
    idx.set_query_order(True)
    k2, d2, c2_ = idx.search_arrays(q, K, EF)
    idx.set_query_order(False)
    k1, d1, c1 = idx.search_arrays(q, K, EF)
    idx.set_query_order(True)
    assert (c1 == K).all()
    assert np.array_equal(k1, k2) and same_bits(d1, d2) and np.array_equal(c1, c2_)
    assert _sorted_by_distance_then_key(k1, d1, c1)
    # each query's results do not depend on its batch or its position in it
    perm = np.random.default_rng(3).permutation(NQ)
    kp, dp, cp = idx.search_arrays(q[perm], K, EF)
    assert np.array_equal(kp, k1[perm]) and same_bits(dp, d1[perm])
    # recall against the exact brute force (the benchmark's operating point)
    bf = H.BruteForceIndex(DIM, H.Metric.L2)
    bf.insert_batch(np.arange(N, dtype=np.uint64), data.base())
    tk, td, tc = bf.search_arrays(q, K)
    rec = O.recall_at_k(tk, tc, k1, c1, K)
    print(f"case recall@10 {rec:.4f}", flush=True)
    assert rec >= 0.94
    # a returned distance is the exact one of its key: where the brute force
    # returned the same key, the bits agree
    both = tk == k1
    assert both.mean() > 0.5 and same_bits(td[both], d1[both])


def test_full_size_int8_and_bf16_rerank(c2):
    data, idx = c2
    q = data.queries()
    idx.enable_quantization(on_device=True)
    idx.enable_bf16()
    g = oracle_graph(idx)
    ok, od, oc, _ = g.search_batch(q[:300], K, EF, quantized=True, threads=THREADS)
    ik, idist, ic = idx.search_arrays(q, K, EF, precision=H.Precision.INT8)
    assert np.array_equal(ic[:300], oc) and np.array_equal(ik[:300], ok)
    assert same_bits(idist[:300], od)
    bk, bd, bc = idx.search_arrays(q, K, EF, precision=H.Precision.BF16)
    fk, fd, fc = idx.search_arrays(q, K, EF)
    for k_, d_, c_ in ((ik, idist, ic), (bk, bd, bc)):
        assert (c_ == K).all() and _sorted_by_distance_then_key(k_, d_, c_)
        # after the exact re-rank, a key found by both traversals has the f32 distance
        for i in range(0, NQ, 97):
            common = np.intersect1d(k_[i], fk[i])
            a = {int(x): y for x, y in zip(k_[i], d_[i])}
            b = {int(x): y for x, y in zip(fk[i], fd[i])}
            assert all(np.float32(a[int(x)]).view(np.uint32) == np.float32(b[int(x)]).view(np.uint32)
                       for x in common)
