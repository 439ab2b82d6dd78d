"""Host side of the engine (CPU only): the C ABI exports, the CPU graph
construction (the crate's insert / insert_batch restated), the host
quantizer and generators, each checked against the oracle or the crate's
own published figures."""
import ctypes
import json
import math
import os
import re

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import build, oracle_graph, same_bits

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KAT = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "hnsw_amd.h")).read()
    declared = set(re.findall(r"\b(hh_[a-z0-9_]+)\s*\(", header))
    assert len(declared) > 30
    lib = ctypes.CDLL(os.path.join(ROOT, "hnsw-index_amd", "libhnsw_amd.so"))
    missing = [s for s in sorted(declared) if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.hh_abi_version() == 1


def test_device_count_is_not_an_error_without_a_gpu():
    assert H.lib.hh_device_count() >= 0


def test_host_kernels_are_the_avx2_sequence_bit_for_bit():
    """hh_cpu_distance runs the crate's AVX2 intrinsic sequence
    (src/kernels/x86.rs); the oracle restates it lane by lane in C."""
    rng = np.random.default_rng(5)
    for n in list(range(0, 70)) + [96, 100, 128, 255, 256, 960]:
        for _ in range(3):
            a = rng.standard_normal(n).astype(np.float32) * 3
            b = rng.standard_normal(n).astype(np.float32) * 3
            for metric in (0, 1, 2):
                host = H.lib.hh_cpu_distance(metric, a.ctypes.data, b.ctypes.data, n)
                want = O.distance(metric, a, b, O.AVX2)
                assert same_bits([host], [want]), (n, metric)


def test_generators_match_the_oracle():
    b, q = O.dataset_gaussian(300, 20, 17, 42)
    d = H.synthetic.Dataset.gaussian(300, 20, 17, 42)
    assert np.array_equal(d.base(), b) and np.array_equal(d.queries(), q)
    b, q = O.dataset_clustered(300, 20, 9, 7, 3)
    d = H.synthetic.Dataset.clustered(300, 20, 9, 7, 3)
    assert np.array_equal(d.base(), b) and np.array_equal(d.queries(), q)


def test_benchmark_generators_have_the_stated_distributions():
    d = H.synthetic.Dataset.uniform_clusters(20000, 2000, 32, 50, 0.1, 0.12, 42, threads=4)
    again = H.synthetic.Dataset.uniform_clusters(20000, 2000, 32, 50, 0.1, 0.12, 42, threads=1)
    assert np.array_equal(d.base(), again.base()), "independent of the thread count"
    b = d.base().astype(np.float64)
    assert -1.5 < b.min() and b.max() < 1.5
    # nearest of 50 centres is within ~0.1*sqrt(32) of every row
    u = H.synthetic.Dataset.unit_uniform(1000, 10, 960, 1)
    assert np.allclose(np.linalg.norm(u.base().astype(np.float64), axis=1), 1.0, atol=1e-5)
    assert (u.base() >= 0).all()
    ln = H.synthetic.Dataset.lognormal(5000, 10, 64, 0.5, 20, 9)
    norms = np.linalg.norm(ln.base().astype(np.float64), axis=1) / 8.0
    assert abs(np.log(norms).std() - 0.5) < 0.05
    assert set(np.unique(ln.categories)) == set(range(20))


def test_host_quantizer_matches_the_oracle_bit_for_bit():
    for metric in (H.Metric.L2, H.Metric.InnerProduct, H.Metric.Cosine):
        b, _ = O.dataset_gaussian(500, 1, 24, 4)
        idx = build(b, metric, threads=2)
        idx.enable_quantization()
        v = idx.graph_arrays()
        off, scale = O.sq_train(int(metric), v["vectors"])
        assert np.array_equal(v["sq_offset"], off) and v["sq_scale"] == scale
        assert np.array_equal(v["codes"], O.sq_encode_rows(off, scale, v["vectors"]))


def test_installed_quantizer_is_validated_and_encodes_like_the_oracle():
    """ScalarQuantizer::from_parts validity (src/quantization.rs:58-62) and
    encode_into of every stored vector when no codes are given."""
    import ctypes
    from hnsw_amd import _lib as L
    b, _ = O.dataset_gaussian(300, 1, 20, 8)
    idx = build(b, H.Metric.L2, threads=2)
    off, scale = O.sq_train(int(H.Metric.L2), idx.graph_arrays()["vectors"])
    g = idx._g
    for bad_off, bad_scale in [(off, 0.0), (off, -1.0), (off, float("inf")), (off, float("nan")),
                               (np.where(np.arange(20) == 3, np.inf, off).astype(np.float32), scale)]:
        assert L.lib.hh_graph_set_quantizer(g, bad_off.ctypes.data, bad_scale, None) == \
            H.InvalidParameter.status
        assert not idx.is_quantized()
    assert L.lib.hh_graph_set_quantizer(g, off.ctypes.data, float(scale), None) == 0
    v = idx.graph_arrays()
    assert np.array_equal(v["sq_offset"], off) and np.float32(v["sq_scale"]) == scale
    assert np.array_equal(v["codes"], O.sq_encode_rows(off, scale, v["vectors"]))


def _assert_invariants(v):
    """src/hnsw/tests.rs:28-53"""
    n, m, m0 = v["n"], v["m"], v["m0"]
    lay0 = v["layer0"].reshape(n, m0 + 1)
    for node in range(n):
        level = int(v["levels"][node])
        for layer in range(level + 1):
            if layer == 0:
                cnt = min(int(lay0[node, 0]), m0)
                nb = lay0[node, 1:1 + cnt]
                cap = m0
            else:
                o = int(v["upper_off"][node]) + (layer - 1) * (m + 1)
                cnt = min(int(v["upper"][o]), m)
                nb = v["upper"][o + 1:o + 1 + cnt]
                cap = m
            assert cnt <= cap
            assert (nb < n).all(), "dangling id"
            assert (nb != node).all(), "self loop"
            assert (v["levels"][nb] >= layer).all()
            assert len(set(nb.tolist())) == len(nb), "duplicate neighbour"
    if n:
        assert int(v["levels"][v["entry_id"]]) == v["top"] == int(v["levels"].max())


def _reachable(v):
    """src/hnsw/tests.rs:57-79"""
    n, m0 = v["n"], v["m0"]
    lay0 = v["layer0"].reshape(n, m0 + 1)
    seen = np.zeros(n, bool)
    queue = [i for i in range(n) if i == v["entry_id"] or v["levels"][i] > 0]
    seen[queue] = True
    while queue:
        nxt = []
        for node in queue:
            for nb in lay0[node, 1:1 + min(int(lay0[node, 0]), m0)]:
                if not seen[nb]:
                    seen[nb] = True
                    nxt.append(int(nb))
        queue = nxt
    return int(seen.sum())


@pytest.mark.parametrize("selection", [H.NeighborSelection.simple(), H.NeighborSelection(),
                                       H.NeighborSelection.heuristic_with(True, True)])
@pytest.mark.parametrize("threads", [1, 4])
def test_graphs_satisfy_invariants_and_stay_connected(selection, threads):
    """src/hnsw/tests.rs:137-166"""
    b, _ = O.dataset_clustered(3000, 1, 16, 20, 5)
    idx = build(b, params=H.HnswParams(8).with_selection(selection), threads=threads)
    v = idx.graph_arrays()
    _assert_invariants(v)
    if selection.heuristic:
        assert _reachable(v) >= 2970
    deg = np.minimum(v["layer0"].reshape(3000, 17)[:, 0], 16).mean()
    assert 4.0 < deg <= 16.0


def test_sequential_build_is_deterministic():
    """src/hnsw/tests.rs:129-135"""
    b, _ = O.dataset_gaussian(500, 10, 8, 1)
    ia, ic = build(b), build(b)  # graph_arrays() borrows from the live index
    a, c = ia.graph_arrays(), ic.graph_arrays()
    for key in ("layer0", "upper", "upper_off", "levels", "keys"):
        assert np.array_equal(a[key], c[key])
    assert a["entry_id"] == c["entry_id"]


def test_levels_follow_the_crate_hash_and_distribution():
    """sample_level (src/hnsw/build.rs:19-28) and its tests (:229-275)."""
    b = np.zeros((20000, 1), np.float32) + np.arange(20000, dtype=np.float32)[:, None]
    idx = build(b, params=H.HnswParams(16).with_seed(42), threads=8)
    levels = idx.graph_arrays()["levels"]
    ml = 1.0 / math.log(16)
    for i in (0, 1, 2, 12345, 19999):
        h = O.mix64(42 ^ O.mix64((i + 0x9E3779B97F4A7C15) % 2**64))
        u = ((h >> 11) + 1) * 2.0**-53
        assert levels[i] == min(31, math.floor(-math.log(u) * ml))
    for lv in (1, 2):
        expected = 20000 * 16.0**-lv
        assert abs((levels >= lv).sum() - expected) < 5 * math.sqrt(expected) + 2


def test_parameter_validation():
    """src/params.rs:204-223, src/hnsw/tests.rs:97-111"""
    for p in [H.HnswParams(1), H.HnswParams(513), H.HnswParams(16).with_m0(15),
              H.HnswParams(16).with_m0(4096), H.HnswParams(16).with_ef_construction(0),
              H.HnswParams(16).with_level_multiplier(0.0),
              H.HnswParams(16).with_level_multiplier(float("nan")),
              H.HnswParams(16).with_level_multiplier(11.0)]:
        with pytest.raises(H.InvalidParameter):
            H.HnswIndex(4, H.Metric.L2, p)
    with pytest.raises(H.InvalidDimension):
        H.HnswIndex(0)
    with pytest.raises(H.InvalidDimension):
        H.HnswIndex(65537)


def test_insert_rejects_invalid_input_without_side_effects():
    """src/hnsw/tests.rs:168-206"""
    idx = H.HnswIndex(2, H.Metric.Cosine)
    idx.insert(1, [1.0, 0.0])
    with pytest.raises(H.DuplicateKey):
        idx.insert(1, [0.0, 1.0])
    with pytest.raises(H.DimensionMismatch):
        idx.insert(2, [1.0])
    with pytest.raises(H.NonFiniteVector):
        idx.insert(2, [np.nan, 1.0])
    with pytest.raises(H.ZeroVector):
        idx.insert(2, [0.0, 0.0])
    with pytest.raises(H.Error):
        idx.insert_batch([2, 3], np.array([1.0, 1.0, 0.0, 0.0]))
    with pytest.raises(H.DuplicateKey):
        idx.insert_batch([2, 2], np.array([1.0, 1.0, 1.0, 2.0]))
    assert len(idx) == 1 and idx.node_count() == 1


def test_recall_table_is_reproduced_exactly():
    """docs/benchmarks.md:162-169: sequential builds on gaussian 10k x 128,
    200 queries, k=10 — the oracle's search on the host-built graph gives
    the published figure at every ef."""
    k = KAT["recall_table_gaussian_10k"]
    base, queries = O.dataset_gaussian(*k["data"])
    tk, _, tc = O.brute_search_batch(O.L2, np.arange(10000), base, queries, 10)
    for m, want in (("m16", k["m16"]), ("m32", k["m32"])):
        g = oracle_graph(build(base, m=int(m[1:]), threads=1))
        got = []
        for ef in k["efs"]:
            fk, _, fc, _ = g.search_batch(queries, 10, ef)
            got.append(O.recall_at_k(tk, tc, fk, fc, 10))
        # 200 queries x k=10: recall is a multiple of 1/2000, published to
        # four decimals, so equality after rounding is exact equality
        assert np.round(got, 4).tolist() == want, (m, got)


def _table_100k(name, make):
    """docs/benchmarks.md:107-121: recall@10 of the sequentially built M=16
    index over 1,000 queries (a multiple of 1/10000, published to four
    decimals), f32 traversal and int8 traversal + exact re-rank on that same
    index.  This pins the generator (draw order, Lemire index, Box-Muller),
    the AVX2-order kernels, the sequential build, the search, the
    quantizer's training, encoding and approximate distance, and the
    re-rank together: each of the 10 figures is reproduced exactly."""
    k = KAT[name]
    base, queries = make(*k["data"])
    n = base.shape[0]
    tk, _, tc = O.brute_search_batch(O.L2, np.arange(n), base, queries, 10)
    idx = build(base, m=16, threads=1)
    for col, quantized in (("f32", False), ("int8", True)):
        if quantized:
            idx.enable_quantization()
        g = oracle_graph(idx)
        got = []
        for ef in k["efs"]:
            fk, _, fc, _ = g.search_batch(queries, 10, ef, quantized=quantized)
            got.append(O.recall_at_k(tk, tc, fk, fc, 10))
        assert np.round(got, 4).tolist() == k[col], (name, col, got)


def test_clustered_100k_table_is_reproduced_exactly():
    _table_100k("recall_table_clustered_100k", O.dataset_clustered)


def test_gaussian_100k_table_is_reproduced_exactly():
    _table_100k("recall_table_gaussian_100k", O.dataset_gaussian)


def test_recall_floors_of_the_crate():
    """tests/recall.rs:74-128 on host-built graphs, searched by the oracle."""
    f = KAT["recall_floors"]
    base, queries = O.dataset_clustered(10000, 200, 128, 50, 42)
    tk, _, tc = O.brute_search_batch(O.L2, np.arange(10000), base, queries, 10)
    g = oracle_graph(build(base, threads=1))
    fk, _, fc, _ = g.search_batch(queries, 10, f["ef"])
    assert O.recall_at_k(tk, tc, fk, fc, 10) >= f["clustered50_m16"]
    base, queries = O.dataset_gaussian(5000, 100, 64, 7)
    for metric in (O.COSINE, O.IP):
        prepared = np.stack([O.prepare(metric, 64, r)[1] for r in base])
        tk, _, tc = O.brute_search_batch(metric, np.arange(5000), prepared, queries, 10)
        g = oracle_graph(build(base, H.Metric(metric), threads=1))
        fk, _, fc, _ = g.search_batch(queries, 10, f["ef"])
        assert O.recall_at_k(tk, tc, fk, fc, 10) >= f["cosine_ip_5k_64"]


def test_parallel_build_and_int8_recall_floors():
    """tests/recall.rs:98-128"""
    f = KAT["recall_floors"]
    base, queries = O.dataset_gaussian(10000, 200, 128, 42)
    tk, _, tc = O.brute_search_batch(O.L2, np.arange(10000), base, queries, 10)
    idx = build(base, m=32, threads=4)
    g = oracle_graph(idx)
    fk, _, fc, _ = g.search_batch(queries, 10, f["ef"])
    assert O.recall_at_k(tk, tc, fk, fc, 10) >= f["parallel4_m32"]
    idx.enable_quantization()
    g = oracle_graph(idx)
    fk, _, fc, _ = g.search_batch(queries, 10, f["ef"], quantized=True)
    assert O.recall_at_k(tk, tc, fk, fc, 10) >= f["int8_rerank_m32"]


def test_oracle_search_edge_cases_on_host_graphs():
    """src/hnsw/tests.rs:113-127 and :208-218 with the oracle as searcher."""
    b, q = O.dataset_gaussian(16, 20, 8, 3)
    g = oracle_graph(build(b, params=H.HnswParams(16).with_selection(H.NeighborSelection.simple())))
    for i in range(20):
        ok, od, _, _ = g.search(q[i], 5, 5)
        bk, bd = O.brute_search(O.L2, np.arange(16), b, q[i], 5)
        assert np.array_equal(ok, bk) and same_bits(od, bd)
    b, q = O.dataset_gaussian(100, 1, 4, 2)
    g = oracle_graph(build(b))
    assert len(g.search(q[0], 0, 10)[0]) == 0
    assert len(g.search(q[0], 10, 1)[0]) == 10
    assert len(g.search(q[0], 500, 500)[0]) == 100


def test_search_needs_a_gpu_and_never_falls_back_to_the_cpu():
    if H.lib.hh_device_count() > 0:
        pytest.skip("a GPU is visible")
    b, q = O.dataset_gaussian(50, 1, 4, 2)
    idx = build(b)
    with pytest.raises(H.DeviceError):
        idx.search(q[0], 5, 10)


def test_take_dirty_reports_every_changed_link_list():
    """hh_graph_take_dirty: every node whose layer-0 or upper block changed
    since the last call is reported (the resident index relies on it)."""
    import ctypes
    from hnsw_amd import _lib as L
    b, _ = O.dataset_gaussian(2600, 1, 16, 19)
    idx = build(b[:2000], H.Metric.L2, m=8, threads=4)
    first = idx._take_dirty()
    assert first.size > 0 and np.all(np.diff(first.astype(np.int64)) > 0)
    assert idx._take_dirty().size == 0  # cleared
    v0 = {k: np.array(x, copy=True) for k, x in idx.graph_arrays().items()
          if isinstance(x, np.ndarray)}
    idx.insert_batch(np.arange(2000, 2600, dtype=np.uint64), b[2000:], threads=4)
    dirty = set(idx._take_dirty().tolist())
    v1 = idx.graph_arrays()
    m0 = v1["m0"]
    l0a = v0["layer0"].reshape(2000, m0 + 1)
    l0b = v1["layer0"].reshape(-1, m0 + 1)[:2000]
    changed = set(np.nonzero((l0a != l0b).any(axis=1))[0].tolist())
    m = v1["m"]
    for node in range(2000):
        lv = int(v1["levels"][node])
        if lv:
            o = int(v1["upper_off"][node])
            if not np.array_equal(v0["upper"][o:o + lv * (m + 1)], v1["upper"][o:o + lv * (m + 1)]):
                changed.add(node)
    assert changed and changed <= dirty
    assert set(range(2000, 2600)) <= dirty
    # a partial take returns the smallest ids and clears only those
    idx.insert_batch(np.arange(5000, 5010, dtype=np.uint64), b[:10] + 0.5, threads=1)
    cnt = ctypes.c_uint64()
    ids = np.zeros(3, np.uint32)
    assert L.lib.hh_graph_take_dirty(idx._g, ids.ctypes.data, 3, ctypes.byref(cnt)) == 0
    rest = idx._take_dirty()
    assert cnt.value == 3 + rest.size and ids.max() < rest.min()


def test_append_then_link_equals_insert_batch():
    """hh_graph_append + hh_graph_link_batch without candidates is the
    host INSERT: sequentially it builds the same graph as insert_batch."""
    import ctypes
    from hnsw_amd import _lib as L
    b, _ = O.dataset_gaussian(1500, 1, 20, 23)
    a = build(b, H.Metric.L2, m=8, threads=1)
    c = H.HnswIndex(20, H.Metric.L2, H.HnswParams(8))
    keys = np.arange(1500, dtype=np.uint64)
    first, bad = ctypes.c_uint64(), ctypes.c_uint64()
    assert L.lib.hh_graph_append(c._g, keys.ctypes.data, b.ctypes.data, 1500, ctypes.byref(first),
                                 ctypes.byref(bad)) == 0 and first.value == 0
    assert L.lib.hh_graph_link_batch(c._g, 0, 700, None, None, None, 0, 1) == 0
    assert L.lib.hh_graph_link_batch(c._g, 700, 800, None, None, None, 0, 1) == 0
    va, vc = a.graph_arrays(), c.graph_arrays()
    for f in ("levels", "layer0", "upper_off", "upper", "keys"):
        assert np.array_equal(va[f], vc[f]), f
    assert va["entry_id"] == vc["entry_id"] and va["top"] == vc["top"]
    # append validates like insert_batch: nothing is stored on error
    dup = np.array([5, 9999], np.uint64)
    assert L.lib.hh_graph_append(c._g, dup.ctypes.data, b[:2].ctypes.data, 2, ctypes.byref(first),
                                 ctypes.byref(bad)) == H.DuplicateKey.status
    assert bad.value == 5 and c.node_count() == 1500
    assert L.lib.hh_graph_link_batch(c._g, 1400, 200, None, None, None, 0, 1) == \
        H.InvalidArgument.status


def _hh_recall(truth, tc, found, fc, k):
    truth = np.ascontiguousarray(truth, np.uint64)
    found = np.ascontiguousarray(found, np.uint64)
    tc = np.ascontiguousarray(tc, np.uint32)
    fc = np.ascontiguousarray(fc, np.uint32)
    return H.lib.hh_recall_at_k(truth.shape[0], truth.ctypes.data, tc.ctypes.data,
                                truth.shape[1] if truth.ndim == 2 else 1, found.ctypes.data,
                                fc.ctypes.data, found.shape[1] if found.ndim == 2 else 1, k)


def test_hh_recall_at_k_known_answers_and_oracle():
    """hh_recall_at_k (the bench's recall) against src/eval.rs:59-72 and
    against the oracle's restatement on ragged random result lists."""
    k = KAT["recall_at_k"]
    t = np.array(k["truth"], np.uint64)
    f = np.array(k["found"], np.uint64)
    c = np.array([3, 3], np.uint32)
    assert abs(_hh_recall(t, c, f, c, 3) - k["k3"]) < 1e-12
    assert _hh_recall(t, c, f, c, 1) == k["k1"]
    one = np.array([[1]], np.uint64)
    assert _hh_recall(one, [1], one, [1], 10) == 1.0
    assert _hh_recall(np.zeros((0, 1), np.uint64), np.zeros(0), np.zeros((0, 1), np.uint64),
                      np.zeros(0), 10) == 1.0
    assert _hh_recall(np.array([[1, 2]], np.uint64), [2], np.zeros((1, 2), np.uint64), [0], 2) == 0.0
    rng = np.random.default_rng(3)
    for trial in range(50):
        nq, stride, kk = int(rng.integers(1, 40)), int(rng.integers(1, 16)), int(rng.integers(1, 20))
        truth = np.stack([rng.permutation(60)[:stride] for _ in range(nq)]).astype(np.uint64)
        found = np.stack([rng.permutation(60)[:stride] for _ in range(nq)]).astype(np.uint64)
        tc = rng.integers(0, stride + 1, nq).astype(np.uint32)
        fc = rng.integers(0, stride + 1, nq).astype(np.uint32)
        want = O.recall_at_k(truth, tc, found, fc, kk)
        # and from the definition, in Python
        hits = exp = 0
        for q in range(nq):
            tt = set(truth[q, :min(kk, tc[q])].tolist())
            exp += len(tt)
            hits += sum(1 for x in found[q, :min(kk, fc[q])].tolist() if x in tt)
        assert want == (1.0 if exp == 0 else hits / exp)
        assert _hh_recall(truth, tc, found, fc, kk) == want
