"""GPU HNSW search against the oracle on the same graph.

The GPU kernel reproduces the crate's AVX2 summation order, so on graphs
without distance ties every query must return the oracle's keys and
bit-identical distances, and the kernel's distance-evaluation counters must
equal the oracle's (src/hnsw/search.rs:81-142 evaluated step by step).
"""
import ctypes

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import assert_same_results, build, oracle_graph, same_bits

pytestmark = pytest.mark.gpu


def _compare(idx, queries, k, ef, quantized=False, allow=None, precision=None):
    g = oracle_graph(idx)
    prec = precision if precision is not None else (H.Precision.INT8 if quantized else H.Precision.F32)
    filt = None if allow is None else allow.astype(bool)
    idx.search_stats(reset=True)
    gk, gd, gc = idx.search_arrays(queries, k, ef, precision=prec, filter=filt)
    stats = idx.search_stats()
    ok, od, oc, cnt = g.search_batch(queries, k, ef, quantized=quantized,
                                      allow=None if allow is None else allow.astype(np.uint8))
    return (gk, gd, gc), (ok, od, oc), stats, cnt


@pytest.mark.parametrize("metric", [H.Metric.L2, H.Metric.InnerProduct, H.Metric.Cosine])
@pytest.mark.parametrize("dim", [128, 96, 100, 37, 3, 256])
def test_search_matches_oracle_bit_exact(metric, dim):
    print(f"case metric={metric} dim={dim}", flush=True)
    b, q = O.dataset_gaussian(3000, 64, dim, 7 + dim)
    idx = build(b, metric)
    for k, ef in [(10, 10), (10, 64), (1, 16), (32, 200)]:
        (gk, gd, gc), (ok, od, oc), st, cnt = _compare(idx, q, k, ef)
        assert np.array_equal(gc, oc)
        assert np.array_equal(gk, ok), f"k={k} ef={ef}"
        assert same_bits(gd, od), f"k={k} ef={ef}"
        assert st["n_dist"] + st["n_upper"] == cnt.n_dist, (st, cnt.n_dist)
        assert st["n_expand"] == cnt.n_expand
        assert st["n_overflow"] == 0


def test_search_batch_equals_individual_searches():
    """tests/concurrency.rs:60-69"""
    b, q = O.dataset_gaussian(2000, 50, 16, 5)
    idx = build(b)
    batch = idx.search_batch(q, 10, 40)
    for i in range(len(q)):
        assert batch[i] == idx.search(q[i], 10, 40)
    with pytest.raises(H.DimensionMismatch):
        idx.search_batch(q.ravel()[:15], 10, 40)


def test_quantized_search_matches_oracle():
    """int8 traversal + exact re-rank (src/hnsw/mod.rs:483-500)."""
    for metric in (H.Metric.L2, H.Metric.InnerProduct, H.Metric.Cosine):
        b, q = O.dataset_gaussian(2000, 40, 32, 12)
        idx = build(b, metric)
        with pytest.raises(H.QuantizationDisabled):
            idx.search_with(q[0], 10, H.SearchOptions(64).quantized(True))
        idx.enable_quantization()
        (gk, gd, gc), (ok, od, oc), st, cnt = _compare(idx, q, 10, 64, quantized=True)
        assert np.array_equal(gk, ok) and same_bits(gd, od) and np.array_equal(gc, oc)
        assert st["n_rerank"] == cnt.n_rerank


@pytest.mark.parametrize("dim", [130, 159, 200, 270, 287, 520, 1000])
def test_int8_search_reads_every_code_chunk(dim):
    """int8 traversal sums the whole code row, every 128-byte chunk of it:
    dims whose codes run one chunk past 32 * (dim // 32) rounded up (129-159,
    257-287) and the wide rows equal the oracle's quantized search
    (regression: the 129-159 and 257-287 ranges once summed one chunk short)."""
    print(f"case dim={dim}", flush=True)
    for metric in (H.Metric.L2, H.Metric.Cosine):
        b, q = O.dataset_gaussian(1500, 40, dim, 3 + dim)
        idx = build(b, metric, threads=8)
        idx.enable_quantization(on_device=True)
        (gk, gd, gc), (ok, od, oc), st, cnt = _compare(idx, q, 10, 64, quantized=True)
        assert np.array_equal(gk, ok) and same_bits(gd, od) and np.array_equal(gc, oc), metric
        assert st["n_dist"] + st["n_upper"] == cnt.n_dist and st["n_rerank"] == cnt.n_rerank


def test_one_percent_filter_at_ef_512_keeps_every_candidate():
    """A 1% filter at ef=512 keeps about ef / pass-rate (tens of thousands)
    candidates live until its results fill, far past the LDS list: the
    per-wave HBM pool holds them all (n_overflow 0), so keys, distances and
    expansions equal the oracle's unbounded heap (regression: an 8,192-entry
    pool dropped candidates on such queries)."""
    n, dim = 60000, 16
    d = H.synthetic.Dataset.lognormal(n, 4, dim, 0.5, 100, 17)
    b, q, cat = d.base(), d.queries(), d.categories
    idx = build(b, H.Metric.InnerProduct, m=8, threads=8)
    g = oracle_graph(idx)
    allow = cat == 7
    assert 0.005 < allow.mean() < 0.02
    idx.search_stats(reset=True)
    gk, gd, gc = idx.search_arrays(q, 100, 512, filter=allow)
    st = idx.search_stats()
    ok, od, oc, cnt = g.search_batch(q, 100, 512, allow=allow.astype(np.uint8))
    print(f"case 1% filter: {st}", flush=True)
    assert st["n_overflow"] == 0 and st["n_cand_spill"] > 8192, st
    assert st["n_expand"] == cnt.n_expand
    assert np.array_equal(gc, oc) and np.array_equal(gk, ok) and same_bits(gd, od)


def test_search_above_1024_dims_is_unsupported_not_wrong():
    """Rows of 1025-1055 dims fit the widest row layout but not the code
    chunks of the kernels; the search refuses them (HH_UNSUPPORTED) instead
    of returning partial sums, and 1024 dims search normally."""
    rng = np.random.default_rng(1)
    for dim, ok in ((1024, True), (1030, False)):
        b = rng.standard_normal((40, dim)).astype(np.float32)
        idx = build(b)
        if ok:
            gk, gd, gc = idx.search_arrays(b[:3], 5, 16)
            assert (gc == 5).all() and np.array_equal(gk[:, 0], np.arange(3))
        else:
            with pytest.raises(H.Unsupported):
                idx.search_arrays(b[:3], 5, 16)


def test_gather_probe_measures_and_refuses_wide_rows():
    """hh_index_gather_probe times the row gather on the index's rows and
    reports a positive rate; rows wider than 256 floats are refused."""
    import ctypes
    from hnsw_amd import _lib as L
    b, _ = O.dataset_gaussian(5000, 1, 128, 3)
    idx = build(b)
    h = idx.device_index()
    gbps, ms = ctypes.c_double(), ctypes.c_float()
    assert L.lib.hh_index_gather_probe(h, 4, ctypes.byref(gbps), ctypes.byref(ms)) == 0, L.last_error()
    assert gbps.value > 0 and ms.value > 0
    assert L.lib.hh_index_gather_probe(h, 0, ctypes.byref(gbps), ctypes.byref(ms)) != 0
    wide = build(np.random.default_rng(2).standard_normal((50, 300)).astype(np.float32))
    assert L.lib.hh_index_gather_probe(wide.device_index(), 4, ctypes.byref(gbps), None) != 0


def test_filter_matches_oracle_and_is_respected():
    """src/hnsw/tests.rs:293-313"""
    b, q = O.dataset_gaussian(2000, 10, 8, 8)
    idx = build(b)
    keys = np.arange(2000)
    allow = (keys % 10 == 3)
    (gk, gd, gc), (ok, od, oc), st, cnt = _compare(idx, q, 10, 64, allow=allow)
    assert np.array_equal(gk, ok) and same_bits(gd, od) and np.array_equal(gc, oc)
    assert (gk % 10 == 3).all() and (gc == 10).all()
    bf = H.BruteForceIndex(8)
    bf.insert_batch(keys, b)
    for i in range(len(q)):
        only = idx.search_filtered(q[i], 3, 10, lambda k: k == 1234)
        assert [n.key for n in only] == [1234]
        assert idx.search_filtered(q[i], 3, 10, lambda k: False) == []
        assert only[0] == bf.search_filtered(q[i], 1, lambda k: k == 1234)[0]


def test_tombstones_are_never_returned():
    """src/hnsw/tests.rs:220-248, with the entry point deleted."""
    b, q = O.dataset_gaussian(300, 5, 8, 9)
    idx = build(b)
    v = idx.graph_arrays()
    entry_key = int(v["keys"][v["entry_id"]])
    idx.search(q[0], 1, 10)  # uploads, so removal goes through the device update
    for key in range(0, 300, 2):
        assert idx.remove(key)
    if idx.contains(entry_key):
        idx.remove(entry_key)
    g = oracle_graph(idx)
    for i in range(len(q)):
        hits = idx.search(q[i], 20, 50)
        assert len(hits) == 20
        assert all(h.key % 2 == 1 and h.key != entry_key for h in hits)
        ok, od, _, _ = g.search(q[i], 20, 50)
        assert [h.key for h in hits] == ok.tolist()


def test_k_and_ef_edge_cases():
    """src/hnsw/tests.rs:208-218, empty index (:88-95)."""
    b, q = O.dataset_gaussian(100, 1, 4, 2)
    idx = build(b)
    assert idx.search(q[0], 0, 10) == []
    assert len(idx.search(q[0], 10, 1)) == 10
    assert len(idx.search(q[0], 500, 500)) == 100
    empty = H.HnswIndex(3)
    assert empty.search([0.0, 0.0, 0.0], 5, 10) == []


def test_invalid_queries_report_the_crate_errors():
    b, q = O.dataset_gaussian(100, 4, 4, 2)
    idx = build(b, H.Metric.Cosine)
    with pytest.raises(H.ZeroVector):
        idx.search([0, 0, 0, 0], 1, 10)
    with pytest.raises(H.NonFiniteVector):
        idx.search([np.nan, 0, 0, 1], 1, 10)
    with pytest.raises(H.NonFiniteVector):
        idx.search([3e38, 3e38, 0, 1], 1, 10)
    qq = q.copy()
    qq[2, 1] = np.inf
    with pytest.raises(H.NonFiniteVector) as e:
        idx.search_batch(qq, 3, 10)
    assert e.value.query == 2


def test_complete_graph_with_simple_selection_is_exact():
    """src/hnsw/tests.rs:113-127"""
    b, q = O.dataset_gaussian(16, 20, 8, 3)
    params = H.HnswParams(16).with_selection(H.NeighborSelection.simple())
    idx = build(b, params=params)
    bf = H.BruteForceIndex(8)
    bf.insert_batch(np.arange(16), b)
    for i in range(len(q)):
        assert idx.search(q[i], 5, 5) == bf.search(q[i], 5)


def test_inner_product_and_cosine_rank_correctly():
    """src/hnsw/tests.rs:370-388"""
    ip = H.HnswIndex(2, H.Metric.InnerProduct)
    ip.insert_batch([1, 2, 3], np.array([1.0, 0.0, 3.0, 0.0, 0.0, 1.0]))
    hits = ip.search([1.0, 0.0], 3, 10)
    assert [h.key for h in hits] == [2, 1, 3]
    assert hits[0].distance == -3.0
    cos = H.HnswIndex(2, H.Metric.Cosine)
    cos.insert_batch([1, 2], np.array([10.0, 0.0, 1.0, 1.0]))
    hits = cos.search([0.1, 0.0], 2, 10)
    assert hits[0].key == 1 and abs(hits[0].distance) < 1e-6
    assert abs(hits[1].distance - (1.0 - 0.7071067811865476)) < 1e-6


def test_bf16_rows_are_round_to_nearest_even():
    """The resident bf16 rows (hh_index_enable_bf16) are the f32 rows
    rounded to nearest even: equal to the oracle's rounding and to torch's
    CPU float32 -> bfloat16 conversion, including ties, on rows built so that
    many values sit exactly half-way between two bf16 values."""
    import torch
    rng = np.random.default_rng(3)
    b = rng.standard_normal((2000, 40)).astype(np.float32)
    # exact ties (low half 0x8000) with even and odd kept halves, and tiny values
    u = b.view(np.uint32)
    u[:500, :20] = (u[:500, :20] & 0xFFFF0000) | 0x8000
    b[500:600] *= 1e-30
    idx = build(b)
    got = idx.bf16_rows()
    want = (O.bf16_round(b).view(np.uint32) >> 16).astype(np.uint16)
    ref = torch.from_numpy(b).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(want, ref)
    assert np.array_equal(got, want)


def test_bf16_traversal_matches_the_oracle_bit_for_bit():
    """bf16 traversal is the crate's search over bf16-rounded rows with the
    f32 query and the AVX2-order sums, then the exact f32 re-rank: keys,
    distances and counters equal the oracle's bf16 mode."""
    for metric in (H.Metric.L2, H.Metric.InnerProduct, H.Metric.Cosine):
        for dim in (64, 100, 128):
            print(f"case bf16 metric={metric} dim={dim}", flush=True)
            b, q = O.dataset_gaussian(3000, 64, dim, 21 + dim)
            idx = build(b, metric, threads=4)
            idx.enable_bf16()
            g = oracle_graph(idx, bf16=True)
            for k, ef in ((10, 64), (50, 200)):
                idx.search_stats(reset=True)
                res = idx.search_arrays(q, k, ef, precision=H.Precision.BF16)
                st = idx.search_stats()
                ok, od, oc, cnt = g.search_batch(q, k, ef, precision=O.PREC_BF16)
                assert_same_results(res, (ok, od, oc), f"bf16 {metric} {dim} k={k} ef={ef}")
                assert st["n_dist"] + st["n_upper"] == cnt.n_dist
                assert st["n_expand"] == cnt.n_expand and st["n_rerank"] == cnt.n_rerank


def test_bf16_traversal_reranks_exactly():
    b, q = O.dataset_gaussian(3000, 64, 64, 21)
    idx = build(b)
    gk, gd, gc = idx.search_arrays(q, 10, 64, precision=H.Precision.BF16)
    v = idx.graph_arrays()
    key_row = {int(kk): i for i, kk in enumerate(v["keys"])}
    for i in range(len(q)):
        for j in range(int(gc[i])):
            want = O.distance(O.L2, q[i], v["vectors"][key_row[int(gk[i, j])]])
            assert same_bits([gd[i, j]], [want])
    tk, td, tc = O.brute_search_batch(O.L2, np.arange(3000), b, q, 10)
    fk, fd, fc = idx.search_arrays(q, 10, 64)
    r32 = O.recall_at_k(tk, tc, fk, fc, 10)
    r16 = O.recall_at_k(tk, tc, gk, gc, 10)
    assert r16 >= r32 - 0.005, (r16, r32)


def test_large_ef_spills_visited_set_and_stays_exact():
    """Queries that outgrow the LDS visited table continue on their HBM
    bitset and clear it afterwards (word by word from the spill log, or
    whole once the log overflows); every later query on the same
    wavefront must start clean, so all results stay the oracle's."""
    b, q = O.dataset_gaussian(20000, 40, 16, 31)
    idx = build(b, m=32, threads=8)
    h = idx.device_index()
    for cap, wpb in ((0, 0), (2048, 2)):  # the table spills past 1536 nodes
        print(f"case cap={cap} wpb={wpb}", flush=True)
        assert H.lib.hh_index_set_launch_shape(h, 512, cap, wpb) == 0
        (gk, gd, gc), (ok, od, oc), st, cnt = _compare(idx, q, 50, 512)
        assert np.array_equal(gk, ok) and same_bits(gd, od)
        assert st["n_overflow"] == 0 and st["n_visited_spill"] > 0
        assert st["n_dist"] + st["n_upper"] == cnt.n_dist
        # mostly more than 4096 nodes marked after the spill: full clears
        assert cnt.n_dist > 40 * (4096 + 1536)
    assert H.lib.hh_index_set_launch_shape(h, 512, 0, 0) == 0


def test_mixed_batch_equals_individual_searches():
    """One launch with per-query k, ef, precision and filters: each result
    equals the same query searched alone (north-star batching)."""
    b, q = O.dataset_gaussian(4000, 96, 32, 17)
    idx = build(b, threads=4)
    idx.enable_quantization()
    rng = np.random.default_rng(1)
    ks = rng.choice([1, 10, 100], 96)
    efs = rng.integers(16, 300, 96)
    precs = rng.choice([0, 1, 2], 96)
    keys = np.arange(4000)
    filters = [keys % 7 == 0, keys % 11 < 2, keys % 3 == 1]
    fids = rng.choice([-1, 0, 1, 2], 96)
    gk, gd, gc, gs = idx.search_mixed(q, ks, efs, precs, fids, filters)
    assert (gs == 0).all()
    for i in range(96):
        flt = None if fids[i] < 0 else filters[fids[i]]
        ak, ad, ac = idx.search_arrays(q[i:i + 1], int(ks[i]), int(efs[i]), H.Precision(int(precs[i])), flt)
        assert gc[i] == ac[0]
        assert np.array_equal(gk[i, :gc[i]], ak[0, :ac[0]]), i
        assert same_bits(gd[i, :gc[i]], ad[0, :ac[0]]), i


def test_mixed_batch_reports_per_query_errors():
    b, q = O.dataset_gaussian(500, 4, 8, 3)
    idx = build(b)
    qq = q.copy()
    qq[1, 0] = np.nan
    gk, gd, gc, gs = idx.search_mixed(qq, [5, 5, 5, 5], [20, 20, 20, 20], [0, 0, 1, 0])
    assert gs.tolist() == [0, 4, 8, 0]  # NonFiniteVector, QuantizationDisabled
    assert gc.tolist() == [5, 0, 0, 5]


def test_graph_captured_plan_replays_match_eager():
    import ctypes
    from hnsw_amd.device import Buffer, Stream
    b, q = O.dataset_gaussian(3000, 200, 64, 23)
    idx = build(b, threads=4)
    h = idx.device_index()
    nq, k = 200, 10
    qd = Buffer.from_array(q)
    outs = [Buffer(nq * k * 8), Buffer(nq * k * 4), Buffer(nq * 4), Buffer(nq * 4)]
    plan = ctypes.c_void_p()
    assert H.lib.hh_search_plan_create(h, qd.ptr, nq, k, 48, 0, None, None, 0, None, 0,
                                       *[t.ptr for t in outs], ctypes.byref(plan)) == 0
    ek, ed, ec = idx.search_arrays(q, k, 48)
    stream = Stream()
    try:
        for rep in range(4):
            for t in outs:
                t.zero()
            assert H.lib.hh_search_plan_launch(plan, stream.handle) == 0
            stream.synchronize()
            gk = outs[0].download((nq, k), np.uint64)
            gd = outs[1].download((nq, k), np.float32)
            gc = outs[2].download(nq, np.uint32)
            assert np.array_equal(gc, ec) and np.array_equal(gk, ek) and same_bits(gd, ed), rep
            _ = idx.search_arrays(q[:5], 3, 16)  # eager work between replays
    finally:
        H.lib.hh_search_plan_free(plan)
        stream.close()


@pytest.mark.parametrize("ef", [32, 64, 128, 300])
def test_selective_filters_match_the_oracle_exactly(ef):
    """Category filters passing 5-15% of an inner-product index keep the
    crate's unbounded candidate heap busy (results fill slowly, so almost
    every visited node becomes a candidate); candidates beyond the register
    list move to the per-query HBM pool, so results, distances and work
    counters stay identical to the oracle's, and recall@10 against the exact
    filtered truth is the oracle's."""
    d = H.synthetic.Dataset.lognormal(20000, 60, 32, 0.5, 20, 5)
    b, q, cat = d.base(), d.queries(), d.categories
    idx = build(b, H.Metric.InnerProduct, m=24, threads=8)
    g = oracle_graph(idx)
    rng = np.random.default_rng(ef)
    keys = np.arange(20000)
    gpu_found, ora_found, truth = [], [], []
    idx.search_stats(reset=True)
    n_expand = 0
    for i in range(60):
        cats = rng.choice(20, int(rng.integers(1, 4)), replace=False)
        allow = np.isin(cat, cats)
        gk, gd, gc = idx.search_arrays(q[i:i + 1], 10, ef, filter=allow)
        ok, od, oc, cnt = g.search_batch(q[i:i + 1], 10, ef, allow=allow.astype(np.uint8))
        n_expand += cnt.n_expand
        assert gc[0] == oc[0] and np.array_equal(gk[0, :gc[0]], ok[0, :oc[0]]), i
        assert same_bits(gd[0, :gc[0]], od[0, :oc[0]]), i
        tk, td = O.brute_search(O.IP, keys, b, q[i], 10, allow=allow.astype(np.uint8))
        gpu_found.append((gk[0], gc[0]))
        ora_found.append((ok[0], oc[0]))
        truth.append(tk)
    st = idx.search_stats()
    assert st["n_overflow"] == 0 and st["n_expand"] == n_expand, st
    if ef <= 64:
        assert st["n_cand_spill"] > 0, st  # the pool was exercised
    tk = np.stack(truth)
    tc = np.full(60, 10, np.uint32)
    rg = O.recall_at_k(tk, tc, np.stack([x[0] for x in gpu_found]),
                       np.array([x[1] for x in gpu_found], np.uint32), 10)
    ro = O.recall_at_k(tk, tc, np.stack([x[0] for x in ora_found]),
                       np.array([x[1] for x in ora_found], np.uint32), 10)
    assert abs(rg - ro) <= 0.005, (rg, ro)


def test_launch_shapes_do_not_change_results():
    """Visited-set capacity (down to one that spills almost every query to
    its HBM bitset) and waves per block only change speed: results,
    distances and counters stay the oracle's, for the exact-layout kernel
    (D = 128) and the general one (D = 100)."""
    for dim in (128, 100):
        b, q = O.dataset_gaussian(3000, 64, dim, 31)
        idx = build(b)
        h = idx.device_index()
        for ef, shapes in ((64, [(256, 1), (1024, 1), (1536, 2), (2048, 4)]),
                           (100, [(512, 1), (2048, 2), (4096, 1)])):
            print(f"case dim={dim} ef={ef}", flush=True)
            for cap, wpb in shapes:
                assert H.lib.hh_index_set_launch_shape(h, ef, cap, wpb) == 0, (cap, wpb)
                (gk, gd, gc), (ok, od, oc), st, cnt = _compare(idx, q, 10, ef)
                assert np.array_equal(gk, ok) and same_bits(gd, od) and np.array_equal(gc, oc), \
                    (dim, ef, cap, wpb)
                assert st["n_dist"] + st["n_upper"] == cnt.n_dist and st["n_expand"] == cnt.n_expand
                if cap <= 512:
                    assert st["n_visited_spill"] > 0  # the bitset path ran
            assert H.lib.hh_index_set_launch_shape(h, ef, 0, 0) == 0
        assert H.lib.hh_index_set_launch_shape(h, 64, 2048, 3) != 0
        assert H.lib.hh_index_set_launch_shape(h, 300, 512, 1) != 0


def test_loaded_index_file_searches_like_the_original():
    """An index written in the crate's file format and loaded back searches
    identically on the GPU (f32, int8 + re-rank, with tombstones)."""
    b, q = O.dataset_gaussian(3000, 50, 32, 44)
    idx = build(b, threads=4)
    for key in (5, 77, 1234):
        idx.remove(key)
    idx.enable_quantization()
    loaded = H.HnswIndex.from_bytes(idx.to_bytes())
    g = oracle_graph(loaded)
    for prec in (H.Precision.F32, H.Precision.INT8):
        ek, ed, ec = idx.search_arrays(q, 10, 48, precision=prec)
        gk, gd, gc = loaded.search_arrays(q, 10, 48, precision=prec)
        assert np.array_equal(ec, gc) and np.array_equal(ek, gk) and same_bits(ed, gd), prec
        # and the loaded graph searched by the oracle
        ok, od, oc, _ = g.search_batch(q, 10, 48, precision=int(prec))
        assert_same_results((gk, gd, gc), (ok, od, oc), f"loaded file {prec.name}")


def test_longest_first_order_returns_what_one_pass_returns():
    """Batches of 1-16 queries per resident wavefront at ef > 64 run in two passes
    (upper-layer descent of every query, an order by descending descent
    distance, then layer 0 in that order).  Results and counters must equal
    the one-pass launch's, and the oracle's, including queries that end
    early (k = 0, a NaN query) and mixed specs."""
    b, q = O.dataset_gaussian(4000, 20000, 24, 31)
    idx = build(b, H.Metric.L2, m=8, threads=8)
    assert idx.graph_arrays()["top"] >= 1
    runs = {}
    for longest_first in (False, True):
        print(f"case longest_first={longest_first}", flush=True)
        idx.set_query_order(longest_first)
        idx.search_stats(reset=True)
        res = idx.search_arrays(q, 10, 96)
        runs[longest_first] = res, idx.search_stats()
    (k1, d1, c1), s1 = runs[False]
    (k2, d2, c2), s2 = runs[True]
    assert np.array_equal(c1, c2) and np.array_equal(k1, k2) and same_bits(d1, d2)
    for f in ("queries", "n_dist", "n_upper", "n_expand", "n_overflow"):
        assert s1[f] == s2[f], f
    g = oracle_graph(idx)
    ok, od, oc, cnt = g.search_batch(q[:400], 10, 96)
    assert np.array_equal(k2[:400], ok) and same_bits(d2[:400], od)
    # mixed specs with queries that end early, in both orders
    nq = q.shape[0]
    rng = np.random.default_rng(5)
    ks = rng.choice([0, 1, 10], nq)
    efs = rng.integers(10, 200, nq)
    qq = q.copy()
    qq[7, 3] = np.nan
    out = {}
    for longest_first in (False, True):
        idx.set_query_order(longest_first)
        out[longest_first] = idx.search_mixed(qq, ks, efs)
    for a, b_ in zip(out[False], out[True]):
        assert np.array_equal(a.view(np.uint8), b_.view(np.uint8))
    assert out[True][3][7] == 4 and out[True][2][7] == 0
    assert (out[True][2][ks == 0] == 0).all()


def test_autotune_keeps_results_in_every_shape_and_order():
    """hh_index_autotune times the launch shapes, then one pass against two,
    and keeps the fastest for the ef bucket: the tuned index returns what the
    untuned one did, in plain and in spec batches, and a zero capacity
    restores the defaults."""
    import ctypes
    from hnsw_amd import _lib as L
    from hnsw_amd.device import Buffer
    b, q = O.dataset_gaussian(4000, 20000, 24, 31)
    idx = build(b, H.Metric.L2, m=8, threads=8)
    nq, k, ef = q.shape[0], 10, 96
    rng = np.random.default_rng(2)
    ks, efs = rng.choice([1, 10], nq), rng.integers(10, 200, nq)
    ref = idx.search_arrays(q, k, ef)
    ref_mixed = idx.search_mixed(q, ks, efs)
    h = idx.device_index()
    qd = Buffer.from_array(np.ascontiguousarray(q, np.float32))
    bk, bd, bc, bs = Buffer(nq * k * 8), Buffer(nq * k * 4), Buffer(nq * 4), Buffer(nq * 4)
    cap, wpb = ctypes.c_uint32(), ctypes.c_int()
    print("case autotune", flush=True)
    assert L.lib.hh_index_autotune(h, qd.ptr, nq, k, ef, 0, bk.ptr, bd.ptr, bc.ptr, bs.ptr, None,
                                   ctypes.byref(cap), ctypes.byref(wpb)) == 0, L.last_error()
    assert cap.value >= 512 and wpb.value in (1, 2, 4)
    tk = bk.download((nq, k), np.uint64)
    assert np.array_equal(tk, ref[0])
    got = idx.search_arrays(q, k, ef)
    assert np.array_equal(got[0], ref[0]) and same_bits(got[1], ref[1])
    assert np.array_equal(got[2], ref[2])
    got_mixed = idx.search_mixed(q, ks, efs)
    for a, b_ in zip(ref_mixed, got_mixed):
        assert np.array_equal(a.view(np.uint8), b_.view(np.uint8))
    assert L.lib.hh_index_set_launch_shape(h, ef, 0, 0) == 0
    got = idx.search_arrays(q, k, ef)
    assert np.array_equal(got[0], ref[0]) and same_bits(got[1], ref[1])


def test_work_ordered_mixed_batch_matches_one_pass_and_oracle():
    """A mixed batch larger than 16 queries per resident wave, with per-query
    k, ef up to 512 and filters of 3-60% pass rate, runs in two passes
    ordered by estimated work (ef over the sampled pass rate); results and
    statuses equal the one-pass launch's, and a sample equals the oracle."""
    d = H.synthetic.Dataset.lognormal(6000, 40000, 16, 0.5, 20, 23)
    b, q, cat = d.base(), d.queries(), d.categories
    idx = build(b, H.Metric.InnerProduct, m=8, threads=8)
    rng = np.random.default_rng(9)
    nq = q.shape[0]
    masks = [cat < c for c in (1, 2, 4, 8, 12)] + [rng.random(6000) < 0.03]
    ks = rng.choice([1, 10, 50], nq)
    efs = rng.integers(10, 513, nq)
    fids = rng.integers(-1, len(masks), nq)
    out = {}
    for longest_first in (False, True):
        print(f"case mixed longest_first={longest_first}", flush=True)
        idx.set_query_order(longest_first)
        out[longest_first] = idx.search_mixed(q, ks, efs, filter_ids=fids, filters=masks)
    for a, b_ in zip(out[False], out[True]):
        assert np.array_equal(a.view(np.uint8), b_.view(np.uint8))
    gk, gd, gc, gs = out[True]
    assert (gs == 0).all()
    g = oracle_graph(idx)
    for i in range(0, nq, 997):
        allow = None if fids[i] < 0 else masks[fids[i]].astype(np.uint8)
        ok, od, oc, _ = g.search_batch(q[i:i + 1], int(ks[i]), int(efs[i]), allow=allow)
        n = int(oc[0])
        assert gc[i] == n and np.array_equal(gk[i, :n], ok[0, :n]) and same_bits(gd[i, :n], od[0, :n]), i


@pytest.mark.parametrize("metric", [H.Metric.L2, H.Metric.InnerProduct, H.Metric.Cosine])
@pytest.mark.parametrize("dim", [128, 256])
def test_descent_kernel_two_pass_matches_one_pass(metric, dim):
    """Exact-layout rows with m <= 16 take the descent kernel (four queries
    per wavefront, whole distances per lane in the AVX2 order) as the first
    pass at every ef: the results and every counter equal the one-pass
    launch's, and the oracle's."""
    print(f"case metric={metric} dim={dim}", flush=True)
    b, q = O.dataset_gaussian(4000, 12000, dim, 83 + dim)
    idx = build(b, metric, m=16, threads=8)
    assert idx.graph_arrays()["top"] >= 1
    runs = {}
    for longest_first in (False, True):
        idx.set_query_order(longest_first)
        for ef in (32, 64):
            idx.search_stats(reset=True)
            res = idx.search_arrays(q, 10, ef)
            runs[longest_first, ef] = res, idx.search_stats()
    for ef in (32, 64):
        (k1, d1, c1), s1 = runs[False, ef]
        (k2, d2, c2), s2 = runs[True, ef]
        assert np.array_equal(c1, c2) and np.array_equal(k1, k2) and same_bits(d1, d2)
        for f in ("queries", "n_dist", "n_upper", "n_expand", "n_overflow"):
            assert s1[f] == s2[f], (ef, f)
    g = oracle_graph(idx)
    ok, od, oc, cnt = g.search_batch(q[:300], 10, 64)
    (k2, d2, c2), s2 = runs[True, 64]
    assert np.array_equal(k2[:300], ok) and same_bits(d2[:300], od)


@pytest.mark.parametrize("dim,m,metric", [
    (96, 24, H.Metric.L2), (100, 32, H.Metric.InnerProduct), (130, 16, H.Metric.Cosine),
    (300, 8, H.Metric.L2), (960, 32, H.Metric.Cosine), (1024, 24, H.Metric.L2)])
def test_descent_kernel_any_layout_and_m32(dim, m, metric):
    """The descent kernel also serves rows with a partial last 4-block, an
    8-wide remainder and a tail (dim 96, 100, 130, 300, 960) and m up to 32
    (32 lanes per query): f32, bf16 and int8 two-pass launches equal the
    one-pass launch (results and counters) and the oracle."""
    print(f"case dim={dim} m={m} metric={metric}", flush=True)
    b, q = O.dataset_gaussian(3000, 9000, dim, 7 + dim + m)
    idx = build(b, metric, m=m, threads=8)
    assert idx.graph_arrays()["top"] >= 1
    idx.enable_quantization(on_device=True)
    idx.enable_bf16()
    for prec in (H.Precision.F32, H.Precision.BF16, H.Precision.INT8):
        runs = {}
        for longest_first in (False, True):
            idx.set_query_order(longest_first)
            idx.search_stats(reset=True)
            runs[longest_first] = idx.search_arrays(q, 10, 48, precision=prec), idx.search_stats()
        (k1, d1, c1), s1 = runs[False]
        (k2, d2, c2), s2 = runs[True]
        assert np.array_equal(c1, c2) and np.array_equal(k1, k2) and same_bits(d1, d2), prec
        for f in ("queries", "n_dist", "n_upper", "n_expand", "n_rerank"):
            assert s1[f] == s2[f], (prec, f)
        if prec == H.Precision.F32:
            f32 = k2, d2
    g = oracle_graph(idx)
    ok, od, oc, _ = g.search_batch(q[:200], 10, 48)
    assert np.array_equal(f32[0][:200], ok) and same_bits(f32[1][:200], od)
    ok, od, oc, _ = g.search_batch(q[:200], 10, 48, quantized=True)
    assert np.array_equal(k2[:200], ok) and same_bits(d2[:200], od)


@pytest.mark.parametrize("metric", [H.Metric.L2, H.Metric.InnerProduct, H.Metric.Cosine])
def test_descent_kernel_two_pass_bf16_and_int8(metric):
    """The descent kernel's bf16 and int8 distances have the search
    kernel's bits: two-pass and one-pass launches agree in those
    precisions too (results after the exact re-rank, and counters)."""
    print(f"case metric={metric}", flush=True)
    b, q = O.dataset_gaussian(4000, 12000, 128, 89)
    idx = build(b, metric, m=16, threads=8)
    idx.enable_quantization(on_device=True)
    idx.enable_bf16()
    for prec in (H.Precision.BF16, H.Precision.INT8):
        runs = {}
        for longest_first in (False, True):
            idx.set_query_order(longest_first)
            idx.search_stats(reset=True)
            runs[longest_first] = idx.search_arrays(q, 10, 64, precision=prec), idx.search_stats()
        (k1, d1, c1), s1 = runs[False]
        (k2, d2, c2), s2 = runs[True]
        assert np.array_equal(c1, c2) and np.array_equal(k1, k2) and same_bits(d1, d2), prec
        for f in ("queries", "n_dist", "n_upper", "n_expand", "n_rerank"):
            assert s1[f] == s2[f], (prec, f)
    g = oracle_graph(idx)
    ok, od, oc, _ = g.search_batch(q[:300], 10, 64, quantized=True)
    assert np.array_equal(k2[:300], ok) and same_bits(d2[:300], od)


def test_pool_overflow_reruns_the_query_exactly():
    """A 0.2% filter at ef=512 over 150,000 nodes never fills its results,
    so every visited node stays a live candidate: the oracle's heap peaks
    above 90,000, more than a wave's 65,536-entry pool plus its LDS list
    hold.  Such a query is listed and run again by the re-run pass with an
    n-entry pool, so its results, distances and expansions are the oracle's
    unbounded heap's (src/hnsw/search.rs:125-135), no candidate is reported
    dropped, and the re-runs are counted."""
    n, dim = 150000, 32
    b, q = O.dataset_gaussian(n, 24, dim, 23)
    idx = build(b, H.Metric.L2, m=16, threads=16)
    g = oracle_graph(idx)
    allow = (np.arange(n) % 500) == 3
    idx.search_stats(reset=True)
    res = idx.search_arrays(q, 100, 512, filter=allow)
    st = idx.search_stats()
    print(f"case 0.2% filter over 150k nodes: {st}", flush=True)
    ok, od, oc, cnt = g.search_batch(q, 100, 512, allow=allow.astype(np.uint8))
    assert cnt.max_cands > 65536 + 1024, cnt.max_cands  # the first pool must overflow
    assert st["n_redo"] > 0 and st["n_overflow"] == 0, st
    assert st["n_expand"] == cnt.n_expand and st["n_dist"] + st["n_upper"] == cnt.n_dist, st
    assert_same_results(res, (ok, od, oc), "0.2% filter")


@pytest.mark.parametrize("metric", [H.Metric.L2, H.Metric.InnerProduct, H.Metric.Cosine])
@pytest.mark.parametrize("dim", [37, 128, 256, 300])
def test_perdim_quantizer_and_search_match_the_oracle(metric, dim):
    """Per-dimension min/max int8 traversal: the device-trained parameters,
    codes and L2 row terms equal the oracle's restatement bit for bit, and
    searches (traversal on the codes, exact f32 re-rank) return the
    oracle's keys, distances and counters."""
    print(f"case perdim metric={metric} dim={dim}", flush=True)
    b, q = O.dataset_gaussian(3000, 64, dim, 5 + dim)
    idx = build(b, metric, threads=8)
    with pytest.raises(H.QuantizationDisabled):
        idx.search_arrays(q[:1], 10, 64, precision=H.Precision.INT8_PERDIM)
    mid, scale = idx.enable_quantization_perdim()
    v = idx.graph_arrays()
    omid, oscale = O.pd_train(v["vectors"])
    assert same_bits(mid, omid) and same_bits(scale, oscale)
    codes, rn = idx.perdim_codes()
    ocodes = O.pd_encode_rows(omid, oscale, v["vectors"])
    assert np.array_equal(codes, ocodes)
    if metric == H.Metric.L2:
        assert same_bits(rn, O.pd_rnorm_rows(oscale, ocodes))
    g = oracle_graph(idx, perdim=True)
    for k, ef in ((10, 64), (1, 16), (32, 200)):
        idx.search_stats(reset=True)
        res = idx.search_arrays(q, k, ef, precision=H.Precision.INT8_PERDIM)
        st = idx.search_stats()
        ok, od, oc, cnt = g.search_batch(q, k, ef, precision=O.PREC_INT8_PERDIM)
        assert_same_results(res, (ok, od, oc), f"perdim k={k} ef={ef}")
        assert st["n_dist"] + st["n_upper"] == cnt.n_dist and st["n_expand"] == cnt.n_expand
        assert st["n_rerank"] == cnt.n_rerank


def test_perdim_two_pass_mixed_and_synced_inserts_match_the_oracle():
    """Per-dimension codes through the descent kernel (two-pass launches),
    in mixed batches beside the other precisions, and after inserts synced
    to the resident index (new rows encoded with the resident parameters)."""
    b, q = O.dataset_gaussian(5000, 12000, 128, 61)
    idx = build(b[:4000], H.Metric.L2, m=16, threads=8)
    pd = idx.enable_quantization_perdim()
    g = oracle_graph(idx, perdim=pd)
    runs = {}
    for longest_first in (False, True):
        idx.set_query_order(longest_first)
        runs[longest_first] = idx.search_arrays(q, 10, 64, precision=H.Precision.INT8_PERDIM)
    assert_same_results(runs[True], runs[False], "two-pass vs one-pass")
    ok, od, oc, _ = g.search_batch(q[:300], 10, 64, precision=O.PREC_INT8_PERDIM)
    assert_same_results(tuple(x[:300] for x in runs[True]), (ok, od, oc), "two-pass vs oracle")
    rng = np.random.default_rng(2)
    nq = 200
    ks = rng.choice([1, 10, 40], nq)
    efs = rng.integers(16, 200, nq)
    precs = rng.choice([0, 3], nq)
    gk, gd, gc, gs = idx.search_mixed(q[:nq], ks, efs, precs)
    assert (gs == 0).all()
    for i in range(nq):
        okk, odd, _, _ = g.search(q[i], int(ks[i]), int(efs[i]), precision=int(precs[i]))
        assert np.array_equal(gk[i, :gc[i]], okk) and same_bits(gd[i, :gc[i]], odd), i
    idx.insert_batch(np.arange(4000, 5000, dtype=np.uint64), b[4000:], threads=8)
    res = idx.search_arrays(q[:300], 10, 64, precision=H.Precision.INT8_PERDIM)
    g2 = oracle_graph(idx, perdim=pd)
    codes, _ = idx.perdim_codes()
    assert np.array_equal(codes, O.pd_encode_rows(pd[0], pd[1], idx.graph_arrays()["vectors"]))
    ok, od, oc, _ = g2.search_batch(q[:300], 10, 64, precision=O.PREC_INT8_PERDIM)
    assert_same_results(res, (ok, od, oc), "after synced inserts")


def test_changes_are_refused_while_served_and_stale_plans_refuse_to_launch():
    """A serving loop reads the resident arrays: while one is open, every
    change of the index is refused (Python and C ABI), so no launch can see
    freed memory; a graph-captured plan made before a change refuses to
    launch after it."""
    from hnsw_amd.device import Buffer, Stream
    b, q = O.dataset_gaussian(3000, 40, 32, 5)
    idx = build(b[:2500], threads=4)
    h = idx.device_index()
    with H.Server(idx, k_max=10) as srv:
        with pytest.raises(H.InvalidArgument):
            idx.insert_batch(np.arange(2500, 3000, dtype=np.uint64), b[2500:])
        with pytest.raises(H.InvalidArgument):
            idx.remove(3)
        assert H.lib.hh_index_enable_bf16(h) == H.InvalidArgument.status
        assert H.lib.hh_index_quantize_perdim(h, None, None) == H.InvalidArgument.status
        keys, dist, st, _ = srv.poll(srv.submit(q[0], 5, 32))
        assert st == 0 and len(keys) == 5
    nq, k = 40, 10
    qd = Buffer.from_array(q)
    outs = [Buffer(nq * k * 8), Buffer(nq * k * 4), Buffer(nq * 4), Buffer(nq * 4)]
    plan = ctypes.c_void_p()
    assert H.lib.hh_search_plan_create(h, qd.ptr, nq, k, 32, 0, None, None, 0, None, 0,
                                       *[t.ptr for t in outs], ctypes.byref(plan)) == 0
    stream = Stream()
    try:
        assert H.lib.hh_search_plan_launch(plan, stream.handle) == 0
        stream.synchronize()
        idx.insert_batch(np.arange(2500, 3000, dtype=np.uint64), b[2500:], threads=4)
        assert idx.device_index().value == h.value  # synced in place
        assert H.lib.hh_search_plan_launch(plan, stream.handle) == H.InvalidArgument.status
    finally:
        H.lib.hh_search_plan_free(plan)
        stream.close()
