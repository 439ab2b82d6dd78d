"""GPU parity at the workloads of the benchmark configurations, each at a
size the oracle checks in seconds: the same data generators, metrics, graph
parameters, k, ef, precisions, filters and launch forms, against the oracle
on the same graph (keys, distance bits, counts and work counters).

C1  gaussian(10000, 1000, 128, 42), L2, M=16, efC=200, ef 32/64/128 — the
    crate's own workload, full size; recall on its first 200 queries equals
    the published table (docs/benchmarks.md:162-169).
C3  U(0,1) rows L2-normalised, cosine, D=960, M=32, k 10/100, ef 64/256, in
    f32, the crate's int8, bf16 and per-dimension int8 traversal.
C4  N(0,1) rows with lognormal norms, inner product, D=256, M=24, one mixed
    batch (k 1/10/100, ef 32..512, 1-3 of 20 categories per query) replayed
    from a HIP graph: the D=256 kernel at the 512-entry list bucket.
C5  8 Gaussian clusters, sigma 0.3, L2, D=96, M=16, bf16 rows + int8 codes
    (the crate's and the per-dimension ones), served by the continuous-
    batching loop with Poisson arrivals, k 1/10/50, ef 16/64/256 and an
    optional 10% key bitmap.
"""
import ctypes
import json
import os

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import assert_same_results, build, oracle_graph, same_bits

pytestmark = pytest.mark.gpu

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def _counters_match(st, cnt, what):
    assert st["n_dist"] + st["n_upper"] == cnt.n_dist, (what, st, cnt.n_dist)
    assert st["n_expand"] == cnt.n_expand, (what, st["n_expand"], cnt.n_expand)
    assert st["n_overflow"] == 0, (what, st)


def test_c1_gaussian_10k_workload_matches_oracle_and_published_recall():
    print("case C1 gaussian 10k x 128, 1000 queries", flush=True)
    k = KAT["recall_table_gaussian_10k"]
    base, queries = O.dataset_gaussian(10000, 1000, 128, 42)
    idx = build(base, m=16, threads=1)  # sequential: the published table's build
    g = oracle_graph(idx)
    brute = H.BruteForceIndex(128)
    brute.insert_batch(np.arange(10000), base)
    tk, td, tc = brute.search_arrays(queries, 10)
    ok_, od_, oc_ = O.brute_search_batch(O.L2, np.arange(10000), base, queries[:200], 10)
    assert np.array_equal(tk[:200], ok_) and same_bits(td[:200], od_)
    published = dict(zip(k["efs"], k["m16"]))
    for ef in (32, 64, 128):
        idx.search_stats(reset=True)
        gk, gd, gc = idx.search_arrays(queries, 10, ef)
        st = idx.search_stats()
        ok, od, oc, cnt = g.search_batch(queries, 10, ef)
        assert_same_results((gk, gd, gc), (ok, od, oc), f"C1 ef={ef}")
        _counters_match(st, cnt, f"C1 ef={ef}")
        # the first 200 queries of this stream are the table's 200 queries
        r200 = O.recall_at_k(tk[:200], tc[:200], gk[:200], gc[:200], 10)
        assert round(r200, 4) == published[ef], (ef, r200)


@pytest.fixture(scope="module")
def c3_index():
    d = H.synthetic.Dataset.unit_uniform(12000, 200, 960, 7)
    idx = build(d.base(), H.Metric.Cosine, m=32, threads=16)
    idx.enable_quantization()
    idx.enable_bf16()
    mid, scale = idx.enable_quantization_perdim()
    g = oracle_graph(idx, bf16=True, perdim=True)
    return idx, g, d.queries(), mid, scale


@pytest.mark.parametrize("prec", [H.Precision.F32, H.Precision.INT8, H.Precision.BF16,
                                  H.Precision.INT8_PERDIM])
def test_c3_cosine_960_every_precision_matches_oracle(c3_index, prec):
    idx, g, q, mid, scale = c3_index
    print(f"case C3 cosine 960 prec={prec.name}", flush=True)
    if prec == H.Precision.INT8_PERDIM:
        omid, oscale = O.pd_train(idx.graph_arrays()["vectors"])
        assert same_bits(mid, omid) and same_bits(scale, oscale)
    for k, ef in ((10, 64), (10, 256), (100, 64), (100, 256)):
        idx.search_stats(reset=True)
        res = idx.search_arrays(q, k, ef, precision=prec)
        st = idx.search_stats()
        ok, od, oc, cnt = g.search_batch(q, k, ef, precision=int(prec))
        assert_same_results(res, (ok, od, oc), f"C3 {prec.name} k={k} ef={ef}")
        _counters_match(st, cnt, f"C3 {prec.name} k={k} ef={ef}")
        if prec != H.Precision.F32:
            assert st["n_rerank"] == cnt.n_rerank


def test_c3_recall_and_distance_error_against_fp64(c3_index):
    """Recall@k of every precision against the FP64 brute-force oracle, and
    the returned distances' error against FP64 (all are exact f32
    re-scores): each approximate traversal stays within 0.02 of f32's
    recall, and every returned distance is within 1e-5 relative of FP64."""
    idx, g, q, _, _ = c3_index
    v = idx.graph_arrays()
    brute = H.BruteForceIndex(960, H.Metric.Cosine)
    brute.insert_batch(v["keys"], v["vectors"])
    for k, ef in ((10, 64), (100, 256)):
        tk, td, tc = brute.search_arrays(q, k, f64=True)
        rec = {}
        for prec in H.Precision:
            gk, gd, gc = idx.search_arrays(q, k, ef, precision=prec)
            rec[prec.name] = O.recall_at_k(tk, tc, gk, gc, k)
            row = {int(kk): i for i, kk in enumerate(v["keys"])}
            for i in range(0, len(q), 17):
                for j in range(int(gc[i])):
                    x = v["vectors"][row[int(gk[i, j])]].astype(np.float64)
                    exact = 1.0 - float(np.dot(q[i].astype(np.float64) / np.linalg.norm(q[i].astype(np.float64)), x))
                    assert abs(float(gd[i, j]) - exact) <= 1e-5 * max(1.0, abs(exact)), (prec, i, j)
        print(f"C3 k={k} ef={ef} recall vs FP64: {rec}", flush=True)
        for name, r in rec.items():
            assert r >= rec["F32"] - 0.02, (k, ef, rec)


def test_c4_mixed_filtered_batch_replayed_from_a_graph_matches_oracle():
    from hnsw_amd.device import Buffer, Stream
    print("case C4 IP 50k x 256, M=24, mixed filtered batch", flush=True)
    d = H.synthetic.Dataset.lognormal(50000, 1500, 256, 0.5, 20, 11)
    b, q, cat = d.base(), d.queries(), d.categories
    idx = build(b, H.Metric.InnerProduct, m=24, threads=16)
    g = oracle_graph(idx)
    rng = np.random.default_rng(4)
    nq = q.shape[0]
    ks = rng.choice([1, 10, 100], nq).astype(np.uint32)
    efs = rng.integers(32, 513, nq).astype(np.uint32)
    nfil = 64
    cats = [rng.choice(20, int(rng.integers(1, 4)), replace=False) for _ in range(nfil)]
    masks = [np.isin(cat, c) for c in cats]
    rates = np.array([m.mean() for m in masks])
    assert 0.03 < rates.min() and rates.max() < 0.2
    fids = rng.integers(0, nfil, nq).astype(np.uint32)
    specs = np.stack([ks, efs, np.zeros(nq, np.uint32), fids], axis=1).astype(np.uint32)
    words = np.ascontiguousarray(np.stack([H._bits_from_mask(m) for m in masks]))
    kst = 100
    h = idx.device_index()
    qd, sd, fd = Buffer.from_array(q), Buffer.from_array(specs), Buffer.from_array(words)
    outs = [Buffer(nq * kst * 8), Buffer(nq * kst * 4), Buffer(nq * 4), Buffer(nq * 4)]
    plan = ctypes.c_void_p()
    assert H.lib.hh_search_plan_create(h, qd.ptr, nq, 0, 512, 0, None, sd.ptr, kst, fd.ptr, nfil,
                                       *[t.ptr for t in outs], ctypes.byref(plan)) == 0, \
        H._lib.last_error()
    ek, ed, ec, es = idx.search_mixed(q, ks, efs, None, fids.astype(np.int64), masks)
    assert (es == 0).all()
    stream = Stream()
    try:
        for rep in range(2):
            for t in outs:
                t.zero()
            idx.search_stats(reset=True)
            assert H.lib.hh_search_plan_launch(plan, stream.handle) == 0
            stream.synchronize()
            st = idx.search_stats()
            gk = outs[0].download((nq, kst), np.uint64)
            gd = outs[1].download((nq, kst), np.float32)
            gc = outs[2].download(nq, np.uint32)
            gs = outs[3].download(nq, np.int32)
            assert (gs == 0).all() and np.array_equal(gc, ec), rep
            for i in range(nq):
                n = int(gc[i])
                assert np.array_equal(gk[i, :n], ek[i, :n]) and same_bits(gd[i, :n], ed[i, :n]), (rep, i)
            assert st["n_overflow"] == 0 and st["n_redo"] == 0, st
    finally:
        H.lib.hh_search_plan_free(plan)
        stream.close()
    # a sample of queries against the oracle, each with its own k, ef and filter
    for i in rng.choice(nq, 200, replace=False):
        allow = masks[fids[i]].astype(np.uint8)
        okk, odd, _, _ = g.search(q[i], int(ks[i]), int(efs[i]), allow=allow)
        n = int(ec[i])
        assert n == len(okk) and np.array_equal(ek[i, :n], okk) and same_bits(ed[i, :n], odd), i


def test_c5_serving_bf16_and_int8_matches_oracle():
    print("case C5 L2 20k x 96 served, bf16 + int8", flush=True)
    d = H.synthetic.Dataset.gauss_clusters(20000, 800, 96, 8, 0.3, 13)
    b, q = d.base(), d.queries()
    idx = build(b, H.Metric.L2, m=16, threads=16)
    idx.enable_quantization()
    idx.enable_bf16()
    pd = idx.enable_quantization_perdim()
    g = oracle_graph(idx, bf16=True, perdim=pd)
    keys = idx.graph_arrays()["keys"]
    mask = (keys % 10) == 3  # a 10% key bitmap
    rng = np.random.default_rng(9)
    n = q.shape[0]
    ks = rng.choice([1, 10, 50], n)
    efs = rng.choice([16, 64, 256], n)
    precs = rng.choice([int(H.Precision.BF16), int(H.Precision.INT8),
                        int(H.Precision.INT8_PERDIM)], n)
    fsel = np.where(rng.random(n) < 0.3, 0, -1)
    arrival = np.cumsum(rng.exponential(1e6 / 200000.0, n))  # 200k requests/s
    with H.Server(idx, k_max=50, max_inflight=2) as srv:
        fid = srv.register_filter(mask)
        lat, sk, sdist, sc, sst = srv.replay(q, ks, efs, arrival, precs,
                                             np.where(fsel < 0, -1, fid))
        stats = srv.stats()
    assert stats["served"] == n and stats["error"] == 0
    assert (sst == 0).all() and (lat > 0).all()
    print(f"C5 served {n}: p50 {np.percentile(lat, 50):.3f} ms, p99 {np.percentile(lat, 99):.3f} ms, "
          f"{stats['batches']} launches", flush=True)
    for i in range(n):
        allow = mask.astype(np.uint8) if fsel[i] >= 0 else None
        okk, odd, _, _ = g.search(q[i], int(ks[i]), int(efs[i]), precision=int(precs[i]),
                                  allow=allow)
        c = int(sc[i])
        assert c == len(okk), i
        assert np.array_equal(sk[i, :c], okk) and same_bits(sdist[i, :c], odd), i
