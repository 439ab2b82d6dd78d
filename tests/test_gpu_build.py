"""Construction with the layer-0 searches of INSERT on the GPU.

hh_index_build_candidates must return what INSERT's layer-0 search returns
for an unlinked node on the same graph (the oracle's search with the node's
vector, ef = max(ef_construction, m), tombstones included as candidates),
bit for bit.  A graph built with device=True is of the same kind as a
parallel host build (batches do not see each other), so it is checked by
graph invariants and by recall against a host build of the same data."""
import ctypes

import numpy as np
import pytest

import hnsw_amd as H
from hnsw_amd import _lib as L
import oracle as O
from helpers import build, oracle_graph, same_bits

pytestmark = pytest.mark.gpu


def _append(idx, keys, vecs):
    first, bad = ctypes.c_uint64(), ctypes.c_uint64()
    vecs = np.ascontiguousarray(vecs, np.float32)
    keys = np.ascontiguousarray(keys, np.uint64)
    assert L.lib.hh_graph_append(idx._g, keys.ctypes.data, vecs.ctypes.data, keys.size,
                                 ctypes.byref(first), ctypes.byref(bad)) == 0
    idx._mark_changed()
    return first.value


@pytest.mark.parametrize("metric", [H.Metric.L2, H.Metric.InnerProduct])
def test_candidates_equal_the_layer0_search_of_insert(metric):
    b, _ = O.dataset_gaussian(3300, 1, 32, 71)
    idx = build(b[:3000], metric, m=8, threads=4)
    for key in range(0, 3000, 37):
        idx.remove(key)  # tombstones stay candidates during construction
    first = _append(idx, np.arange(3000, 3300), b[3000:])
    h = idx.device_index()
    ef = 200
    ids = np.zeros((300, ef), np.uint32)
    d = np.zeros((300, ef), np.float32)
    c = np.zeros(300, np.uint32)
    assert L.lib.hh_index_build_candidates(h, first, 300, ef, ids.ctypes.data, d.ctypes.data,
                                           c.ctypes.data) == 0
    v = idx.graph_arrays()
    # the oracle's search with every node admitted (no tombstones)
    g = O.OracleGraph(v["dim"], v["metric"], v["m"], v["m0"], v["vectors"], v["keys"],
                      v["levels"], v["layer0"], v["upper_off"], v["upper"], v["entry_id"],
                      v["top"])
    lv = v["levels"][3000:]
    rows = v["vectors"][3000:]
    ok, od, oc, _ = g.search_batch(rows, ef, ef)
    keys = v["keys"]
    for i in range(300):
        if lv[i] != 0:
            continue
        assert c[i] == oc[i]
        assert np.array_equal(keys[ids[i, :c[i]]], ok[i, :oc[i]]), i
        assert same_bits(d[i, :c[i]], od[i, :oc[i]]), i


def _invariants(v):
    n, m, m0 = v["n"], v["m"], v["m0"]
    lay0 = v["layer0"].reshape(n, m0 + 1)
    for node in range(n):
        for layer in range(int(v["levels"][node]) + 1):
            if layer == 0:
                cnt = min(int(lay0[node, 0]), m0)
                nb = lay0[node, 1:1 + cnt]
            else:
                o = int(v["upper_off"][node]) + (layer - 1) * (m + 1)
                cnt = min(int(v["upper"][o]), m)
                nb = v["upper"][o + 1:o + 1 + cnt]
            assert node not in nb
            assert len(set(nb.tolist())) == len(nb)
            assert all(int(v["levels"][x]) >= layer for x in nb)


def test_device_build_matches_host_build_recall():
    b, q = O.dataset_gaussian(20000, 300, 32, 73)
    host = build(b, H.Metric.L2, m=16, threads=8)
    dev = H.HnswIndex(32, H.Metric.L2, H.HnswParams(16))
    dev.insert_batch(np.arange(20000, dtype=np.uint64), b, threads=8, device=True)
    v = dev.graph_arrays()
    assert v["n"] == 20000 and np.array_equal(v["levels"], host.graph_arrays()["levels"])
    _invariants(v)
    bf = H.BruteForceIndex(32, H.Metric.L2)
    bf.insert_batch(np.arange(20000, dtype=np.uint64), b)
    tk, _, tc = bf.search_arrays(q, 10)
    for ef in (32, 64):
        rh = H.lib.hh_recall_at_k(300, tk.ctypes.data, tc.ctypes.data, 10,
                                  *[a.ctypes.data for a in host.search_arrays(q, 10, ef)[::2]], 10, 10)
        rd = H.lib.hh_recall_at_k(300, tk.ctypes.data, tc.ctypes.data, 10,
                                  *[a.ctypes.data for a in dev.search_arrays(q, 10, ef)[::2]], 10, 10)
        print(f"ef={ef}: host recall {rh:.4f}, device-assisted recall {rd:.4f}", flush=True)
        assert rd >= rh - 0.02
    # searches on the device-built graph are still exact against the oracle
    g = oracle_graph(dev)
    ok, od, oc, _ = g.search_batch(q, 10, 64)
    gk, gd, gc = dev.search_arrays(q, 10, 64)
    assert np.array_equal(gk, ok) and same_bits(gd, od)


def test_device_build_grows_an_existing_index():
    b, q = O.dataset_gaussian(9000, 100, 24, 79)
    idx = build(b[:5000], H.Metric.Cosine, m=8, threads=4)
    idx.device_index()
    idx.insert_batch(np.arange(5000, 9000, dtype=np.uint64), b[5000:], threads=4, device=True)
    _invariants(idx.graph_arrays())
    assert len(idx) == 9000
    g = oracle_graph(idx)
    ok, od, oc, _ = g.search_batch(q, 10, 48)
    gk, gd, gc = idx.search_arrays(q, 10, 48)
    assert np.array_equal(gk, ok) and same_bits(gd, od)
    with pytest.raises(H.DuplicateKey):
        idx.insert_batch(np.array([42, 100000], np.uint64), b[:2], device=True)
    assert len(idx) == 9000
