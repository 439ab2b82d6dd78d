"""Resident index kept in step with inserts and removals (hh_index_sync):
after any sequence of mutations, searches on the synced index equal
searches on a fresh upload of the same graph, bit for bit and counter for
counter, for f32, int8 (device-trained quantizer) and bf16 traversal."""
import ctypes

import numpy as np
import pytest

import hnsw_amd as H
from hnsw_amd import _lib as L
import oracle as O
from helpers import build, oracle_graph, same_bits

pytestmark = pytest.mark.gpu


def _fresh(idx):
    d = idx._view()
    h = ctypes.c_void_p()
    assert L.lib.hh_index_upload(ctypes.byref(d), 0, ctypes.byref(h)) == 0
    return h


def _search(h, q, k, ef, prec=0):
    q = np.ascontiguousarray(q, np.float32)
    nq = q.shape[0]
    ok = np.zeros((nq, k), np.uint64)
    od = np.zeros((nq, k), np.float32)
    oc = np.zeros(nq, np.uint32)
    bad = ctypes.c_uint64()
    st = L.lib.hh_search_batch(h, q.ctypes.data, nq, k, ef, prec, None, ok.ctypes.data,
                               od.ctypes.data, oc.ctypes.data, ctypes.byref(bad))
    assert st == 0, L.last_error()
    s = L.SearchStats()
    assert L.lib.hh_index_stats(h, ctypes.byref(s)) == 0
    assert L.lib.hh_index_reset_stats(h) == 0
    return ok, od, oc, (s.n_dist, s.n_upper, s.n_expand)


def _same_as_fresh(idx, q, k=10, ef=64, prec=0):
    h = idx.device_index()
    L.lib.hh_index_reset_stats(h)
    a = _search(h, q, k, ef, prec)
    f = _fresh(idx)
    try:
        if prec == 2:
            assert L.lib.hh_index_enable_bf16(f) == 0
        b = _search(f, q, k, ef, prec)
    finally:
        L.lib.hh_index_free(f)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[0], b[0]) and same_bits(a[1], b[1])
    assert a[3] == b[3]
    return a


def test_inserts_and_removals_sync_like_a_fresh_upload():
    b, q = O.dataset_gaussian(6000, 300, 24, 41)
    idx = build(b[:2000], H.Metric.L2, m=8, threads=4)
    h0 = idx.device_index()
    _same_as_fresh(idx, q)
    rng = np.random.default_rng(2)
    start = 2000
    # batches of growing size: the arrays grow several times
    for step, cnt in enumerate([1, 40, 700, 1500, 1759]):
        print(f"case step={step} insert={cnt}", flush=True)
        keys = np.arange(start, start + cnt, dtype=np.uint64)
        idx.insert_batch(keys, b[start:start + cnt], threads=4)
        for key in rng.choice(start, 20, replace=False):
            idx.remove(int(key))
        start += cnt
        assert idx._stale
        _same_as_fresh(idx, q, k=10, ef=48)
        assert idx.device_index().value == h0.value  # synced in place, not re-uploaded
    g = oracle_graph(idx)
    ok, od, oc, _ = g.search_batch(q, 10, 48)
    gk, gd, gc = idx.search_arrays(q, 10, 48)
    assert np.array_equal(gk, ok) and same_bits(gd, od)


def test_sync_keeps_device_quantizer_and_bf16_rows():
    b, q = O.dataset_gaussian(4000, 200, 32, 43)
    for metric in (H.Metric.L2, H.Metric.Cosine):
        print(f"case metric={metric}", flush=True)
        idx = build(b[:1500], metric, m=8, threads=4)
        idx.enable_quantization(on_device=True)
        idx.enable_bf16()
        idx.insert_batch(np.arange(1500, 4000, dtype=np.uint64), b[1500:], threads=4)
        _same_as_fresh(idx, q, prec=1)
        idx.insert_batch(np.arange(9000, 9100, dtype=np.uint64), b[:100] * 1.5, threads=4)
        h = idx.device_index()
        assert L.lib.hh_index_reset_stats(h) == 0
        a = _search(h, q, 10, 64, 2)
        f = _fresh(idx)
        try:
            assert L.lib.hh_index_enable_bf16(f) == 0
            c = _search(f, q, 10, 64, 2)
        finally:
            L.lib.hh_index_free(f)
        assert np.array_equal(a[0], c[0]) and same_bits(a[1], c[1]) and a[3] == c[3]


def test_sync_refuses_a_changed_quantizer_or_graph():
    b, q = O.dataset_gaussian(1200, 10, 16, 47)
    idx = build(b[:1000], H.Metric.L2, m=8)
    d = idx._view()
    h = ctypes.c_void_p()
    assert L.lib.hh_index_upload(ctypes.byref(d), 0, ctypes.byref(h)) == 0
    try:
        assert L.lib.hh_graph_enable_quantization(idx._g) == 0
        d = idx._view()
        assert L.lib.hh_index_sync(h, ctypes.byref(d), None, 0) == H.InvalidArgument.status
        other = build(b[:500], H.Metric.L2, m=8)
        d2 = other._view()
        assert L.lib.hh_index_sync(h, ctypes.byref(d2), None, 0) == H.InvalidArgument.status
    finally:
        L.lib.hh_index_free(h)
    # the Python index falls back to a full upload when the quantizer changes
    idx2 = build(b[:1000], H.Metric.L2, m=8)
    idx2.device_index()
    idx2.insert_batch(np.arange(1000, 1200, dtype=np.uint64), b[1000:])
    idx2.enable_quantization()
    _same_as_fresh(idx2, q, prec=1)


def test_empty_resident_index_grows_by_sync():
    b, q = O.dataset_gaussian(900, 20, 12, 53)
    idx = H.HnswIndex(12, H.Metric.L2, H.HnswParams(8))
    idx.device_index()
    idx.insert_batch(np.arange(900, dtype=np.uint64), b, threads=2)
    _same_as_fresh(idx, q, k=5, ef=32)
