"""Scalar quantizer trained and applied on the GPU (hh_index_quantize)
against the oracle's sequential restatement of ScalarQuantizer::train /
encode_into (src/quantization.rs:32-81) and the host library's
hh_graph_enable_quantization, bit for bit.

The per-dimension mean is a sequential f64 sum in row order; the
cancellation case below is built so that any other summation order (a tree
or a split over row blocks) gives a different f32 offset, which the test
checks first with an exact sum.
"""
import ctypes
import math

import numpy as np
import pytest

import hnsw_amd as H
from hnsw_amd import _lib as L
import oracle as O
from helpers import build, oracle_graph, same_bits

pytestmark = pytest.mark.gpu

METRICS = [H.Metric.L2, H.Metric.InnerProduct, H.Metric.Cosine]


def _check_against_oracle(idx, metric):
    v = idx.graph_arrays()
    off, scale = O.sq_train(int(metric), v["vectors"])
    assert same_bits(v["sq_offset"], off)
    assert same_bits(np.float32(v["sq_scale"]), scale)
    assert np.array_equal(v["codes"], O.sq_encode_rows(off, scale, v["vectors"]))
    return off, scale


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("dim,n", [(128, 3000), (100, 513), (37, 1), (3, 700), (256, 1030),
                                   (1000, 200), (16, 512)])
def test_device_quantizer_matches_oracle_bit_for_bit(metric, dim, n):
    print(f"case metric={metric} dim={dim} n={n}", flush=True)
    b, _ = O.dataset_gaussian(n, 1, dim, 40 + dim)
    idx = build(b, metric, threads=8)
    idx.enable_quantization(on_device=True)
    off, scale = _check_against_oracle(idx, metric)
    # the codes resident on the device are the ones the host graph received
    codes = np.zeros((n, dim), np.int8)
    assert L.lib.hh_index_download_codes(idx.device_index(), codes.ctypes.data) == 0
    assert np.array_equal(codes, idx.graph_arrays()["codes"])
    # the host quantizer reproduces the same parameters and codes
    before = {k: np.array(x, copy=True) for k, x in idx.graph_arrays().items()
              if k in ("sq_offset", "codes")}
    idx.enable_quantization()
    after = idx.graph_arrays()
    assert same_bits(after["sq_offset"], before["sq_offset"])
    assert np.array_equal(after["codes"], before["codes"])
    assert np.float32(after["sq_scale"]) == scale


def test_sequential_sum_order_is_kept():
    """Rows 1e17, 600 small values, -1e17 per dimension: the small values are
    absorbed in row order, so the crate's offset differs from the exactly
    rounded mean, and the device must reproduce the crate's."""
    dim, n_small = 40, 600
    rng = np.random.default_rng(3)
    small = rng.uniform(1.0, 2.0, (n_small, dim)).astype(np.float32)
    base = np.concatenate([np.full((1, dim), 1e17, np.float32), small,
                           np.full((1, dim), -1e17, np.float32)])
    exact = np.array([math.fsum(base[:, d].astype(np.float64)) / base.shape[0]
                      for d in range(dim)]).astype(np.float32)
    off, scale = O.sq_train(int(H.Metric.L2), base)
    assert not np.array_equal(off, exact)  # the order matters for this data
    idx = build(base, H.Metric.L2, threads=4)
    idx.enable_quantization(on_device=True)
    got, _ = _check_against_oracle(idx, H.Metric.L2)
    assert same_bits(got, off)


def test_empty_index_gets_the_neutral_quantizer():
    """rows = 0: offset 0 and scale 1 (src/quantization.rs:35-53)."""
    idx = H.HnswIndex(12, H.Metric.L2, H.HnswParams(8))
    idx.enable_quantization(on_device=True)
    off, scale = idx.quantizer()
    assert np.array_equal(off, np.zeros(12, np.float32)) and scale == 1.0
    assert idx.is_quantized()
    q = np.ones(12, np.float32)
    assert list(idx.search_with(q, 3, H.SearchOptions(16).quantized(True))) == []
    plain = H.HnswIndex(12, H.Metric.L2, H.HnswParams(8))
    with pytest.raises(H.QuantizationDisabled):
        plain.search_with(q, 3, H.SearchOptions(16).quantized(True))


def test_quantized_search_after_device_training_matches_oracle():
    """int8 traversal + exact re-rank on device codes (src/hnsw/mod.rs:483-500)."""
    for metric in METRICS:
        print(f"case metric={metric}", flush=True)
        b, q = O.dataset_gaussian(2000, 40, 32, 12)
        idx = build(b, metric)
        idx.enable_quantization(on_device=True)
        g = oracle_graph(idx)
        gk, gd, gc = idx.search_arrays(q, 10, 64, precision=H.Precision.INT8)
        ok, od, oc, _ = g.search_batch(q, 10, 64, quantized=True)
        assert np.array_equal(gc, oc)
        assert np.array_equal(gk, ok)
        assert same_bits(gd, od)


def test_inserts_after_device_training_keep_the_quantizer():
    """Vectors inserted after enable_quantization are encoded with the
    trained quantizer, which is not retrained (src/hnsw/mod.rs:544-547)."""
    b, q = O.dataset_gaussian(1500, 20, 48, 21)
    idx = build(b[:1000], H.Metric.L2, threads=4)
    idx.enable_quantization(on_device=True)
    off, scale = idx.quantizer()
    idx.insert_batch(np.arange(1000, 1500, dtype=np.uint64), b[1000:], threads=4)
    off2, scale2 = idx.quantizer()
    assert same_bits(off2, off) and np.float32(scale2) == np.float32(scale)
    v = idx.graph_arrays()
    assert np.array_equal(v["codes"], O.sq_encode_rows(off, np.float32(scale), v["vectors"]))
    g = oracle_graph(idx)
    gk, gd, gc = idx.search_arrays(q, 10, 64, precision=H.Precision.INT8)
    ok, od, oc, _ = g.search_batch(q, 10, 64, quantized=True)
    assert np.array_equal(gk, ok) and same_bits(gd, od)


def _rows_only_index(vectors, metric):
    """An index of the given rows with no links (layer-0 counts 0): enough
    to hold rows for the quantizer at full size without building a graph."""
    n, dim = vectors.shape
    m, m0 = 16, 32
    keys = np.arange(n, dtype=np.uint64)
    levels = np.zeros(n, np.uint8)
    layer0 = np.zeros(n * (m0 + 1), np.uint32)
    upper_off = np.zeros(n, np.uint64)
    deleted = np.zeros(n, np.uint8)
    d = L.GraphDesc(dim=dim, metric=int(metric), m=m, m0=m0, n=n, entry_id=0, top_level=0,
                    vectors=vectors.ctypes.data, keys=keys.ctypes.data,
                    levels=levels.ctypes.data, layer0=layer0.ctypes.data,
                    upper_off=upper_off.ctypes.data, upper=None, deleted=deleted.ctypes.data,
                    codes=None, sq_offset=None, sq_scale=1.0, live=n)
    h = ctypes.c_void_p()
    assert L.lib.hh_index_upload(ctypes.byref(d), 0, ctypes.byref(h)) == 0
    return h


@pytest.mark.parametrize("metric", [H.Metric.L2, H.Metric.InnerProduct])
def test_full_size_training_matches_oracle(metric):
    """1M x 128 rows (the benchmark's base set shape): offsets and scale bit
    for bit, and the codes of sampled rows."""
    n, dim = 1_000_000, 128
    rng = np.random.default_rng(11)
    x = rng.standard_normal((n, dim), dtype=np.float32)
    x += rng.uniform(-3, 3, dim).astype(np.float32)  # per-dimension shift
    h = _rows_only_index(x, metric)
    try:
        off = np.zeros(dim, np.float32)
        sc = ctypes.c_float()
        print(f"case full-size metric={metric}", flush=True)
        assert L.lib.hh_index_quantize(h, off.ctypes.data, ctypes.byref(sc)) == 0
        codes = np.zeros((n, dim), np.int8)
        assert L.lib.hh_index_download_codes(h, codes.ctypes.data) == 0
    finally:
        L.lib.hh_index_free(h)
    ooff, oscale = O.sq_train(int(metric), x)
    assert same_bits(off, ooff)
    assert np.float32(sc.value) == oscale
    sample = rng.choice(n, 2000, replace=False)
    assert np.array_equal(codes[sample], O.sq_encode_rows(ooff, oscale, x[sample]))
    assert codes.min() >= -127 and codes.max() <= 127


def test_int8_search_on_an_index_that_searched_before_training():
    """The query workspace made by earlier f32 searches holds no code rows;
    training on the resident index must size it for the query codes."""
    b, q = O.dataset_gaussian(3000, 600, 64, 61)
    idx = build(b, H.Metric.L2, threads=4)
    idx.search_arrays(q, 10, 64)                 # workspace sized without codes
    idx.search_arrays(q[:100], 10, 64)
    idx.enable_quantization(on_device=True)
    g = oracle_graph(idx)
    for n in (600, 100):
        print(f"case nq={n}", flush=True)
        gk, gd, gc = idx.search_arrays(q[:n], 10, 64, precision=H.Precision.INT8)
        ok, od, oc, _ = g.search_batch(q[:n], 10, 64, quantized=True)
        assert np.array_equal(gk, ok) and same_bits(gd, od) and np.array_equal(gc, oc)
