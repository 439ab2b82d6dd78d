"""hh_search_batch_device_noalloc: the batched search with its workspace in
the library's static device array instead of allocations.  Through the C
ABI it must return exactly what hh_search_batch_device returns (keys,
distance bits, counts and statuses), call after call: the array keeps the
previous call's contents, so two calls in a row, the second smaller, must
both match; and a batch that does not fit is refused before any launch."""
import ctypes

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import build
from hnsw_amd.device import Buffer, Stream

pytestmark = pytest.mark.gpu


def _search(fn, h, qd, nq, k, ef, prec, allow, stream):
    outs = [Buffer(max(nq * k * 8, 8)), Buffer(max(nq * k * 4, 4)), Buffer(nq * 4), Buffer(nq * 4)]
    for o in outs:
        o.zero()
    st = fn(h, qd.ptr, nq, k, ef, int(prec), None if allow is None else allow.ptr,
            *[o.ptr for o in outs], stream.handle)
    assert st == 0, H._lib.last_error()
    stream.synchronize()
    return (outs[0].download((nq, k), np.uint64), outs[1].download((nq, k), np.float32),
            outs[2].download(nq, np.uint32), outs[3].download(nq, np.int32))


def _same(a, b, what):
    ak, ad, ac, ast = a
    bk, bd, bc, bst = b
    assert np.array_equal(ac, bc) and np.array_equal(ast, bst), what
    for i in range(len(ac)):
        n = int(ac[i])
        assert np.array_equal(ak[i, :n], bk[i, :n]), (what, i)
        assert np.array_equal(ad[i, :n].view(np.uint32), bd[i, :n].view(np.uint32)), (what, i)


def test_noalloc_search_equals_the_allocating_search_call_after_call():
    print("case no-alloc: 70k x 32 L2, 2000 then 300 queries", flush=True)
    b, q = O.dataset_gaussian(70000, 2000, 32, 19)
    idx = build(b, H.Metric.L2, m=16, threads=16)
    idx.enable_quantization()
    h = idx.device_index()
    q[7, 3] = np.nan  # an invalid query keeps its status and no results
    qd = Buffer.from_array(q)
    allow_mask = (np.arange(70000) % 3) != 0
    allow = Buffer.from_array(H._bits_from_mask(allow_mask))
    stream = Stream()
    try:
        for what, nq, k, ef, prec, flt in (("2000 f32", 2000, 10, 64, H.Precision.F32, None),
                                          ("300 int8 filtered", 300, 20, 128, H.Precision.INT8, allow),
                                          ("300 f32 again", 300, 10, 64, H.Precision.F32, None)):
            want = _search(H.lib.hh_search_batch_device, h, qd, nq, k, ef, prec, flt, stream)
            got = _search(H.lib.hh_search_batch_device_noalloc, h, qd, nq, k, ef, prec, flt, stream)
            _same(got, want, what)
            assert (got[3][:nq] != 0).sum() == (1 if nq > 7 else 0)
        # a batch whose prepared queries alone exceed the array: refused,
        # nothing launched (the pointer is never read)
        st = H.lib.hh_search_batch_device_noalloc(h, qd.ptr, 1 << 20, 10, 64, 0, None, qd.ptr,
                                                  qd.ptr, qd.ptr, qd.ptr, stream.handle)
        assert st == H.Unsupported.status
    finally:
        stream.close()


def test_noalloc_search_reruns_pool_overflows_like_the_allocating_search():
    """The re-run pools live in the array too: a 0.2% filter over 150,000
    nodes at ef=512 (queries overflow their first pool) returns the same
    results through both entry points."""
    print("case no-alloc: 150k x 32, 0.2% filter, ef 512", flush=True)
    n = 150000
    b, q = O.dataset_gaussian(n, 16, 32, 23)
    idx = build(b, H.Metric.L2, m=16, threads=16)
    h = idx.device_index()
    qd = Buffer.from_array(q)
    allow = Buffer.from_array(H._bits_from_mask((np.arange(n) % 500) == 3))
    stream = Stream()
    try:
        idx.search_stats(reset=True)
        want = _search(H.lib.hh_search_batch_device, h, qd, 16, 100, 512, 0, allow, stream)
        assert idx.search_stats()["n_redo"] > 0
        got = _search(H.lib.hh_search_batch_device_noalloc, h, qd, 16, 100, 512, 0, allow, stream)
        _same(got, want, "overflowing filter")
        got = _search(H.lib.hh_search_batch_device_noalloc, h, qd, 5, 100, 512, 0, allow, stream)
        _same(got, tuple(x[:5] for x in want), "overflowing filter, 5 queries")
    finally:
        stream.close()
