"""Beam steps of the bf16 and per-dimension int8 traversals, against the
oracle's restatement of the rule (search_layer0_beam in oracle/oracle.c):
keys, distance bits and counters for every width, list width and metric,
with filters, in mixed batches, through the visited set's move to HBM, the
candidate pool's re-run and the two-pass launch."""
import ctypes

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import assert_same_results, build, oracle_graph

pytestmark = pytest.mark.gpu

PRECS = {H.Precision.BF16: O.PREC_BF16, H.Precision.INT8_PERDIM: O.PREC_INT8_PERDIM}


def _ready(idx, prec):
    if prec == H.Precision.BF16:
        idx.enable_bf16()
    else:
        idx.enable_quantization_perdim()


def _oracle(idx, prec, width):
    return oracle_graph(idx, bf16=prec == H.Precision.BF16,
                        perdim=prec == H.Precision.INT8_PERDIM, beam_width=width)


@pytest.mark.parametrize("prec", list(PRECS))
@pytest.mark.parametrize("m,metric,dim", [(16, H.Metric.L2, 128), (24, H.Metric.InnerProduct, 100),
                                          (32, H.Metric.Cosine, 128), (48, H.Metric.L2, 64),
                                          (8, H.Metric.L2, 37), (72, H.Metric.L2, 32)])
def test_beam_steps_match_the_oracle(prec, m, metric, dim):
    """Widths 1-8 on lists of 16, 32, 48, 64, 96 and 144 slots (at most
    256 slots per step, so 96-slot lists take two per step and lists of 144
    keep the crate's rule whatever the width): results, distance bits,
    expansions and distance counts equal the oracle's."""
    b, q = O.dataset_gaussian(3000, 64, dim, 3 * m + dim)
    idx = build(b, metric, m=m, threads=8)
    _ready(idx, prec)
    for width in (1, 2, 3, 4, 5, 8):
        idx.set_beam_width(width)
        g = _oracle(idx, prec, width)
        for k, ef in ((10, 64), (1, 16), (40, 150)):
            print(f"case beam prec={int(prec)} m={m} metric={metric} dim={dim} width={width} "
                  f"k={k} ef={ef}", flush=True)
            idx.search_stats(reset=True)
            res = idx.search_arrays(q, k, ef, precision=prec)
            st = idx.search_stats()
            ok, od, oc, cnt = g.search_batch(q, k, ef, precision=PRECS[prec])
            assert_same_results(res, (ok, od, oc), f"width={width} k={k} ef={ef}")
            assert st["n_expand"] == cnt.n_expand, (st, cnt.n_expand)
            assert st["n_dist"] + st["n_upper"] == cnt.n_dist
            assert st["n_rerank"] == cnt.n_rerank and st["n_overflow"] == 0


@pytest.mark.parametrize("prec", list(PRECS))
def test_beam_with_filters_and_tombstones(prec):
    b, q = O.dataset_gaussian(4000, 48, 96, 404)
    idx = build(b, H.Metric.L2, m=16, threads=8)
    for key in range(0, 4000, 7):
        idx.remove(key)
    _ready(idx, prec)
    g = _oracle(idx, prec, idx.beam_width())
    rng = np.random.default_rng(4)
    for rate in (0.5, 0.1, 0.02):
        mask = rng.random(4000) < rate
        print(f"case beam filter prec={int(prec)} rate={rate}", flush=True)
        res = idx.search_arrays(q, 10, 80, precision=prec, filter=mask)
        ok, od, oc, _ = g.search_batch(q, 10, 80, precision=PRECS[prec],
                                       allow=mask.astype(np.uint8))
        assert_same_results(res, (ok, od, oc), f"filter {rate}")


def test_beam_queries_return_the_same_alone_and_in_a_mixed_batch():
    """A mixed launch holding f32, crate-int8, bf16 and per-dimension
    queries with their own k, ef and filters: each query's results are the
    oracle's for its own precision (beam steps for bf16 and per-dimension
    codes, the crate's rule for the others)."""
    b, q = O.dataset_gaussian(5000, 300, 128, 77)
    idx = build(b, H.Metric.L2, m=16, threads=8)
    idx.enable_quantization(on_device=True)
    idx.enable_bf16()
    pd = idx.enable_quantization_perdim()
    g = oracle_graph(idx, bf16=True, perdim=pd)
    rng = np.random.default_rng(9)
    nq = len(q)
    ks = rng.choice([1, 10, 50], nq)
    efs = rng.integers(16, 300, nq)
    precs = rng.choice([0, 1, 2, 3], nq)
    masks = [rng.random(5000) < r for r in (0.3, 0.05)]
    fids = rng.choice([-1, 0, 1], nq)
    gk, gd, gc, gs = idx.search_mixed(q, ks, efs, precs, filter_ids=fids, filters=masks)
    assert (gs == 0).all()
    for i in range(nq):
        allow = None if fids[i] < 0 else masks[fids[i]].astype(np.uint8)
        okk, odd, _, _ = g.search(q[i], int(ks[i]), int(efs[i]), precision=int(precs[i]),
                                  allow=allow)
        n = int(gc[i])
        assert n == len(okk), i
        assert np.array_equal(gk[i, :n], okk), (i, int(precs[i]))
        assert np.array_equal(gd[i, :n].view(np.uint32), odd.view(np.uint32)), i
    # the bf16 queries alone give the same
    sel = np.nonzero((precs == 2) & (fids < 0) & (ks == 10))[0]
    if len(sel):
        ak, ad, ac = idx.search_arrays(q[sel], 10, 64, precision=H.Precision.BF16)
        ok, od, oc, _ = g.search_batch(q[sel], 10, 64, precision=O.PREC_BF16)
        assert_same_results((ak, ad, ac), (ok, od, oc), "bf16 alone")


@pytest.mark.parametrize("prec", list(PRECS))
def test_beam_visited_set_leaves_lds_without_filling_it(prec):
    """A 256-entry visited table (the smallest launch shape at ef <= 64)
    leaves LDS before a step of up to 256 new neighbours could fill it; the
    search continues on the HBM bitset with the oracle's results."""
    b, q = O.dataset_gaussian(6000, 96, 64, 15)
    idx = build(b, H.Metric.L2, m=16, threads=8)
    _ready(idx, prec)
    g = _oracle(idx, prec, idx.beam_width())
    lib = H.lib
    assert lib.hh_index_set_launch_shape(idx.device_index(), 64, 256, 1) == 0
    idx.search_stats(reset=True)
    res = idx.search_arrays(q, 10, 64, precision=prec)
    st = idx.search_stats()
    assert st["n_visited_spill"] > 0, st
    ok, od, oc, _ = g.search_batch(q, 10, 64, precision=PRECS[prec])
    assert_same_results(res, (ok, od, oc), "small visited table")
    assert lib.hh_index_set_launch_shape(idx.device_index(), 64, 0, 0) == 0


def test_beam_pool_overflow_reruns_exactly():
    """A 0.2% filter at ef=512 over 150,000 nodes keeps nearly every visited
    node a live candidate, beyond a wave's 65,536-entry pool: the query is
    re-run with an n-entry pool and returns the oracle's beam results."""
    n, dim = 150000, 32
    b, q = O.dataset_gaussian(n, 16, dim, 29)
    idx = build(b, H.Metric.L2, m=16, threads=16)
    idx.enable_quantization_perdim()
    g = _oracle(idx, H.Precision.INT8_PERDIM, idx.beam_width())
    allow = (np.arange(n) % 500) == 7
    idx.search_stats(reset=True)
    res = idx.search_arrays(q, 100, 512, precision=H.Precision.INT8_PERDIM, filter=allow)
    st = idx.search_stats()
    print(f"case beam 0.2% filter over 150k nodes: {st}", flush=True)
    ok, od, oc, cnt = g.search_batch(q, 100, 512, precision=O.PREC_INT8_PERDIM,
                                     allow=allow.astype(np.uint8))
    assert st["n_overflow"] == 0 and st["n_expand"] == cnt.n_expand, st
    assert_same_results(res, (ok, od, oc), "0.2% filter, beam")


@pytest.mark.parametrize("prec", list(PRECS))
def test_beam_two_pass_equals_one_pass(prec):
    b, q = O.dataset_gaussian(4000, 12000, 128, 55)
    idx = build(b, H.Metric.L2, m=16, threads=8)
    _ready(idx, prec)
    runs = {}
    for longest_first in (False, True):
        idx.set_query_order(longest_first)
        idx.search_stats(reset=True)
        runs[longest_first] = idx.search_arrays(q, 10, 64, precision=prec), idx.search_stats()
    assert_same_results(runs[True][0], runs[False][0], "two-pass vs one-pass")
    for f in ("queries", "n_dist", "n_upper", "n_expand", "n_rerank"):
        assert runs[True][1][f] == runs[False][1][f], f
    g = _oracle(idx, prec, idx.beam_width())
    ok, od, oc, _ = g.search_batch(q[:300], 10, 64, precision=PRECS[prec])
    assert_same_results(tuple(x[:300] for x in runs[True][0]), (ok, od, oc), "vs oracle")


def test_beam_width_is_validated_and_refused_while_served():
    b, _ = O.dataset_gaussian(500, 1, 16, 3)
    idx = build(b)
    for w in (0, 9):
        with pytest.raises(H.InvalidArgument):
            idx.set_beam_width(w)
    steps = ctypes.c_uint32()
    assert H.lib.hh_index_beam_steps(idx.device_index(), ctypes.byref(steps)) == 0
    assert steps.value == 4  # the default width
    idx.set_beam_width(8)
    assert H.lib.hh_index_beam_steps(idx.device_index(), ctypes.byref(steps)) == 0
    assert steps.value == 8  # m0 = 32: eight lists fill a step's 256 slots
    with H.Server(idx, k_max=10):
        with pytest.raises(H.InvalidArgument):
            idx.set_beam_width(2)
    idx.set_beam_width(2)
    assert H.lib.hh_index_beam_steps(idx.device_index(), ctypes.byref(steps)) == 0
    assert steps.value == 2
