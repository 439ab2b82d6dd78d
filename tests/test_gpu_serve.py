"""Serving loop: requests arriving while others run are admitted as launch
slots free up, and every result equals the same query searched alone."""
import threading
import time

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import build, oracle_graph, same_bits

pytestmark = pytest.mark.gpu


def test_results_are_the_same_alone_and_in_the_mix():
    b, q = O.dataset_gaussian(4000, 400, 32, 29)
    idx = build(b, threads=4)
    idx.enable_quantization()
    idx.enable_bf16()
    keys = np.arange(4000)
    masks = [keys % 10 == 1, keys % 4 != 0]
    rng = np.random.default_rng(7)
    ks = rng.choice([1, 10, 50], 400)
    efs = rng.choice([16, 64, 256], 400)
    precs = rng.choice([0, 1, 2], 400)
    fsel = rng.choice([-1, -1, 0, 1], 400)
    gaps = rng.exponential(1.0 / 20000.0, 400)  # Poisson arrivals, 20k requests/s
    tickets = [None] * 400
    with H.Server(idx, k_max=50, max_batch=32, max_inflight=2) as srv:
        fids = [srv.register_filter(m) for m in masks]

        def producer():
            for i in range(400):
                time.sleep(gaps[i])
                fid = -1 if fsel[i] < 0 else fids[fsel[i]]
                tickets[i] = srv.submit(q[i], ks[i], efs[i], precs[i], fid)

        t = threading.Thread(target=producer)
        t.start()
        t.join()
        got = [srv.poll(tk, wait=True) for tk in tickets]
        st = srv.stats()
    assert st["served"] == 400 and st["error"] == 0
    assert st["batches"] > 1, "requests were admitted in several launches"
    lat = np.array([g[3] for g in got])
    assert (lat > 0).all()
    g = oracle_graph(idx, bf16=True)
    for i in range(400):
        flt = None if fsel[i] < 0 else masks[fsel[i]]
        ak, ad, ac = idx.search_arrays(q[i:i + 1], int(ks[i]), int(efs[i]), H.Precision(int(precs[i])), flt)
        gk, gd, gs, _ = got[i]
        assert gs == 0
        assert np.array_equal(gk, ak[0, :ac[0]]), i
        assert same_bits(gd, ad[0, :ac[0]]), i
        # and the oracle's search of the same request
        ok, od, _, _ = g.search(q[i], int(ks[i]), int(efs[i]), precision=int(precs[i]),
                                allow=None if flt is None else flt.astype(np.uint8))
        assert np.array_equal(gk, ok) and same_bits(gd, od), i


def test_queue_limit_and_invalid_requests():
    b, q = O.dataset_gaussian(500, 4, 8, 3)
    idx = build(b)
    with H.Server(idx, k_max=10, max_batch=4, max_inflight=1) as srv:
        bad = q[0].copy()
        bad[0] = np.inf
        t = srv.submit(bad, 5, 20)
        keys, dist, status, _ = srv.poll(t)
        assert status == 4 and len(keys) == 0
        with pytest.raises(H.InvalidParameter):
            srv.submit(q[0], 11, 20)


def test_launches_take_one_ef_bucket_at_a_time():
    """Requests queue per ef bucket and a launch takes one bucket: a burst
    of mixed ef 16 / 256 requests is served in launches that each hold one
    bucket (so at least two launches), and every result is the oracle's."""
    b, q = O.dataset_gaussian(3000, 200, 32, 41)
    idx = build(b, threads=4)
    g = oracle_graph(idx)
    efs = np.where(np.arange(200) % 2 == 0, 16, 256)
    ks = np.full(200, 10)
    with H.Server(idx, k_max=10, max_inflight=1) as srv:
        lat, sk, sd, sc, sst = srv.replay(q, ks, efs, np.zeros(200))
        st = srv.stats()
    assert (sst == 0).all() and st["batches"] >= 2
    for i in range(200):
        ok, od, _, _ = g.search(q[i], 10, int(efs[i]))
        c = int(sc[i])
        assert np.array_equal(sk[i, :c], ok) and same_bits(sd[i, :c], od), i


def test_registered_filters_are_copied_whole_and_respected():
    """The filter table keeps exactly the bitmaps registered, whatever the
    host allocator does with the caller's memory afterwards (the binding
    keeps its bitmap alive until the library has copied it), and every
    served result of a filtered request passes its filter and equals the
    oracle's: several servers in a row, each with every request filtered."""
    b, q = O.dataset_gaussian(6000, 300, 32, 31)
    idx = build(b, threads=4)
    g = oracle_graph(idx)
    keys = idx.graph_arrays()["keys"]
    mask = (keys % 10) == 3
    allow = mask.astype(np.uint8)
    want = [g.search(q[i], 10, 64, allow=allow)[0] for i in range(len(q))]
    for rep in range(3):
        with H.Server(idx, k_max=10, max_inflight=2) as srv:
            fid = srv.register_filter(mask)
            junk = [np.random.default_rng(rep).random(5000) for _ in range(20)]  # reuse freed memory
            del junk
            arrival = np.cumsum(np.full(len(q), 5.0))
            lat, sk, sd, sc, sst = srv.replay(q, np.full(len(q), 10), np.full(len(q), 64), arrival,
                                              filter_ids=np.full(len(q), fid))
        assert (sst == 0).all()
        for i in range(len(q)):
            c = int(sc[i])
            assert mask[sk[i, :c].astype(np.int64)].all(), (rep, i)
            assert np.array_equal(sk[i, :c], want[i]), (rep, i)
