"""Serving loop: requests arriving while others run are admitted as launch
slots free up, and every result equals the same query searched alone."""
import threading
import time

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import build, same_bits

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
    for i in range(400):
        flt = None if fsel[i] < 0 else masks[fsel[i]]
        ak, ad, ac = idx.search_arrays(q[i:i + 1], int(ks[i]), int(efs[i]), H.Precision(int(precs[i])), flt)
        gk, gd, gs, _ = got[i]
        assert gs == 0
        assert np.array_equal(gk, ak[0, :ac[0]]), i
        assert same_bits(gd, ad[0, :ac[0]]), i


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
