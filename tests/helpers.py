"""Shared test helpers: build graphs with the host library and hand the
same graph to the oracle."""
import numpy as np

import hnsw_amd as H
import oracle as O


def oracle_graph(idx, bf16=False, perdim=False, beam_width=None):
    """The oracle's view of the index's current graph (arrays copied, so the
    view outlives later mutation of the index).  bf16: the oracle also
    traverses bf16-rounded rows; perdim: True trains the per-dimension
    quantizer on the rows (the oracle's restatement), a (mid, scale) tuple
    uses those parameters, for precision 3.  beam_width: the beam steps of
    those two traversals (default: the index's)."""
    v = {k: (np.array(x, copy=True) if isinstance(x, np.ndarray) else x)
         for k, x in idx.graph_arrays().items()}
    if isinstance(perdim, tuple):
        pd = perdim  # (mid, scale) of a quantizer trained earlier
    else:
        pd = O.pd_train(v["vectors"]) if perdim else None
    return O.OracleGraph(v["dim"], v["metric"], v["m"], v["m0"], v["vectors"], v["keys"],
                         v["levels"], v["layer0"], v["upper_off"], v["upper"], v["entry_id"],
                         v["top"], deleted=v["deleted"], codes=v["codes"],
                         sq_offset=v["sq_offset"], sq_scale=v["sq_scale"], bf16=bf16,
                         perdim=pd,
                         beam_width=idx.beam_width() if beam_width is None else beam_width)


def build(base, metric=H.Metric.L2, m=16, keys=None, threads=1, params=None):
    base = np.ascontiguousarray(base, np.float32)
    idx = H.HnswIndex(base.shape[1], metric, params or H.HnswParams(m))
    if keys is None:
        keys = np.arange(base.shape[0], dtype=np.uint64)
    idx.insert_batch(keys, base, threads=threads)
    return idx


def same_bits(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def assert_same_results(g, o, what=""):
    """(keys, dists, counts) of the GPU equal the oracle's: counts, keys, and
    distance bits of every returned entry."""
    gk, gd, gc = g
    ok, od, oc = o
    assert np.array_equal(gc, oc), f"{what}: counts differ"
    for i in range(len(gc)):
        n = int(gc[i])
        assert np.array_equal(gk[i, :n], ok[i, :n]), f"{what}: keys differ at query {i}"
        assert same_bits(gd[i, :n], od[i, :n]), f"{what}: distances differ at query {i}"
