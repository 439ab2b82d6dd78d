"""Shared test helpers: build graphs with the host library and hand the
same graph to the oracle."""
import numpy as np

import hnsw_amd as H
import oracle as O


def oracle_graph(idx):
    """The oracle's view of the index's current graph (arrays copied, so the
    view outlives later mutation of the index)."""
    v = {k: (np.array(x, copy=True) if isinstance(x, np.ndarray) else x)
         for k, x in idx.graph_arrays().items()}
    return O.OracleGraph(v["dim"], v["metric"], v["m"], v["m0"], v["vectors"], v["keys"],
                         v["levels"], v["layer0"], v["upper_off"], v["upper"], v["entry_id"],
                         v["top"], deleted=v["deleted"], codes=v["codes"],
                         sq_offset=v["sq_offset"], sq_scale=v["sq_scale"])


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
