"""ctypes binding of the C oracle (oracle/liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py's CPU-baseline leg.  The product library never
imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

L2, IP, COSINE = 0, 1, 2
SCALAR, PORTABLE, AVX2 = 0, 1, 3

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
U64 = ctypes.c_uint64
SZ = ctypes.c_size_t


class Graph(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_uint32), ("metric", ctypes.c_uint32),
        ("m", ctypes.c_uint32), ("m0", ctypes.c_uint32),
        ("n", ctypes.c_uint64), ("entry_id", ctypes.c_uint32), ("top", ctypes.c_int32),
        ("vectors", P), ("keys", P), ("levels", P), ("layer0", P),
        ("upper_off", P), ("upper", P), ("deleted", P),
        ("codes", P), ("sq_offset", P), ("sq_scale", ctypes.c_float),
        ("bf16_rows", P), ("pd_codes", P), ("pd_mid", P), ("pd_scale", P), ("pd_rnorm", P),
        ("beam_width", ctypes.c_uint32),
    ]


PREC_F32, PREC_INT8, PREC_BF16, PREC_INT8_PERDIM = 0, 1, 2, 3


class Counters(ctypes.Structure):
    _fields_ = [("n_dist", U64), ("n_expand", U64), ("n_rerank", U64), ("n_upper", U64),
                ("max_cands", U64)]


def _declare(L):
    L.or_mix64.restype = U64
    L.or_mix64.argtypes = [U64]
    L.or_splitmix_next.restype = U64
    L.or_splitmix_next.argtypes = [P]
    L.or_next_gaussian.restype = ctypes.c_double
    L.or_next_gaussian.argtypes = [P]
    L.or_next_below.restype = U64
    L.or_next_below.argtypes = [P, U64]
    L.or_dataset_gaussian.argtypes = [SZ, SZ, SZ, U64, P, P]
    L.or_dataset_clustered.argtypes = [SZ, SZ, SZ, SZ, U64, P, P]
    for f in ("or_dot", "or_l2"):
        getattr(L, f).restype = ctypes.c_float
        getattr(L, f).argtypes = [ctypes.c_int, P, P, SZ]
    L.or_distance.restype = ctypes.c_float
    L.or_distance.argtypes = [ctypes.c_int, ctypes.c_int, P, P, SZ]
    for f in ("or_i8_dot", "or_i8_l2"):
        getattr(L, f).restype = ctypes.c_int64
        getattr(L, f).argtypes = [P, P, SZ]
    L.or_prepare.restype = ctypes.c_int
    L.or_prepare.argtypes = [ctypes.c_int, SZ, P, SZ, P]
    L.or_sq_train.argtypes = [ctypes.c_int, SZ, SZ, P, P, P]
    L.or_sq_encode.argtypes = [P, ctypes.c_float, SZ, P, P]
    L.or_sq_distance.restype = ctypes.c_float
    L.or_sq_distance.argtypes = [ctypes.c_int, ctypes.c_float, P, P, SZ]
    L.or_brute_search.restype = ctypes.c_long
    L.or_brute_search.argtypes = [ctypes.c_int, ctypes.c_int, SZ, SZ, P, P, P, SZ, P, P, P]
    L.or_brute_search_f64.restype = ctypes.c_long
    L.or_brute_search_f64.argtypes = [ctypes.c_int, SZ, SZ, P, P, P, SZ, P, P, P]
    L.or_brute_search_batch.restype = ctypes.c_int
    L.or_brute_search_batch.argtypes = [ctypes.c_int, ctypes.c_int, SZ, SZ, P, P, P, SZ, SZ,
                                        ctypes.c_int, ctypes.c_int, P, P, P]
    L.or_hnsw_search.restype = ctypes.c_long
    L.or_hnsw_search.argtypes = [ctypes.POINTER(Graph), P, SZ, SZ, SZ, ctypes.c_int, P, P, P, P,
                                 ctypes.POINTER(Counters)]
    L.or_hnsw_search_batch.restype = ctypes.c_long
    L.or_hnsw_search_batch.argtypes = [ctypes.POINTER(Graph), P, SZ, SZ, SZ, ctypes.c_int, P,
                                       ctypes.c_int, P, P, P, ctypes.POINTER(Counters),
                                       ctypes.POINTER(SZ)]
    L.or_bf16_round.argtypes = [P, SZ, P]
    L.or_pd_train.argtypes = [SZ, SZ, P, P, P]
    L.or_pd_encode.argtypes = [P, P, SZ, P, P]
    L.or_pd_rnorm.restype = ctypes.c_float
    L.or_pd_rnorm.argtypes = [P, SZ, P]
    L.or_pd_query.argtypes = [ctypes.c_int, SZ, P, P, P, P, P, P]
    L.or_pd_distance.restype = ctypes.c_float
    L.or_pd_distance.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 P, P, SZ]
    L.or_recall_at_k.restype = ctypes.c_double
    L.or_recall_at_k.argtypes = [SZ, P, P, SZ, P, P, SZ, SZ]
    L.or_candidate_cmp.restype = ctypes.c_int
    L.or_candidate_cmp.argtypes = [ctypes.c_float, ctypes.c_uint32, ctypes.c_float, ctypes.c_uint32]


def _p(a):
    return None if a is None else a.ctypes.data


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


# ----------------------------------------------------------------- random
class SplitMix64:
    def __init__(self, seed):
        self.state = np.array([seed], dtype=np.uint64)

    def next_u64(self):
        return int(lib().or_splitmix_next(self.state.ctypes.data))

    def next_gaussian(self):
        return float(lib().or_next_gaussian(self.state.ctypes.data))

    def next_below(self, n):
        return int(lib().or_next_below(self.state.ctypes.data, n))


def mix64(z):
    return int(lib().or_mix64(z))


def dataset_gaussian(base, queries, dim, seed):
    b = np.empty((base, dim), np.float32)
    q = np.empty((queries, dim), np.float32)
    lib().or_dataset_gaussian(base, queries, dim, seed, _p(b), _p(q))
    return b, q


def dataset_clustered(base, queries, dim, clusters, seed):
    b = np.empty((base, dim), np.float32)
    q = np.empty((queries, dim), np.float32)
    lib().or_dataset_clustered(base, queries, dim, clusters, seed, _p(b), _p(q))
    return b, q


# ---------------------------------------------------------------- kernels
def dot(a, b, backend=AVX2):
    a, b = f32(a), f32(b)
    return np.float32(lib().or_dot(backend, _p(a), _p(b), a.size))


def l2(a, b, backend=AVX2):
    a, b = f32(a), f32(b)
    return np.float32(lib().or_l2(backend, _p(a), _p(b), a.size))


def distance(metric, a, b, backend=AVX2):
    a, b = f32(a), f32(b)
    return np.float32(lib().or_distance(metric, backend, _p(a), _p(b), a.size))


def i8_dot(a, b):
    a, b = np.ascontiguousarray(a, np.int8), np.ascontiguousarray(b, np.int8)
    return int(lib().or_i8_dot(_p(a), _p(b), a.size))


def i8_l2(a, b):
    a, b = np.ascontiguousarray(a, np.int8), np.ascontiguousarray(b, np.int8)
    return int(lib().or_i8_l2(_p(a), _p(b), a.size))


def prepare(metric, dim, v):
    """Returns (status, prepared vector)."""
    v = f32(v)
    out = np.empty(max(dim, v.size), np.float32)
    st = lib().or_prepare(metric, dim, _p(v), v.size, _p(out))
    return st, out[:dim]


# -------------------------------------------------------------- quantizer
def sq_train(metric, vectors):
    vectors = f32(vectors)
    rows, dim = vectors.shape
    off = np.empty(dim, np.float32)
    scale = np.empty(1, np.float32)
    lib().or_sq_train(metric, dim, rows, _p(vectors), _p(off), _p(scale))
    return off, np.float32(scale[0])


def sq_encode(offset, scale, v):
    v = f32(v)
    out = np.empty(v.size, np.int8)
    lib().or_sq_encode(_p(f32(offset)), scale, v.size, _p(v), _p(out))
    return out


def sq_encode_rows(offset, scale, rows):
    rows = f32(rows)
    return np.stack([sq_encode(offset, scale, r) for r in rows]) if len(rows) else \
        np.empty((0, rows.shape[1]), np.int8)


def sq_distance(metric, scale, a, b):
    a, b = np.ascontiguousarray(a, np.int8), np.ascontiguousarray(b, np.int8)
    return np.float32(lib().or_sq_distance(metric, scale, _p(a), _p(b), a.size))


def bf16_round(a):
    """f32 -> bf16 (round to nearest even), widened back to f32."""
    a = f32(a)
    out = np.empty_like(a)
    lib().or_bf16_round(_p(a), a.size, _p(out))
    return out


# ------------------------------------------- per-dimension int8 quantizer
def pd_train(vectors):
    vectors = f32(vectors)
    rows, dim = vectors.shape
    mid = np.empty(dim, np.float32)
    scale = np.empty(dim, np.float32)
    lib().or_pd_train(dim, rows, _p(vectors), _p(mid), _p(scale))
    return mid, scale


def pd_encode_rows(mid, scale, rows):
    rows = f32(rows)
    mid, scale = f32(mid), f32(scale)
    out = np.empty(rows.shape, np.int8)
    for i in range(rows.shape[0]):
        lib().or_pd_encode(_p(mid), _p(scale), rows.shape[1], _p(rows[i]), _p(out[i]))
    return out


def pd_rnorm_rows(scale, codes):
    scale = f32(scale)
    codes = np.ascontiguousarray(codes, np.int8)
    return np.array([lib().or_pd_rnorm(_p(scale), codes.shape[1], _p(codes[i]))
                     for i in range(codes.shape[0])], np.float32)


def pd_query(metric, mid, scale, q):
    mid, scale, q = f32(mid), f32(scale), f32(q)
    code = np.empty(q.size, np.int8)
    sg = np.empty(1, np.float32)
    qc = np.empty(1, np.float32)
    lib().or_pd_query(metric, q.size, _p(mid), _p(scale), _p(q), _p(code), _p(sg), _p(qc))
    return code, np.float32(sg[0]), np.float32(qc[0])


def pd_distance(metric, sigma, qconst, rnorm, qcode, code):
    qcode = np.ascontiguousarray(qcode, np.int8)
    code = np.ascontiguousarray(code, np.int8)
    return np.float32(lib().or_pd_distance(metric, sigma, qconst, rnorm, _p(qcode), _p(code),
                                           code.size))


# ------------------------------------------------------------ brute force
def brute_search(metric, keys, vectors, query, k, allow=None, backend=AVX2):
    """vectors are prepared rows. Returns (keys, dists) or raises ValueError(status)."""
    keys = np.ascontiguousarray(keys, np.uint64)
    vectors = f32(vectors)
    n, dim = vectors.shape
    query = f32(query)
    kk = min(k, n)
    ok = np.empty(max(kk, 1), np.uint64)
    od = np.empty(max(kk, 1), np.float32)
    al = None if allow is None else np.ascontiguousarray(allow, np.uint8)
    r = lib().or_brute_search(metric, backend, dim, n, _p(keys), _p(vectors), _p(query), k,
                              _p(al), _p(ok), _p(od))
    if r < 0:
        raise ValueError(-r)
    return ok[:r].copy(), od[:r].copy()


def brute_search_f64(metric, keys, vectors, query, k, allow=None):
    keys = np.ascontiguousarray(keys, np.uint64)
    vectors = f32(vectors)
    n, dim = vectors.shape
    query = f32(query)
    kk = min(k, n)
    ok = np.empty(max(kk, 1), np.uint64)
    od = np.empty(max(kk, 1), np.float64)
    al = None if allow is None else np.ascontiguousarray(allow, np.uint8)
    r = lib().or_brute_search_f64(metric, dim, n, _p(keys), _p(vectors), _p(query), k, _p(al),
                                  _p(ok), _p(od))
    if r < 0:
        raise ValueError(-r)
    return ok[:r].copy(), od[:r].copy()


def brute_search_batch(metric, keys, vectors, queries, k, f64=False, threads=8, backend=AVX2):
    keys = np.ascontiguousarray(keys, np.uint64)
    vectors = f32(vectors)
    queries = f32(queries)
    n, dim = vectors.shape
    nq = queries.shape[0]
    ok = np.zeros((nq, max(k, 1)), np.uint64)
    od = np.zeros((nq, max(k, 1)), np.float64)
    oc = np.zeros(nq, np.uint32)
    st = lib().or_brute_search_batch(metric, backend, dim, n, _p(keys), _p(vectors), _p(queries),
                                     nq, k, int(f64), threads, _p(ok), _p(od), _p(oc))
    if st:
        raise ValueError(st)
    return ok[:, :k], od[:, :k], oc


# ------------------------------------------------------------ HNSW search
class OracleGraph:
    """Holds numpy arrays in the crate's storage layout and the ctypes view."""

    def __init__(self, dim, metric, m, m0, vectors, keys, levels, layer0, upper_off, upper,
                 entry_id, top, deleted=None, codes=None, sq_offset=None, sq_scale=1.0,
                 bf16=False, perdim=None, beam_width=1):
        """bf16: also traverse bf16-rounded rows (precision 2).  perdim:
        (mid, scale) of a per-dimension quantizer; its codes and row norms
        are computed here from the rows (precision 3).  beam_width: the
        candidates expanded per step by those two traversals (1: the crate's
        rule)."""
        vectors = f32(vectors)
        bf16_rows = bf16_round(vectors) if bf16 else None
        pd_codes = pd_mid = pd_scale = pd_rnorm = None
        if perdim is not None:
            pd_mid, pd_scale = f32(perdim[0]), f32(perdim[1])
            pd_codes = pd_encode_rows(pd_mid, pd_scale, vectors)
            pd_rnorm = pd_rnorm_rows(pd_scale, pd_codes)
        self._keep = dict(
            vectors=f32(vectors), keys=np.ascontiguousarray(keys, np.uint64),
            levels=np.ascontiguousarray(levels, np.uint8),
            layer0=np.ascontiguousarray(layer0, np.uint32),
            upper_off=np.ascontiguousarray(upper_off, np.uint64),
            upper=np.ascontiguousarray(upper if len(upper) else np.zeros(1), np.uint32),
            deleted=None if deleted is None else np.ascontiguousarray(deleted, np.uint8),
            codes=None if codes is None else np.ascontiguousarray(codes, np.int8),
            sq_offset=None if sq_offset is None else f32(sq_offset),
            bf16_rows=bf16_rows, pd_codes=pd_codes, pd_mid=pd_mid, pd_scale=pd_scale,
            pd_rnorm=pd_rnorm)
        k = self._keep
        self.g = Graph(dim, metric, m, m0, len(k["keys"]), entry_id, top,
                       _p(k["vectors"]), _p(k["keys"]), _p(k["levels"]), _p(k["layer0"]),
                       _p(k["upper_off"]), _p(k["upper"]), _p(k["deleted"]), _p(k["codes"]),
                       _p(k["sq_offset"]), float(sq_scale), _p(k["bf16_rows"]),
                       _p(k["pd_codes"]), _p(k["pd_mid"]), _p(k["pd_scale"]), _p(k["pd_rnorm"]),
                       int(beam_width))
        self.dim = dim

    def set_beam_width(self, width):
        self.g.beam_width = int(width)

    @staticmethod
    def _prec(quantized, precision):
        return int(precision) if precision is not None else (PREC_INT8 if quantized else PREC_F32)

    def search(self, query, k, ef, quantized=False, allow=None, precision=None):
        query = f32(query)
        ok = np.empty(max(k, 1), np.uint64)
        od = np.empty(max(k, 1), np.float32)
        oi = np.empty(max(k, 1), np.uint32)
        cnt = Counters()
        al = None if allow is None else np.ascontiguousarray(allow, np.uint8)
        r = lib().or_hnsw_search(ctypes.byref(self.g), _p(query), query.size, k, ef,
                                 self._prec(quantized, precision), _p(al), _p(ok), _p(od), _p(oi),
                                 ctypes.byref(cnt))
        if r < 0:
            raise ValueError(-r)
        return ok[:r].copy(), od[:r].copy(), oi[:r].copy(), cnt

    def search_batch(self, queries, k, ef, quantized=False, allow=None, threads=8,
                     precision=None):
        queries = f32(queries)
        nq = queries.shape[0]
        ok = np.zeros((nq, max(k, 1)), np.uint64)
        od = np.zeros((nq, max(k, 1)), np.float32)
        oc = np.zeros(nq, np.uint32)
        cnt = Counters()
        bad = SZ(0)
        al = None if allow is None else np.ascontiguousarray(allow, np.uint8)
        r = lib().or_hnsw_search_batch(ctypes.byref(self.g), _p(queries), nq, k, ef,
                                       self._prec(quantized, precision), _p(al), threads, _p(ok),
                                       _p(od), _p(oc),
                                       ctypes.byref(cnt), ctypes.byref(bad))
        if r < 0:
            raise ValueError(-r, bad.value)
        return ok[:, :k], od[:, :k], oc, cnt


def recall_at_k(truth_keys, truth_count, found_keys, found_count, k):
    tk = np.ascontiguousarray(truth_keys, np.uint64)
    fk = np.ascontiguousarray(found_keys, np.uint64)
    tc = np.ascontiguousarray(truth_count, np.uint32)
    fc = np.ascontiguousarray(found_count, np.uint32)
    nq = len(tc)
    return float(lib().or_recall_at_k(nq, _p(tk), _p(tc), tk.shape[1] if tk.ndim == 2 else 0,
                                      _p(fk), _p(fc), fk.shape[1] if fk.ndim == 2 else 0, k))


def candidate_cmp(da, ia, db, ib):
    return int(lib().or_candidate_cmp(da, ia, db, ib))
