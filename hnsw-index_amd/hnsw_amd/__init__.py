"""Python mirror of the hnsw-index crate's public API over the MI355X engine.

Names, argument meaning and error behaviour follow the crate
(aozmekik/hnsw-index): `HnswIndex` (src/hnsw/mod.rs), `BruteForceIndex`
(src/brute.rs), `SearchOptions`, `Metric`, `Neighbor`, `HnswParams`,
`NeighborSelection`, `Error` variants (src/error.rs), `eval.recall_at_k`,
`synthetic.Dataset`, and `kernels.Backend` with the new `Backend.HIP`.

Graph construction (`insert`, `insert_batch`) runs on the CPU, as in the
crate; every search runs on the GPU through libhnsw_amd.so.  There is no CPU
search path: without a GPU, searching raises `DeviceError`.
"""
import ctypes
import enum
import os
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import lib

__all__ = [
    "Backend", "BruteForceIndex", "Error", "HnswIndex", "HnswParams", "Metric", "Neighbor",
    "NeighborSelection", "SearchOptions", "Precision", "Server", "eval", "synthetic", "MAX_DIM",
    "MAX_NODES",
]

MAX_DIM = 1 << 16           # src/lib.rs:70
MAX_NODES = (1 << 32) - 2   # src/lib.rs:74
# candidates per beam step of the bf16 / per-dimension int8 traversals
DEFAULT_BEAM_WIDTH = 4


# ---------------------------------------------------------------- errors
class Error(Exception):
    """Base of every error (src/error.rs:9-51); `.status` is the C ABI code."""
    status = -1

    def __init__(self, msg="", **fields):
        super().__init__(msg)
        for k, v in fields.items():
            setattr(self, k, v)


class DimensionMismatch(Error): status = 1
class InvalidDimension(Error): status = 2
class InvalidParameter(Error): status = 3
class NonFiniteVector(Error): status = 4
class ZeroVector(Error): status = 5
class DuplicateKey(Error): status = 6
class CapacityExceeded(Error): status = 7
class QuantizationDisabled(Error): status = 8
class FormatError(Error):
    """Error::Format (src/error.rs:56-75): `.kind` is one of "BadMagic",
    "UnsupportedVersion", "Truncated", "ChecksumMismatch", "Invalid",
    "TrailingBytes"; `.what`, `.version`, `.stored`, `.computed` as the
    variant carries."""
    status = 9


class MemoryLimitExceeded(Error):
    """Error::MemoryLimitExceeded: `.required` and `.limit` in bytes."""
    status = 10


class DeviceError(Error): status = 100
class OutOfMemory(Error): status = 101
class Unsupported(Error): status = 102
class InvalidArgument(Error): status = 103


_BY_STATUS = {c.status: c for c in (DimensionMismatch, InvalidDimension, InvalidParameter,
                                     NonFiniteVector, ZeroVector, DuplicateKey,
                                     CapacityExceeded, QuantizationDisabled, FormatError,
                                     MemoryLimitExceeded, DeviceError,
                                     OutOfMemory, Unsupported, InvalidArgument)}


def _check(st, **fields):
    if st:
        raise _BY_STATUS.get(st, Error)(L.last_error(), **fields)


# ------------------------------------------------------------ plain types
class Metric(enum.IntEnum):
    """src/metric.rs:11-30 (tags as :44-50)."""
    L2 = 0
    InnerProduct = 1
    Cosine = 2


class Backend(enum.Enum):
    """kernels::Backend (src/kernels/mod.rs:55-66) plus the HIP engine."""
    Scalar = "scalar"
    Portable = "portable"
    Neon = "neon"
    Avx2 = "avx2"
    Hip = "hip"

    def is_available(self):
        if self is Backend.Hip:
            return lib.hh_device_count() > 0
        if self is Backend.Neon:
            return False
        return True


class Precision(enum.IntEnum):
    """Traversal precision (hh_precision).  INT8 is the crate's quantized
    traversal; BF16 and INT8_PERDIM (per-dimension min/max codes) have no
    crate counterpart.  Every mode but F32 re-ranks exactly."""
    F32 = 0
    INT8 = 1
    BF16 = 2
    INT8_PERDIM = 3


@dataclass(frozen=True)
class Neighbor:
    """src/neighbor.rs:7-12."""
    key: int
    distance: float


@dataclass(frozen=True)
class NeighborSelection:
    """src/params.rs:19-46."""
    heuristic: bool = True
    extend_candidates: bool = False
    keep_pruned: bool = False

    @staticmethod
    def simple():
        return NeighborSelection(False)

    @staticmethod
    def heuristic_with(extend_candidates=False, keep_pruned=False):
        return NeighborSelection(True, extend_candidates, keep_pruned)


class HnswParams:
    """src/params.rs:61-161."""

    def __init__(self, m=16):
        p = L.Params()
        lib.hh_params_default(m, ctypes.byref(p))
        self._p = p
        if m < 2:  # keep what the crate keeps; validation happens at HnswIndex()
            self._p.m0 = m * 2

    def with_m0(self, m0):
        self._p.m0 = m0
        return self

    def with_ef_construction(self, ef):
        self._p.ef_construction = ef
        return self

    def with_level_multiplier(self, ml):
        self._p.level_multiplier = ml
        return self

    def with_selection(self, sel):
        self._p.selection = 1 if sel.heuristic else 0
        self._p.extend_candidates = int(sel.extend_candidates)
        self._p.keep_pruned = int(sel.keep_pruned)
        return self

    def with_seed(self, seed):
        self._p.seed = seed
        return self

    def m(self): return self._p.m
    def m0(self): return self._p.m0
    def ef_construction(self): return self._p.ef_construction
    def level_multiplier(self): return self._p.level_multiplier
    def seed(self): return self._p.seed


class SearchOptions:
    """src/hnsw/mod.rs:56-98, plus `precision` for the bf16 traversal mode."""

    def __init__(self, ef):
        self._ef = ef
        self._filter = None
        self._precision = Precision.F32

    def filter(self, f):
        self._filter = f
        return self

    def quantized(self, q):
        self._precision = Precision.INT8 if q else Precision.F32
        return self

    def precision(self, p):
        self._precision = Precision(p)
        return self

    def ef(self):
        return self._ef


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _bits_from_mask(mask):
    """bool array over ids -> uint32 words, bit i of word i/32 = mask[i]."""
    mask = np.ascontiguousarray(mask, dtype=bool)
    n = mask.size
    padded = np.zeros(((n + 31) // 32) * 32, dtype=bool)
    padded[:n] = mask
    return np.packbits(padded, bitorder="little").view(np.uint32).copy()


def _results(keys, dists, counts):
    return [[Neighbor(int(keys[i, j]), float(dists[i, j])) for j in range(int(counts[i]))]
            for i in range(len(counts))]


# ----------------------------------------------------------------- index
class HnswIndex:
    """The HNSW index (src/hnsw/mod.rs:148-172): built on the CPU, searched
    on the GPU.  The device copy is uploaded on the first search after a
    mutation."""

    def __init__(self, dim, metric=Metric.L2, params=None, device=0):
        params = params or HnswParams()
        h = ctypes.c_void_p()
        _check(lib.hh_graph_new(dim, int(metric), ctypes.byref(params._p), ctypes.byref(h)))
        self._g = h
        self._dim = dim
        self._metric = Metric(metric)
        self._params = params
        self._device = device
        self._idx = None
        self._dirty = True
        self._stale = False
        self._bf16 = False
        self._perdim = False
        self._beam = DEFAULT_BEAM_WIDTH
        self._servers = 0

    def __del__(self):
        try:
            if getattr(self, "_idx", None):
                lib.hh_index_free(self._idx)
            if getattr(self, "_g", None):
                lib.hh_graph_free(self._g)
        except Exception:
            pass

    # -- persistence (src/hnsw/mod.rs:611-700, src/hnsw/io.rs)
    def to_bytes(self):
        """The crate's version-1 index file, byte for byte."""
        n = ctypes.c_uint64()
        _check(lib.hh_graph_to_bytes(self._g, None, 0, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint8)
        _check(lib.hh_graph_to_bytes(self._g, buf.ctypes.data, n.value, ctypes.byref(n)))
        return buf.tobytes()

    def write_to(self, f):
        data = self.to_bytes()
        if isinstance(f, (str, os.PathLike)):
            with open(f, "wb") as fh:
                fh.write(data)
        else:
            f.write(data)

    @classmethod
    def from_bytes(cls, data, max_memory=None, device=0):
        """Loads an index file, validating it as untrusted input; raises
        FormatError or MemoryLimitExceeded (max_memory None = the crate's
        default limit, 64x the input size and at least 64 MiB)."""
        buf = np.frombuffer(bytes(data), np.uint8)
        h = ctypes.c_void_p()
        info = L.FormatInfo()
        limit = (1 << 64) - 1 if max_memory is None else min(max(int(max_memory), 0), (1 << 64) - 2)
        st = lib.hh_graph_from_bytes(buf.ctypes.data if buf.size else None, buf.size, limit,
                                     ctypes.byref(h), ctypes.byref(info))
        if st:
            kinds = {1: "BadMagic", 2: "UnsupportedVersion", 3: "Truncated",
                     4: "ChecksumMismatch", 5: "Invalid", 6: "TrailingBytes"}
            _check(st, kind=kinds.get(info.kind), what=info.what.decode(), version=info.version,
                   stored=info.stored, computed=info.computed, required=info.required,
                   limit=info.limit)
        self = cls.__new__(cls)
        dim, metric = ctypes.c_uint32(), ctypes.c_uint32()
        p = L.Params()
        _check(lib.hh_graph_info(h, ctypes.byref(dim), ctypes.byref(metric), ctypes.byref(p)))
        params = HnswParams(p.m)
        params._p = p
        self._g = h
        self._dim = dim.value
        self._metric = Metric(metric.value)
        self._params = params
        self._device = device
        self._idx = None
        self._dirty = True
        self._stale = False
        self._bf16 = False
        self._perdim = False
        self._beam = DEFAULT_BEAM_WIDTH
        self._servers = 0
        return self

    @classmethod
    def read_from(cls, f, **kw):
        if isinstance(f, (str, os.PathLike)):
            with open(f, "rb") as fh:
                return cls.from_bytes(fh.read(), **kw)
        return cls.from_bytes(f.read(), **kw)

    def memory_usage(self):
        return int(lib.hh_graph_memory_usage(self._g))

    # -- properties (src/hnsw/mod.rs:210-264)
    def dim(self): return self._dim
    def metric(self): return self._metric
    def params(self): return self._params
    def backend(self): return Backend.Hip
    def __len__(self): return int(lib.hh_graph_len(self._g))
    def len(self): return len(self)
    def is_empty(self): return len(self) == 0
    def node_count(self): return int(lib.hh_graph_node_count(self._g))
    def deleted_len(self): return self.node_count() - len(self)

    def contains(self, key):
        return key in self._key_map()

    def _key_map(self):
        v = self.graph_arrays()
        live = ~v["deleted"].astype(bool)
        return set(v["keys"][live].tolist())

    # -- mutation (CPU, src/hnsw/mod.rs:290-400)
    def _refuse_if_served(self):
        """A Server launches searches on the resident index: while one is
        open the index (host graph and device copy) must not change."""
        if self._servers > 0:
            raise InvalidArgument("a Server is open on this index: close it before changing "
                                  "the index")

    def insert(self, key, vector):
        self.insert_batch([key], np.asarray(vector, np.float32).reshape(1, -1), threads=1)

    def insert_batch(self, keys, vectors, threads=None, device=False):
        """HnswIndex::insert_batch (src/hnsw/mod.rs:290-349).  With device,
        the layer-0 searches of INSERT run on the GPU (see _link_on_device);
        the graph is of the same kind as a parallel host build, not equal to
        it."""
        self._refuse_if_served()
        keys = np.ascontiguousarray(keys, np.uint64)
        vectors = _f32(vectors)
        if vectors.size != keys.size * self._dim:
            raise DimensionMismatch(
                f"dimension mismatch: expected {keys.size * self._dim}, got {vectors.size}",
                expected=keys.size * self._dim, actual=vectors.size)
        if threads is None:
            threads = min(16, os.cpu_count() or 1)
        bad = ctypes.c_uint64(0)
        if device:
            first = ctypes.c_uint64()
            _check(lib.hh_graph_append(self._g, _ptr(keys), _ptr(vectors), keys.size,
                                       ctypes.byref(first), ctypes.byref(bad)), key=bad.value)
            self._mark_changed()
            self._link_on_device(first.value, keys.size, threads)
            return
        st = lib.hh_graph_insert_batch(self._g, _ptr(keys), _ptr(vectors), keys.size, threads,
                                       ctypes.byref(bad))
        _check(st, key=bad.value)
        self._mark_changed()

    def _mark_changed(self):
        if self._idx is not None and not self._dirty:
            self._stale = True  # the resident index takes just the changes
        else:
            self._dirty = True

    def _link_on_device(self, first, count, threads, host_prefix=4096, fraction=32):
        """Links appended nodes [first, first + count) in rounds.  While the
        graph is small (below host_prefix linked nodes) the host INSERT links
        them; after that each round takes up to 1/fraction of the linked
        graph: the resident index is synced (the round's rows arrive with no
        links, so no search can reach them), the GPU searches each row as
        INSERT does at layer 0 (hh_index_build_candidates), and the host runs
        the selection and linking on those candidates (hh_graph_link_batch;
        nodes of level >= 1 take the full host INSERT)."""
        ef = max(int(self._params._p.ef_construction), int(self._params._p.m))
        done = 0
        while done < count:
            linked = first + done
            start = first + done
            if linked < host_prefix:
                take = min(count - done, host_prefix - linked)
                _check(lib.hh_graph_link_batch(self._g, start, take, None, None, None, 0, threads))
            else:
                take = min(count - done, max(256, linked // fraction))
                h = self.device_index()
                ids = np.zeros((take, ef), np.uint32)
                d = np.zeros((take, ef), np.float32)
                c = np.zeros(take, np.uint32)
                _check(lib.hh_index_build_candidates(h, start, take, ef, _ptr(ids), _ptr(d),
                                                     _ptr(c)))
                _check(lib.hh_graph_link_batch(self._g, start, take, _ptr(ids), _ptr(d), _ptr(c),
                                               ef, threads))
            self._mark_changed()
            done += take

    def remove(self, key):
        self._refuse_if_served()
        r = ctypes.c_int(0)
        _check(lib.hh_graph_remove(self._g, key, ctypes.byref(r)))
        if r.value and self._idx and not self._dirty and not self._stale:
            v = self.graph_arrays()
            _check(lib.hh_index_update_deleted(self._idx, _ptr(v["deleted"]), len(self)))
        return bool(r.value)

    def enable_quantization(self, on_device=False):
        """HnswIndex::enable_quantization (src/hnsw/mod.rs:544-547).  With
        on_device the quantizer is trained and the codes encoded by the GPU
        over the resident rows (bit-identical parameters and codes), and the
        host graph receives both, so later inserts encode with it as the
        crate does."""
        self._refuse_if_served()
        if not on_device:
            _check(lib.hh_graph_enable_quantization(self._g))
            self._dirty = True
            return
        idx = self.device_index()
        off = np.zeros(self._dim, np.float32)
        sc = ctypes.c_float()
        _check(lib.hh_index_quantize(idx, _ptr(off), ctypes.byref(sc)))
        codes = np.zeros((self.node_count(), self._dim), np.int8)
        if codes.size:
            _check(lib.hh_index_download_codes(idx, _ptr(codes)))
        _check(lib.hh_graph_set_quantizer(self._g, _ptr(off), sc.value,
                                          _ptr(codes) if codes.size else None))
        # the resident index already holds exactly these codes

    def quantizer(self):
        """(offset, scale) of the trained quantizer, or None."""
        v = self.graph_arrays()
        if v["codes"] is None:
            return None
        return v["sq_offset"].copy(), float(v["sq_scale"])

    def disable_quantization(self):
        self._refuse_if_served()
        _check(lib.hh_graph_disable_quantization(self._g))
        self._dirty = True

    def enable_quantization_perdim(self):
        """Trains the per-dimension min/max int8 quantizer on the resident
        rows and encodes them on the GPU (hh_index_quantize_perdim; no crate
        counterpart).  Searches with Precision.INT8_PERDIM then traverse on
        those codes and re-rank exactly.  Inserts synced later are encoded
        with the same parameters; a full re-upload of the index trains it
        again.  Returns (mid, scale), dim values each."""
        self._refuse_if_served()
        idx = self.device_index()
        mid = np.zeros(self._dim, np.float32)
        scale = np.zeros(self._dim, np.float32)
        _check(lib.hh_index_quantize_perdim(idx, _ptr(mid), _ptr(scale)))
        self._perdim = True
        return mid, scale

    def perdim_codes(self):
        """(codes n x dim, row terms n or None) of the resident per-dimension
        quantizer."""
        idx = self.device_index()
        n = self.node_count()
        codes = np.zeros((n, self._dim), np.int8)
        rn = np.zeros(n, np.float32) if self._metric == Metric.L2 else None
        if n:
            _check(lib.hh_index_download_perdim(idx, _ptr(codes), _ptr(rn)))
        return codes, rn

    def bf16_rows(self):
        """The resident bf16 rows (n x dim uint16 bit patterns)."""
        if not self._bf16:
            self.enable_bf16()
        out = np.zeros((self.node_count(), self._dim), np.uint16)
        if out.size:
            _check(lib.hh_index_download_bf16(self.device_index(), _ptr(out)))
        return out

    def is_quantized(self):
        return bool(self._view().codes)

    # -- graph access
    def _view(self):
        d = L.GraphDesc()
        _check(lib.hh_graph_view(self._g, ctypes.byref(d)))
        return d

    def graph_arrays(self):
        """Borrowed numpy views of the host graph in the crate's layout."""
        d = self._view()
        n, dim, m, m0 = d.n, d.dim, d.m, d.m0

        def arr(ptr, ctype, count):
            if not ptr or count == 0:
                return np.zeros(0, dtype=np.dtype(ctype))
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (count,))

        levels = arr(d.levels, ctypes.c_uint8, n)
        upper_off = arr(d.upper_off, ctypes.c_uint64, n)
        words = int((upper_off.astype(np.int64) + levels.astype(np.int64) * (m + 1)).max()) if n else 0
        return dict(
            dim=dim, metric=d.metric, m=m, m0=m0, n=n, entry_id=d.entry_id, top=d.top_level,
            vectors=arr(d.vectors, ctypes.c_float, n * dim).reshape(n, dim),
            keys=arr(d.keys, ctypes.c_uint64, n), levels=levels,
            layer0=arr(d.layer0, ctypes.c_uint32, n * (m0 + 1)), upper_off=upper_off,
            upper=arr(d.upper, ctypes.c_uint32, words), deleted=arr(d.deleted, ctypes.c_uint8, n),
            codes=arr(d.codes, ctypes.c_int8, n * dim).reshape(-1, dim) if d.codes else None,
            sq_offset=arr(d.sq_offset, ctypes.c_float, dim) if d.codes else None,
            sq_scale=d.sq_scale, live=d.live)

    # -- device residency
    def device_index(self):
        """The uploaded hh_index handle (uploading first if needed)."""
        if (self._dirty or self._stale or self._idx is None) and self._servers > 0:
            self._refuse_if_served()
        if self._dirty or self._idx is None:
            if self._idx:
                lib.hh_index_free(self._idx)
                self._idx = None
            d = self._view()
            h = ctypes.c_void_p()
            _check(lib.hh_index_upload(ctypes.byref(d), self._device, ctypes.byref(h)))
            self._idx = h
            self._dirty = False
            self._bf16 = False
            self._take_dirty()  # the upload carried every link
            self._stale = False
            if self._perdim:
                _check(lib.hh_index_quantize_perdim(h, None, None))
            _check(lib.hh_index_set_beam_width(h, self._beam))
        elif self._stale:
            # inserts and removals since the upload: move only the changes
            ids = self._take_dirty()
            d = self._view()
            st = lib.hh_index_sync(self._idx, ctypes.byref(d), _ptr(ids) if ids.size else None,
                                   ids.size)
            self._stale = False
            if st:
                self._dirty = True
                return self.device_index()
        return self._idx

    def _take_dirty(self):
        """Ids of the nodes whose links changed since the last call."""
        cnt = ctypes.c_uint64()
        _check(lib.hh_graph_take_dirty(self._g, None, 0, ctypes.byref(cnt)))
        ids = np.zeros(max(cnt.value, 1), np.uint32)
        _check(lib.hh_graph_take_dirty(self._g, _ptr(ids), cnt.value, ctypes.byref(cnt)))
        return ids[:cnt.value]

    def set_query_order(self, longest_first=True):
        """Large batches (1-16 queries per resident wavefront) run in two
        passes with the longest queries first (default), or in one pass in
        query order; results are the same either way."""
        _check(lib.hh_index_set_query_order(self.device_index(), int(bool(longest_first))))

    def set_beam_width(self, width):
        """Candidates per layer-0 step of the bf16 and per-dimension int8
        traversals (1: the crate's one-at-a-time rule; 2-8: beam steps, at
        most 256 neighbour slots per step).  f32 and the crate's int8 codes
        always follow the crate's rule."""
        width = int(width)
        if not 1 <= width <= 8:
            raise InvalidArgument("beam width must be 1..8")
        if self._idx is not None and not self._dirty:
            self._refuse_if_served()
            _check(lib.hh_index_set_beam_width(self._idx, width))
        self._beam = width

    def beam_width(self):
        return self._beam

    def enable_bf16(self):
        if not self._bf16:
            self._refuse_if_served()
        _check(lib.hh_index_enable_bf16(self.device_index()))
        self._bf16 = True

    def search_stats(self, reset=False):
        s = L.SearchStats()
        _check(lib.hh_index_stats(self.device_index(), ctypes.byref(s)))
        if reset:
            _check(lib.hh_index_reset_stats(self.device_index()))
        return {f: getattr(s, f) for f, _ in L.SearchStats._fields_}

    # -- search (src/hnsw/mod.rs:402-536), on the GPU
    def _allow_bits(self, flt):
        if flt is None:
            return None
        v = self.graph_arrays()
        if isinstance(flt, np.ndarray) and flt.dtype == bool:
            mask = flt  # by node id
        else:
            keys = v["keys"]
            mask = np.fromiter((bool(flt(int(kk))) for kk in keys), dtype=bool, count=keys.size)
        return _bits_from_mask(mask)

    def search_arrays(self, queries, k, ef, precision=Precision.F32, filter=None):
        """Batched search into padded arrays: (keys nq x k, dists nq x k, counts nq)."""
        queries = _f32(queries)
        if queries.size % self._dim:
            raise DimensionMismatch(
                f"dimension mismatch: expected {self._dim}, got {queries.size}",
                expected=self._dim, actual=queries.size)
        nq = queries.size // self._dim
        if precision == Precision.BF16 and not self._bf16:
            self.device_index()
            self.enable_bf16()
        idx = self.device_index()
        kk = max(int(k), 1)
        ok = np.zeros((nq, kk), np.uint64)
        od = np.zeros((nq, kk), np.float32)
        oc = np.zeros(nq, np.uint32)
        allow = self._allow_bits(filter)
        bad = ctypes.c_uint64(0)
        st = lib.hh_search_batch(idx, _ptr(queries), nq, int(k), int(ef), int(precision),
                                 _ptr(allow), _ptr(ok), _ptr(od), _ptr(oc), ctypes.byref(bad))
        _check(st, query=bad.value)
        return ok[:, :k], od[:, :k], oc

    def search_mixed(self, queries, ks, efs, precisions=None, filter_ids=None, filters=None):
        """One launch for queries with their own k, ef, precision and filter
        (filter_ids[i] indexes `filters`, a list of bool masks over node ids,
        or -1 for none).  Returns (keys, dists, counts, status) padded to
        max(ks); an invalid query gets its status and no results."""
        queries = _f32(queries).reshape(-1, self._dim)
        nq = queries.shape[0]
        specs = np.zeros((nq, 4), np.uint32)
        specs[:, 0] = ks
        specs[:, 1] = efs
        specs[:, 2] = 0 if precisions is None else np.asarray(precisions)
        specs[:, 3] = 0xFFFFFFFF if filter_ids is None else \
            np.where(np.asarray(filter_ids) < 0, 0xFFFFFFFF, np.asarray(filter_ids))
        if precisions is not None and (np.asarray(precisions) == Precision.BF16).any() and not self._bf16:
            self.device_index()
            self.enable_bf16()
        idx = self.device_index()
        kst = max(int(np.max(ks)) if nq else 1, 1)
        fl = None
        nf = 0
        if filters:
            fl = np.ascontiguousarray(np.stack([_bits_from_mask(m) for m in filters]))
            nf = len(filters)
        ok = np.zeros((nq, kst), np.uint64)
        od = np.zeros((nq, kst), np.float32)
        oc = np.zeros(nq, np.uint32)
        ost = np.zeros(nq, np.int32)
        _check(lib.hh_search_batch_spec(idx, _ptr(queries), nq, _ptr(specs), kst, _ptr(fl), nf,
                                        _ptr(ok), _ptr(od), _ptr(oc), _ptr(ost)))
        # slots past each query's count are not written by the library
        past = np.arange(kst)[None, :] >= oc[:, None]
        ok[past] = 0
        od[past] = 0.0
        return ok, od, oc, ost

    def search_batch(self, queries, k, ef):
        queries = _f32(queries)
        if queries.size % self._dim:
            raise DimensionMismatch(
                f"dimension mismatch: expected {self._dim}, got {queries.size}",
                expected=self._dim, actual=queries.size)
        return _results(*self.search_arrays(queries, k, ef))

    def _single(self, query, k, ef, precision, flt):
        q = _f32(query).ravel()
        if q.size != self._dim:
            raise DimensionMismatch(f"dimension mismatch: expected {self._dim}, got {q.size}",
                                    expected=self._dim, actual=q.size)
        return _results(*self.search_arrays(q.reshape(1, -1), k, ef, precision, flt))[0]

    def search(self, query, k, ef):
        return self._single(query, k, ef, Precision.F32, None)

    def search_filtered(self, query, k, ef, filter):
        return self._single(query, k, ef, Precision.F32, filter)

    def search_with(self, query, k, options):
        return self._single(query, k, options._ef, options._precision, options._filter)


class Server:
    """Continuous-batching request loop over an index's device copy
    (hh_serve_*): submit single queries with their own k, ef, precision and
    registered filter; results come back per ticket with their latency."""

    def __init__(self, index, k_max=100, max_batch=0, max_inflight=2, max_filters=8,
                 max_queue=0):
        self._index = index
        self._dim = index.dim()
        cfg = L.ServeConfig(max_batch, max_inflight, k_max, max_filters, max_queue)
        h = ctypes.c_void_p()
        self._h = None
        _check(lib.hh_serve_open(index.device_index(), ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self._k_max = k_max
        # the index refuses changes (and keeps its device copy) while open
        index._servers += 1

    def register_filter(self, mask):
        """mask: bool array over node ids. Returns the filter id."""
        fid = ctypes.c_uint32()
        # the bitmap stays referenced until the library has copied it
        bits = _bits_from_mask(mask)
        _check(lib.hh_serve_register_filter(self._h, _ptr(bits), ctypes.byref(fid)))
        return fid.value

    def submit(self, query, k, ef, precision=Precision.F32, filter_id=-1):
        q = _f32(query).ravel()
        if q.size != self._dim:
            raise DimensionMismatch(f"dimension mismatch: expected {self._dim}, got {q.size}",
                                    expected=self._dim, actual=q.size)
        spec = L.QuerySpec(int(k), int(ef), int(precision),
                           0xFFFFFFFF if filter_id < 0 else int(filter_id))
        t = ctypes.c_uint64()
        _check(lib.hh_serve_submit(self._h, _ptr(q), ctypes.byref(spec), ctypes.byref(t)))
        return t.value

    def poll(self, ticket, wait=True):
        """(keys, dists, status, latency_ms) or None while pending."""
        keys = np.zeros(self._k_max, np.uint64)
        dist = np.zeros(self._k_max, np.float32)
        cnt = ctypes.c_uint32()
        st = ctypes.c_int32()
        lat = ctypes.c_double()
        done = ctypes.c_int()
        _check(lib.hh_serve_poll(self._h, ticket, int(wait), _ptr(keys), _ptr(dist),
                                 ctypes.byref(cnt), ctypes.byref(st), ctypes.byref(lat),
                                 ctypes.byref(done)))
        if not done.value:
            return None
        return keys[:cnt.value], dist[:cnt.value], st.value, lat.value

    def replay(self, queries, ks, efs, arrival_us, precisions=None, filter_ids=None):
        """Native load generator: submits query i at arrival_us[i] and returns
        (latency_ms, keys, dists, counts, status) per request."""
        queries = _f32(queries).reshape(-1, self._dim)
        n = queries.shape[0]
        specs = np.zeros((n, 4), np.uint32)
        specs[:, 0] = ks
        specs[:, 1] = efs
        specs[:, 2] = 0 if precisions is None else np.asarray(precisions)
        specs[:, 3] = 0xFFFFFFFF if filter_ids is None else \
            np.where(np.asarray(filter_ids) < 0, 0xFFFFFFFF, np.asarray(filter_ids))
        arr = np.ascontiguousarray(arrival_us, np.float64)
        lat = np.zeros(n, np.float64)
        keys = np.zeros((n, self._k_max), np.uint64)
        dist = np.zeros((n, self._k_max), np.float32)
        cnt = np.zeros(n, np.uint32)
        st = np.zeros(n, np.int32)
        _check(lib.hh_serve_replay(self._h, _ptr(queries), _ptr(specs), n, _ptr(arr), _ptr(lat),
                                   _ptr(keys), _ptr(dist), _ptr(cnt), _ptr(st)))
        return lat, keys, dist, cnt, st

    def stats(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        st = lib.hh_serve_stats(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"served": a.value, "batches": b.value, "queued": c.value, "error": st}

    def close(self):
        if self._h:
            h, self._h = self._h, None
            self._index._servers -= 1
            _check(lib.hh_serve_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BruteForceIndex:
    """Exact search (src/brute.rs:28-205), on the GPU."""

    def __init__(self, dim, metric=Metric.L2, device=0):
        if dim == 0 or dim > MAX_DIM:
            raise InvalidDimension(f"invalid dimension {dim}: must be between 1 and {MAX_DIM}")
        self._dim = dim
        self._metric = Metric(metric)
        self._device = device
        self._keys = np.zeros(0, np.uint64)
        self._vecs = np.zeros((0, dim), np.float32)
        self._h = None
        self._mfma = 1

    def __del__(self):
        try:
            if getattr(self, "_h", None):
                lib.hh_brute_free(self._h)
        except Exception:
            pass

    def dim(self): return self._dim
    def metric(self): return self._metric
    def __len__(self): return int(self._keys.size)
    def len(self): return len(self)
    def is_empty(self): return len(self) == 0

    def insert(self, key, vector):
        self.insert_batch([key], np.asarray(vector, np.float32).reshape(1, -1))

    def insert_batch(self, keys, vectors):
        keys = np.ascontiguousarray(keys, np.uint64)
        vectors = _f32(vectors)
        if vectors.size != keys.size * self._dim:
            raise DimensionMismatch(
                f"dimension mismatch: expected {keys.size * self._dim}, got {vectors.size}",
                expected=keys.size * self._dim, actual=vectors.size)
        new_keys = np.concatenate([self._keys, keys])
        new_vecs = np.concatenate([self._vecs, vectors.reshape(-1, self._dim)])
        h = ctypes.c_void_p()
        bad = ctypes.c_uint64(0)
        # validation, duplicate checks and upload happen in the library; the
        # batch is rejected as a whole on error (src/brute.rs:102-114)
        st = lib.hh_brute_upload(self._dim, int(self._metric), _ptr(new_keys), _ptr(new_vecs),
                                 new_keys.size, self._device, ctypes.byref(h), ctypes.byref(bad))
        _check(st, key=bad.value)
        if self._h:
            lib.hh_brute_free(self._h)
        self._h = h
        self._keys, self._vecs = new_keys, new_vecs
        _check(lib.hh_brute_set_mfma(self._h, self._mfma))

    def set_mfma(self, mode):
        """Execution choice for f32 batches: 0 exact streaming kernel only,
        1 automatic, 2 MFMA shortlist whenever D <= 128 and k <= 64.  All
        modes return identical results."""
        if mode not in (0, 1, 2):
            raise ValueError("mode must be 0, 1 or 2")
        self._mfma = mode
        if self._h:
            _check(lib.hh_brute_set_mfma(self._h, mode))

    def mfma_stats(self):
        """(queries of the last batch on the MFMA path, of which re-run exactly)."""
        a, b = ctypes.c_uint64(0), ctypes.c_uint64(0)
        _check(lib.hh_brute_mfma_stats(self._handle(), ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def _handle(self):
        if self._h is None:
            h = ctypes.c_void_p()
            bad = ctypes.c_uint64(0)
            _check(lib.hh_brute_upload(self._dim, int(self._metric), None, None, 0,
                                       self._device, ctypes.byref(h), ctypes.byref(bad)))
            self._h = h
            _check(lib.hh_brute_set_mfma(self._h, self._mfma))
        return self._h

    def search_arrays(self, queries, k, filter=None, f64=False):
        queries = _f32(queries)
        if queries.size % self._dim:
            raise DimensionMismatch(
                f"dimension mismatch: expected {self._dim}, got {queries.size}",
                expected=self._dim, actual=queries.size)
        nq = queries.size // self._dim
        k = int(min(k, max(len(self), 0), 1 << 31))
        kk = max(k, 1)
        ok = np.zeros((nq, kk), np.uint64)
        od = np.zeros((nq, kk), np.float64 if f64 else np.float32)
        oc = np.zeros(nq, np.uint32)
        allow = None
        if filter is not None:
            mask = filter if isinstance(filter, np.ndarray) else np.fromiter(
                (bool(filter(int(x))) for x in self._keys), dtype=bool, count=self._keys.size)
            allow = _bits_from_mask(mask)
        bad = ctypes.c_uint64(0)
        if f64:
            st = lib.hh_brute_search_batch_f64(self._handle(), _ptr(queries), nq, k, _ptr(ok),
                                               _ptr(od), _ptr(oc), ctypes.byref(bad))
        else:
            st = lib.hh_brute_search_batch(self._handle(), _ptr(queries), nq, k, _ptr(allow),
                                           _ptr(ok), _ptr(od), _ptr(oc), ctypes.byref(bad))
        _check(st, query=bad.value)
        return ok[:, :k], od[:, :k], oc

    def search(self, query, k):
        return self.search_filtered(query, k, None)

    def search_filtered(self, query, k, filter):
        q = _f32(query).ravel()
        if q.size != self._dim:
            raise DimensionMismatch(f"dimension mismatch: expected {self._dim}, got {q.size}",
                                    expected=self._dim, actual=q.size)
        return _results(*self.search_arrays(q.reshape(1, -1), k, filter))[0]

    def search_batch(self, queries, k):
        return _results(*self.search_arrays(queries, k))


from . import device, eval, synthetic  # noqa: E402  (public submodules, as in the crate)
