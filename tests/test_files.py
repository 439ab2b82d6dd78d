"""The crate's index file (src/hnsw/io.rs, version 1) on the host: CRC-32
known answers, byte-exact round trips, the loader's error for every
structural violation the crate's own tests plant, memory bounds for crafted
files, and searches on a loaded graph equal to the original's (through the
oracle; the GPU form of that check is in test_gpu_search.py)."""
import struct

import numpy as np
import pytest

import hnsw_amd as H
import oracle as O
from helpers import build, oracle_graph


def bitwise_crc32(data):
    """Independent bitwise CRC-32 (IEEE), as tests/hostile_files.rs:78-91."""
    crc = 0xFFFFFFFF
    for b in data:
        crc ^= b
        for _ in range(8):
            crc = (crc >> 1) ^ 0xEDB88320 if crc & 1 else crc >> 1
    return crc ^ 0xFFFFFFFF


def crc(data):
    data = bytes(data)
    buf = np.frombuffer(data, np.uint8)
    return int(H.lib.hh_crc32(buf.ctypes.data if buf.size else None, buf.size))


def with_valid_crc(b):
    """src/hnsw/tests.rs:481-486: recompute the trailer so a mutation reaches
    structural validation."""
    b = bytearray(b)
    b[-4:] = struct.pack("<I", crc(b[:-4]))
    return bytes(b)


def format_error(b, **kw):
    with pytest.raises(H.FormatError) as e:
        H.HnswIndex.from_bytes(b, **kw)
    return e.value


def test_crc32_known_answers():
    """src/crc32.rs:103-111."""
    assert crc(b"123456789") == 0xCBF43926
    assert crc(b"") == 0
    assert crc(b"The quick brown fox jumps over the lazy dog") == 0x414FA339
    data = bytes(((i * 31) ^ (i >> 3)) & 0xFF for i in range(1000))
    for n in (0, 1, 7, 8, 9, 63, 64, 65, 1000):
        assert crc(data[:n]) == bitwise_crc32(data[:n]), n


def _round_trip_index():
    """src/hnsw/tests.rs:415-449."""
    b, q = O.dataset_clustered(500, 10, 8, 5, 2)
    params = H.HnswParams(6).with_selection(H.NeighborSelection.heuristic_with(True, False))
    idx = build(b, H.Metric.Cosine, params=params)
    for key in (3, 50, 499):
        idx.remove(key)
    idx.enable_quantization()
    return idx, q


def _lists(v):
    m, m0 = v["m"], v["m0"]
    l0 = v["layer0"].reshape(-1, m0 + 1)
    out = []
    for i in range(len(v["keys"])):
        out.append(tuple(l0[i, 1:1 + l0[i, 0]]))
        for layer in range(1, int(v["levels"][i]) + 1):
            blk = v["upper"][v["upper_off"][i] + (layer - 1) * (m + 1):]
            out.append(tuple(blk[1:1 + blk[0]]))
    return out


def test_round_trip_preserves_everything():
    idx, q = _round_trip_index()
    data = idx.to_bytes()
    loaded = H.HnswIndex.from_bytes(data)
    assert loaded.to_bytes() == data
    p, lp = idx.params(), loaded.params()
    assert (lp.m(), lp.m0(), lp.ef_construction(), lp.level_multiplier(), lp.seed()) == \
        (p.m(), p.m0(), p.ef_construction(), p.level_multiplier(), p.seed())
    assert (len(loaded), loaded.deleted_len()) == (497, 3)
    a, b = idx.graph_arrays(), loaded.graph_arrays()
    for name in ("keys", "levels", "deleted", "vectors", "upper_off", "codes", "sq_offset"):
        assert np.array_equal(a[name], b[name]), name
    assert _lists(a) == _lists(b)  # every neighbour list (slots past a count are unused)
    assert a["entry_id"] == b["entry_id"] and a["top"] == b["top"]
    assert a["sq_scale"] == b["sq_scale"]
    # searches (f32 and int8 + re-rank) through the oracle on both graphs
    ga, gb = oracle_graph(idx), oracle_graph(loaded)
    for quantized in (False, True):
        ra, rb = ga.search_batch(q, 5, 32, quantized), gb.search_batch(q, 5, 32, quantized)
        assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])
    empty = H.HnswIndex(3, H.Metric.L2, H.HnswParams())
    assert H.HnswIndex.from_bytes(empty.to_bytes()).is_empty()


def test_write_to_and_read_from_files(tmp_path):
    idx, _ = _round_trip_index()
    path = tmp_path / "index.hnsw"
    idx.write_to(str(path))
    assert path.read_bytes() == idx.to_bytes()
    assert H.HnswIndex.read_from(str(path)).to_bytes() == idx.to_bytes()


def _small_bytes(n=30):
    b, _ = O.dataset_gaussian(n, 1, 4, 1)
    return build(b, H.Metric.L2).to_bytes()


def test_loader_reports_envelope_errors():
    """src/hnsw/tests.rs:453-478."""
    data = _small_bytes(20)
    assert format_error(b"").kind == "Truncated"
    assert format_error(b"HNSW").kind == "Truncated"
    assert format_error(b"not an index file").kind == "BadMagic"
    v2 = bytearray(data)
    v2[8] = 2
    e = format_error(bytes(v2))
    assert e.kind == "UnsupportedVersion" and e.version == 2
    flipped = bytearray(data)
    flipped[100] ^= 0x10
    e = format_error(bytes(flipped))
    assert e.kind == "ChecksumMismatch" and e.stored != e.computed
    for n in (12, 40, len(data) // 2, len(data) - 1):
        with pytest.raises(H.Error):
            H.HnswIndex.from_bytes(data[:n])


def test_loader_rejects_structural_violations():
    """src/hnsw/tests.rs:488-573: every planted violation names its field."""
    data = _small_bytes(30)

    def invalid(offset, value):
        b = bytearray(data)
        b[offset:offset + len(value)] = value
        return format_error(with_valid_crc(b))

    u32 = lambda v: struct.pack("<I", v)
    expect = [
        (12, bytes([9]), "metric"),
        (13, bytes([0, 1]), "neighbor selection"),
        (15, bytes([2]), "quantized flag"),
        (16, u32(0), "dimension"),
        (20, u32(1), "parameters"),
        (32, struct.pack("<d", float("nan")), "parameters"),
        (48, u32(0xFFFFFFFF), "node count"),
        (52, u32(30), "entry point"),
        (56 + 8, struct.pack("<Q", 0), "duplicate live key"),
    ]
    for off, val, what in expect:
        e = invalid(off, val)
        assert e.kind == "Invalid" and e.what == what, (off, e.kind, e.what)
    assert invalid(48, u32(1_000_000)).kind == "Truncated"
    levels = 56 + 30 * 8
    assert invalid(levels, bytes([200])).what == "node level"
    assert invalid(levels + 30, bytes([7])).what == "tombstone flag"
    vectors = levels + 60
    assert invalid(vectors, struct.pack("<f", float("inf"))).what == "vector out of range"
    assert invalid(vectors, struct.pack("<f", 3e38)).what == "vector out of range"
    links = vectors + 30 * 4 * 4
    assert invalid(links, u32(1000)).what == "neighbor count exceeds capacity"
    assert invalid(links + 4, u32(999)).what == "neighbor id"
    count = struct.unpack("<I", data[links:links + 4])[0]
    assert count >= 2, "node 0 needs two neighbours for this case"
    assert invalid(links + 8, data[links + 4:links + 8]).what == "duplicate neighbor id"
    longer = data[:-4] + bytes(5) + data[-4:]
    assert format_error(with_valid_crc(longer)).kind == "TrailingBytes"


def test_loader_rejects_implausible_levels_before_allocating():
    """src/hnsw/tests.rs:576-592."""
    b, _ = O.dataset_gaussian(2000, 1, 1, 1)
    data = bytearray(build(b, H.Metric.L2, m=64).to_bytes())
    levels = 56 + 2000 * 8
    data[levels:levels + 2000] = bytes([31]) * 2000
    assert format_error(with_valid_crc(data)).what == "implausible level distribution"


def test_loader_rejects_vectors_outside_the_cosine_unit_sphere():
    """src/hnsw/tests.rs:595-607."""
    idx = H.HnswIndex(2, H.Metric.Cosine, H.HnswParams())
    idx.insert_batch(np.array([1, 2], np.uint64), np.array([[1.0, 0.0], [0.0, 1.0]], np.float32))
    data = bytearray(idx.to_bytes())
    vectors = 56 + 2 * 8 + 2 + 2
    data[vectors:vectors + 4] = struct.pack("<f", 2.0)
    assert format_error(with_valid_crc(data)).what == "cosine vector is not unit length"


@pytest.mark.parametrize("metric,quantize", [(H.Metric.L2, False), (H.Metric.Cosine, True)])
def test_memory_limit_is_the_loaded_memory_usage(metric, quantize):
    """src/hnsw/tests.rs:610-633, with this library's heap layout."""
    b, _ = O.dataset_gaussian(3000, 1, 24, 4)
    idx = build(b, metric, m=12, threads=4)
    if quantize:
        idx.enable_quantization()
    data = idx.to_bytes()
    with pytest.raises(H.MemoryLimitExceeded) as e:
        H.HnswIndex.from_bytes(data, max_memory=0)
    required = e.value.required
    assert e.value.limit == 0 and required > 0
    with pytest.raises(H.MemoryLimitExceeded):
        H.HnswIndex.from_bytes(data, max_memory=required - 1)
    loaded = H.HnswIndex.from_bytes(data, max_memory=required)
    assert loaded.memory_usage() == required == idx.memory_usage()


def test_default_memory_limit_admits_low_dimensional_indexes_with_m0_256():
    """src/hnsw/tests.rs:636-656."""
    b, _ = O.dataset_gaussian(20000, 1, 1, 5)
    idx = build(b, H.Metric.L2, m=128, threads=8)
    data = idx.to_bytes()
    with pytest.raises(H.MemoryLimitExceeded) as e:
        H.HnswIndex.from_bytes(data, max_memory=0)
    ratio = e.value.required / len(data)
    assert ratio < 48.0, ratio
    assert len(H.HnswIndex.from_bytes(data)) == 20000


def crafted(n, level, layers_present):
    """tests/hostile_files.rs:93-125: n one-dimensional nodes on `level`
    under the largest M, M0 and mL, every list empty; with layers_present
    below level + 1 the adjacency section is cut short."""
    b = bytearray(b"HNSWIDX\0")
    b += struct.pack("<I", 1) + bytes([0, 1, 0, 0])
    b += struct.pack("<IIII", 1, 512, 2048, 200)
    b += struct.pack("<dQII", 10.0, 0, n, 0)
    b += np.arange(n, dtype="<u8").tobytes()
    b += bytes([level]) * n + bytes(n)
    b += np.arange(n, dtype="<f4").tobytes()
    b += bytes(4 * layers_present * n)
    b += struct.pack("<I", bitwise_crc32(b) if n < 2000 else crc(b))
    return bytes(b)


def _rss_kib():
    with open("/proc/self/status") as f:
        for line in f:
            if line.startswith("VmRSS:"):
                return int(line.split()[1])
    return 0


def test_crafted_files_are_rejected_without_large_allocations():
    """tests/hostile_files.rs:128-159: a 1.9 MB file describing about 945 MB
    of graph storage, and the same file cut after the layer-0 counts, fail
    before the loader allocates graph storage."""
    complete = crafted(20000, 19, 20)
    assert len(complete) == 1_880_060
    before = _rss_kib()
    with pytest.raises(H.MemoryLimitExceeded) as e:
        H.HnswIndex.from_bytes(complete)
    assert e.value.required > 900_000_000 and e.value.limit == 64 * len(complete)
    truncated = crafted(20000, 19, 1)
    assert len(truncated) == 360_060
    assert format_error(truncated, max_memory=(1 << 64) - 2).kind == "Truncated"
    assert _rss_kib() - before < 64 * 1024  # far below the 945 MB described


def test_truncated_and_mutated_files_never_crash():
    """tests/properties.rs:188-230: every prefix fails; mutated files (with a
    recomputed checksum) fail or load into a searchable graph."""
    data = _small_bytes(40)
    for n in range(0, len(data), 37):
        with pytest.raises(H.Error):
            H.HnswIndex.from_bytes(data[:n])
    rng = np.random.default_rng(5)
    links = 56 + 40 * (8 + 1 + 1 + 16)
    loaded_ok = 0
    for trial in range(300):
        b = bytearray(data)
        start = links if trial % 2 else 0
        for _ in range(int(rng.integers(1, 8))):
            b[start + int(rng.integers(0, len(b) - 4 - start))] = int(rng.integers(0, 256))
        try:
            g = H.HnswIndex.from_bytes(with_valid_crc(b))
        except (H.FormatError, H.MemoryLimitExceeded):
            continue
        loaded_ok += 1
        og = oracle_graph(g)
        try:
            ok, od, oc, _ = og.search_batch(np.full((1, g.dim()), 0.5, np.float32), 5, 16)
            assert oc[0] <= 5
        except ValueError:
            pass  # e.g. a zero vector under cosine
        assert H.HnswIndex.from_bytes(g.to_bytes()).to_bytes() == g.to_bytes()
    assert loaded_ok > 0
