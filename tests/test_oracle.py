"""The oracle pinned against the reference crate's own known answers
(tests/golden/reference_kats.json), closed forms and properties.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

KAT = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_kats.json")))


def test_splitmix64_known_answers():
    k = KAT["splitmix64"]
    r = O.SplitMix64(k["seed"])
    assert [r.next_u64(), r.next_u64()] == k["outputs"]


def test_splitmix_is_vigna_closed_form():
    # state += golden; output = mix(state) with Vigna's constants
    s = 987654321
    r = O.SplitMix64(s)
    for _ in range(5):
        s = (s + 0x9E3779B97F4A7C15) % 2**64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) % 2**64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) % 2**64
        assert r.next_u64() == z ^ (z >> 31)


def test_gaussian_first_components():
    k = KAT["gaussian_first_components"]
    base, _ = O.dataset_gaussian(*k["args"])
    assert np.allclose(base.ravel()[:3], k["values"], atol=k["tol"], rtol=0)


def test_gaussian_is_box_muller_of_the_stream():
    # closed form: u1 = 1 - next_f64, u2 = next_f64, sqrt(-2 ln u1) cos(2 pi u2)
    r = O.SplitMix64(77)
    s = O.SplitMix64(77)
    for _ in range(50):
        u1 = 1.0 - (s.next_u64() >> 11) * 2.0**-53
        u2 = (s.next_u64() >> 11) * 2.0**-53
        want = math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2)
        assert abs(r.next_gaussian() - want) < 1e-12


def test_gaussian_moments():
    k = KAT["gaussian_moments"]
    r = O.SplitMix64(k["seed"])
    x = np.array([r.next_gaussian() for _ in range(k["n"])])
    assert abs(x.mean()) < k["mean_tol"] and abs(x.var() - 1.0) < k["var_tol"]


def test_next_below_stays_in_range_and_covers():
    r = O.SplitMix64(3)
    seen = {r.next_below(7) for _ in range(1000)}
    assert seen == set(range(7))


def test_clustered_is_deterministic_and_clustered():
    a = O.dataset_clustered(100, 5, 8, 4, 1)
    b = O.dataset_clustered(100, 5, 8, 4, 1)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    c = O.dataset_clustered(100, 5, 8, 4, 2)
    assert not np.array_equal(a[0], c[0])
    # 4 clusters with spread 0.35: every row is within ~5 spread of some other
    # row of its cluster, far closer than unrelated N(0,1) centres
    base = O.dataset_clustered(2000, 0, 16, 4, 9)[0]
    d = ((base[:, None, :] - base[None, :200, :]) ** 2).sum(-1)
    nearest = np.sort(d, axis=1)[:, 1]
    assert np.median(nearest) < 16 * 0.35**2 * 2 * 1.5


def test_scalar_and_doctest_kernels():
    k = KAT["scalar_kernels"]
    assert O.dot(k["a"], k["b"], O.SCALAR) == k["dot"]
    assert O.l2(k["a"], k["b"], O.SCALAR) == k["l2"]
    assert O.dot([], [], O.SCALAR) == 0.0
    d = KAT["kernel_doctest"]
    assert O.dot(d["a"], d["b"], O.SCALAR) == d["dot"]
    assert O.l2(d["a"], d["b"], O.SCALAR) == d["l2"]
    for be in (O.SCALAR, O.PORTABLE, O.AVX2):
        assert O.dot([], [], be) == 0.0 and O.l2([], [], be) == 0.0


def test_metric_definitions():
    k = KAT["metric_definitions"]
    assert O.distance(O.L2, k["a"], k["b"], O.SCALAR) == k["l2"]
    ip = O.distance(O.IP, k["a"], k["b"], O.SCALAR)
    assert ip == 0.0 and math.copysign(1.0, ip) == -1.0  # -0.0
    assert O.distance(O.COSINE, k["a"], k["b"], O.SCALAR) == k["cos"]
    assert O.distance(O.COSINE, k["a"], k["a"], O.SCALAR) == k["cos_self"]


def test_exact_on_small_integers_every_backend():
    """tests/kernels.rs:83-96"""
    a = np.array([(i % 7) - 3 for i in range(133)], np.float32)
    b = np.array([(i % 5) - 2 for i in range(133)], np.float32)
    for be in (O.SCALAR, O.PORTABLE, O.AVX2):
        assert O.dot(a, b, be) == float((a * b).sum())
        assert O.l2(a, b, be) == float(((a - b) ** 2).sum())


def _tol(terms, n):
    return (n + 4) * np.finfo(np.float32).eps * float(np.abs(terms).sum()) + np.finfo(np.float32).tiny


@pytest.mark.parametrize("seed", range(40))
def test_backends_match_scalar_within_the_reference_tolerance(seed):
    """tests/kernels.rs:11-71 (the proptest bound)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 301))
    a = rng.uniform(-100, 100, n).astype(np.float32)
    b = rng.uniform(-100, 100, n).astype(np.float32)
    wd, wl = O.dot(a, b, O.SCALAR), O.l2(a, b, O.SCALAR)
    for be in (O.PORTABLE, O.AVX2):
        assert abs(O.dot(a, b, be) - wd) <= _tol(a * b, n)
        assert abs(O.l2(a, b, be) - wl) <= _tol((a - b) ** 2, n)
        assert O.l2(a, b, be) >= 0


def test_every_length_up_to_two_blocks():
    """tests/kernels.rs:73-81"""
    for n in range(81):
        a = np.sin(np.arange(n) * 0.37).astype(np.float32)
        b = np.cos(np.arange(n) * 0.11).astype(np.float32)
        wd, wl = O.dot(a, b, O.SCALAR), O.l2(a, b, O.SCALAR)
        for be in (O.PORTABLE, O.AVX2):
            assert abs(O.dot(a, b, be) - wd) <= _tol(a * b, n)
            assert abs(O.l2(a, b, be) - wl) <= _tol((a - b) ** 2, n)


def test_int8_kernels_match_naive_sums():
    """src/kernels/int8.rs:68-85"""
    for n in range(80):
        a = np.array([((i * 37) % 255) for i in range(n)], np.int64).astype(np.uint8).view(np.int8)
        b = np.array([((i * 91 + 13) % 255) for i in range(n)], np.int64).astype(np.uint8).view(np.int8)
        assert O.i8_dot(a, b) == int((a.astype(np.int64) * b.astype(np.int64)).sum())
        assert O.i8_l2(a, b) == int(((a.astype(np.int64) - b.astype(np.int64)) ** 2).sum())
    a = np.full(65536, 127, np.int8)
    b = np.full(65536, -127, np.int8)
    assert O.i8_dot(a, b) == -(127 * 127) * 65536
    assert O.i8_l2(a, b) == 254 * 254 * 65536


def test_prepare_known_answers_and_rejections():
    """src/vector.rs:106-179"""
    k = KAT["cosine_normalisation"]
    st, v = O.prepare(O.COSINE, 2, k["input"])
    assert st == 0 and np.allclose(v, k["expect"], atol=k["tol"], rtol=0)
    st, v = O.prepare(O.COSINE, 2, k["input2"])
    assert st == 0 and v.tolist() == k["expect2"]
    cases = [([1.0], O.L2, 1), ([np.nan, 1.0], O.L2, 4), ([np.inf, 1.0], O.IP, 4),
             ([3.4028235e38, 3.4028235e38], O.L2, 4), ([0.0, 0.0], O.COSINE, 5)]
    for v, metric, want in cases:
        assert O.prepare(metric, 2, v)[0] == want
    assert O.prepare(O.L2, 2, [0.0, 0.0])[0] == 0
    assert O.prepare(O.L2, 1, [1.3e19])[0] == 4


def test_quantizer_known_answers():
    """src/quantization.rs:125-145"""
    k = KAT["quantizer"]
    off, scale = O.sq_train(O.L2, np.array(k["vectors"], np.float32))
    assert off.tolist() == k["offset"] and abs(scale - k["scale"]) < k["scale_tol"]
    assert O.sq_encode(off, scale, k["encode"]).tolist() == k["codes"]
    off, _ = O.sq_train(O.IP, np.array(k["vectors"], np.float32))
    assert off.tolist() == [0.0, 0.0]
    _, s = O.sq_train(O.L2, np.array(k["degenerate"], np.float32))
    assert s == k["degenerate_scale"]


def test_quantizer_tracks_exact_distances():
    """src/quantization.rs:101-123"""
    base, q = O.dataset_gaussian(500, 50, 32, 11)
    for metric in (O.L2, O.IP):
        off, sc = O.sq_train(metric, base)
        codes = O.sq_encode_rows(off, sc, base)
        qc = O.sq_encode(off, sc, q[0])
        worst = 0.0
        for i in range(500):
            exact = O.distance(metric, q[0], base[i])
            approx = O.sq_distance(metric, sc, qc, codes[i])
            scale = exact if metric == O.L2 else 32.0
            worst = max(worst, abs(exact - approx) / scale)
        assert worst < 0.05


def test_quantizer_rounds_half_away_from_zero():
    off = np.zeros(3, np.float32)
    assert O.sq_encode(off, 1.0, [0.5, -0.5, 2.5]).tolist() == [1, -1, 3]


def test_brute_force_known_answers():
    """src/brute.rs:211-262"""
    k = KAT["brute_force"]
    rows = np.array(k["rows"], np.float32).reshape(-1, 1)
    keys, dists = O.brute_search(O.L2, k["keys"], rows, [k["query"]], 3)
    assert keys.tolist() == k["top3"] and dists[0] == k["top_distance"]
    allow = np.array([kk > 4 for kk in k["keys"]], np.uint8)
    keys, _ = O.brute_search(O.L2, k["keys"], rows, [k["query"]], 2, allow=allow)
    assert keys.tolist() == k["filtered_gt4_k2"]
    assert len(O.brute_search(O.L2, k["keys"], rows, [k["query"]], 1 << 40)[0]) == 4
    assert len(O.brute_search(O.L2, k["keys"], rows, [k["query"]], 0)[0]) == 0
    tr = np.array(k["tie_rows"], np.float32).reshape(-1, 1)
    assert O.brute_search(O.L2, k["tie_keys"], tr, [0.0], 3)[0].tolist() == k["tie_order"]
    d = KAT["brute_doctest"]
    keys, dists = O.brute_search(O.L2, d["keys"], np.array(d["rows"], np.float32), d["query"], 1)
    assert keys[0] == d["key"] and dists[0] == d["distance"]


def test_brute_f64_matches_numpy_f64():
    base, q = O.dataset_gaussian(1000, 5, 24, 3)
    keys = np.arange(1000)
    for i in range(5):
        k, d = O.brute_search_f64(O.L2, keys, base, q[i], 10)
        ref = ((base.astype(np.float64) - q[i].astype(np.float64)) ** 2).sum(1)
        order = np.lexsort((keys, ref))[:10]
        assert k.tolist() == order.tolist()
        assert np.allclose(d, ref[order], rtol=1e-12)


def test_candidate_order_known_answer():
    """src/hnsw/search.rs:169-178"""
    import functools
    k = KAT["candidate_order"]
    items = [(float("nan") if d == "nan" else d, i) for d, i in k["pushed"]]
    order = sorted(items, key=functools.cmp_to_key(
        lambda a, b: O.candidate_cmp(a[0], a[1], b[0], b[1])))
    assert [i for _, i in order] == k["order"]


def test_recall_at_k_known_answers():
    """src/eval.rs:59-72"""
    k = KAT["recall_at_k"]
    t = np.array(k["truth"], np.uint64)
    f = np.array(k["found"], np.uint64)
    c = np.array([3, 3], np.uint32)
    assert abs(O.recall_at_k(t, c, f, c, 3) - k["k3"]) < 1e-12
    assert O.recall_at_k(t, c, f, c, 1) == k["k1"]
    one = np.array([[1]], np.uint64)
    assert O.recall_at_k(one, [1], one, [1], 10) == 1.0
    assert O.recall_at_k(np.zeros((0, 1), np.uint64), [], np.zeros((0, 1), np.uint64), [], 10) == 1.0
    assert O.recall_at_k(np.array([[1, 2]], np.uint64), [2], np.zeros((1, 2), np.uint64), [0], 2) == 0.0


def _splitmix_py(seed):
    """Vigna's SplitMix64 in Python integers (pinned above by the crate's KAT
    and the closed form), used as an independent stream below."""
    s = seed
    while True:
        s = (s + 0x9E3779B97F4A7C15) % 2**64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) % 2**64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) % 2**64
        yield z ^ (z >> 31)


def test_next_below_is_lemires_multiply_shift():
    """src/synthetic.rs:43-45: ((u as u128 * n) >> 64) over the raw stream.
    A modulo reduction, a 32-bit product or a shift of the wrong width each
    fail this (checked on the last line: the modulo form differs here)."""
    bounds = [1, 2, 7, 100, 1000, 20, 2**32 - 1, 2**32 + 5, 2**40 + 7, 2**63 + 11, 2**64 - 1]
    r = O.SplitMix64(271828)
    raw = _splitmix_py(271828)
    want, modulo = [], []
    for i in range(400):
        n = bounds[i % len(bounds)]
        u = next(raw)
        want.append((u * n) >> 64)
        modulo.append(u % n)
        assert r.next_below(n) == want[-1], (i, n)
    assert want != modulo


def test_clustered_follows_the_crate_draw_order():
    """src/synthetic.rs:102-124 restated in Python floats from the raw
    stream: all centres first (f64 Box-Muller), then per point one Lemire
    index followed by dim spread draws, base before queries, rounded to f32
    once.  Python's math.log/cos are the platform libm the crate calls."""
    base, queries = O.dataset_clustered(6, 4, 5, 3, 99)
    raw = _splitmix_py(99)

    def f64():
        return (next(raw) >> 11) * 2.0**-53

    def gauss():
        u1 = 1.0 - f64()
        u2 = f64()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(math.tau * u2)

    centers = [gauss() for _ in range(3 * 5)]
    for out in (base, queries):
        for row in out:
            c = (next(raw) * 3) >> 64
            want = np.array([centers[c * 5 + d] + 0.35 * gauss() for d in range(5)], np.float32)
            assert np.array_equal(row.view(np.uint32), want.view(np.uint32))


def _sq_distance_crate(metric, scale, a, b):
    """src/quantization.rs:85-92 in numpy: exact integer sum, converted to
    f32 once (round to nearest even), times s2 = scale*scale in f32; cosine
    subtracts that product from 1 in f32 (no fused multiply-add)."""
    a64, b64 = a.astype(np.int64), b.astype(np.int64)
    s2 = np.float32(scale) * np.float32(scale)
    if metric == O.L2:
        return np.float32(int(((a64 - b64) ** 2).sum())) * s2
    dot = np.float32(int((a64 * b64).sum()))
    return -(dot * s2) if metric == O.IP else np.float32(1.0) - dot * s2


def test_sq_distance_rounds_like_the_crate():
    """Sums above 2^24 make the i64 -> f32 conversion round; a product taken
    in f64, a fused cosine step or s2 formed in f64 each change some of
    these values (the last assertions show the f64 form does)."""
    rng = np.random.default_rng(5)
    f64_differs = 0
    for trial in range(300):
        dim = int(rng.choice([1, 16, 128, 960, 1024, 4096]))
        a = rng.integers(-127, 128, dim).astype(np.int8)
        b = rng.integers(-127, 128, dim).astype(np.int8)
        scale = np.float32(rng.uniform(1e-3, 0.1))
        for metric in (O.L2, O.IP, O.COSINE):
            got = O.sq_distance(metric, scale, a, b)
            want = _sq_distance_crate(metric, scale, a, b)
            assert np.array_equal(np.float32(got).view(np.uint32), np.float32(want).view(np.uint32)), \
                (trial, dim, metric)
        l2 = int(((a.astype(np.int64) - b.astype(np.int64)) ** 2).sum())
        once = np.float32(float(l2) * float(scale) * float(scale))
        f64_differs += int(once != _sq_distance_crate(O.L2, scale, a, b))
    assert f64_differs > 0


def test_bf16_round_is_torch_round_to_nearest_even():
    """or_bf16_round against torch's CPU float32 -> bfloat16 conversion
    (round to nearest, ties to even), on random values, exact ties with even
    and odd kept halves, subnormals, the largest finite values and infinities."""
    import torch
    rng = np.random.default_rng(8)
    x = rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-30, 30, 20000).astype(np.float32)
    u = x.view(np.uint32)
    u[:4000] = (u[:4000] & 0xFFFF0000) | 0x8000  # exact ties
    extra = np.array([0.0, -0.0, 1e-40, -1e-42, 3.4028235e38, -3.4028235e38, np.inf, -np.inf,
                      1.0 + 2.0**-8, 1.0 + 3 * 2.0**-8], np.float32)
    x = np.concatenate([x, extra])
    got = O.bf16_round(x)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_perdim_quantizer_closed_forms():
    """The per-dimension quantizer (no crate counterpart): mid and scale
    are the column midpoint and range / 254; every value is reconstructed
    to within half a step (no value of the training rows is clamped); a
    constant column codes to 0 with scale 1; the L2 row term is the f32
    rounding of sum (scale_d * code_d)^2."""
    rng = np.random.default_rng(12)
    rows = (rng.standard_normal((400, 23)) * rng.uniform(0.01, 50, 23)).astype(np.float32)
    rows[:, 5] = 2.5  # constant column
    mid, scale = O.pd_train(rows)
    lo, hi = rows.min(0), rows.max(0)
    assert np.array_equal(mid, ((lo + hi) * np.float32(0.5)))
    want_scale = (hi - lo) / np.float32(254)
    want_scale[5] = 1.0
    assert np.array_equal(scale, want_scale)
    codes = O.pd_encode_rows(mid, scale, rows)
    assert np.abs(codes.astype(np.int64)).max() <= 127 and (codes[:, 5] == 0).all()
    recon = mid.astype(np.float64) + scale.astype(np.float64) * codes
    err = np.abs(rows - recon)
    assert (err <= scale.astype(np.float64) * 0.5 * (1 + 1e-5) + 1e-30).all()
    rn = O.pd_rnorm_rows(scale, codes)
    import math
    for i in range(0, 400, 7):
        exact = math.fsum((float(scale[d]) * int(codes[i, d])) ** 2 for d in range(23))
        assert abs(float(rn[i]) - exact) <= 2.0**-23 * exact


@pytest.mark.parametrize("metric", [O.L2, O.IP, O.COSINE])
def test_perdim_distance_error_bound(metric):
    """The per-dimension traversal distance against the exact (f64)
    distance to the stored row: the query weights are within sigma/2 of
    sigma*qcode, the rows within scale/2 of their reconstruction, so
    |approx - exact| <= sigma/2 * sum|code| + sum|qw_d| * scale_d/2 (+ the
    x - x-hat term for L2) + f32 rounding.  A dropped term (the row norm,
    qconst) or a wrong sign breaks the bound."""
    rng = np.random.default_rng(metric)
    dim = 96
    rows = rng.standard_normal((300, dim)).astype(np.float32)
    qs = rng.standard_normal((20, dim)).astype(np.float32)
    if metric == O.COSINE:
        rows /= np.linalg.norm(rows, axis=1, keepdims=True)
        qs /= np.linalg.norm(qs, axis=1, keepdims=True)
    mid, scale = O.pd_train(rows)
    codes = O.pd_encode_rows(mid, scale, rows)
    rn = O.pd_rnorm_rows(scale, codes) if metric == O.L2 else np.zeros(300, np.float32)
    s64 = scale.astype(np.float64)
    for q in qs:
        qcode, sigma, qconst = O.pd_query(metric, mid, scale, q)
        q64 = q.astype(np.float64)
        w = (q64 - mid) * s64 if metric == O.L2 else q64 * s64
        assert np.abs(w - float(sigma) * qcode).max() <= float(sigma) * 0.5 * (1 + 1e-5)
        for i in range(0, 300, 3):
            x = rows[i].astype(np.float64)
            xh = mid.astype(np.float64) + s64 * codes[i]
            approx = float(O.pd_distance(metric, sigma, qconst, rn[i], qcode, codes[i]))
            if metric == O.L2:
                exact_h = float(((q64 - xh) ** 2).sum())
                exact = float(((q64 - x) ** 2).sum())
                # quantising the weights: 2 * |sum (w - sigma b) c|
                bound_q = 2 * float(sigma) * 0.5 * np.abs(codes[i]).sum()
                bound_x = abs(exact - exact_h)
            else:
                dot_h = float(q64 @ xh)
                exact_h = -dot_h if metric == O.IP else 1.0 - dot_h
                exact = -float(q64 @ x) if metric == O.IP else 1.0 - float(q64 @ x)
                bound_q = float(sigma) * 0.5 * np.abs(codes[i]).sum()
                bound_x = float((np.abs(q64) * s64 * 0.5).sum())
            tol = 1e-5 * (abs(exact_h) + float(np.abs(w).sum()) * 127 + 1.0)
            assert abs(approx - exact_h) <= bound_q + tol, (i, approx, exact_h)
            assert abs(approx - exact) <= bound_q + bound_x + tol, (i, approx, exact)


def _line_graph(xs, lists, m0):
    """A one-level graph over 1-D points with the given layer-0 lists."""
    n = len(xs)
    layer0 = np.full((n, m0 + 1), 0xFFFFFFFF, np.uint32)
    for i, nb in enumerate(lists):
        layer0[i, 0] = len(nb)
        layer0[i, 1:1 + len(nb)] = nb
    vec = np.asarray(xs, np.float32).reshape(n, 1)
    return O.OracleGraph(1, O.L2, m0 // 2, m0, vec, np.arange(n, dtype=np.uint64),
                         np.zeros(n, np.uint8), layer0.reshape(-1), np.zeros(n, np.uint64),
                         np.zeros(1, np.uint32), 0, 0, bf16=True)


def test_beam_steps_hand_worked_graph():
    """Six points on a line, query 0, entry x=10, ef=2 (distances are exact
    squares).  The crate's rule expands 0, 1, 3, 4 (node 2 is never
    popped: after node 1's step the worst result is 9 < 16).  A beam of
    two pops 1 and 2 together (the worst is then 16, so node 2 passes the
    stop rule), visits 5 from node 2, then pops 3 alone (5 at 12.25 is
    beyond the worst, 9) and 4 alone: five expansions, six distances, the
    same two results."""
    xs = [10.0, 3.0, 4.0, 1.0, 2.0, 3.5]
    lists = [[1, 2], [0, 3], [0, 5], [1, 4], [3], [2]]
    g = _line_graph(xs, lists, 4)
    q = np.zeros(1, np.float32)
    want_keys = [3, 4]
    for width, n_expand, n_dist in ((1, 4, 5), (2, 5, 6), (4, 5, 6)):
        g.set_beam_width(width)
        keys, dists, _, cnt = g.search(q, 2, 2, precision=O.PREC_BF16)
        assert list(keys) == want_keys and list(dists) == [1.0, 4.0], width
        assert (cnt.n_expand, cnt.n_dist) == (n_expand, n_dist), (width, cnt.n_expand, cnt.n_dist)
    # f32 traversal ignores the beam width (the crate's rule)
    g.set_beam_width(2)
    _, _, _, cnt = g.search(q, 2, 2)
    assert (cnt.n_expand, cnt.n_dist) == (4, 5)


@pytest.mark.parametrize("width", [2, 3, 4])
def test_beam_steps_on_a_complete_graph_equal_brute_force(width):
    """n <= m0 with every node linked to every other: any search reaches
    every node, so the results are the exact top-k (bf16 traversal, exact
    re-rank)."""
    rng = np.random.default_rng(width)
    n, dim = 24, 16
    vec = rng.standard_normal((n, dim)).astype(np.float32)
    m0 = 32
    layer0 = np.full((n, m0 + 1), 0xFFFFFFFF, np.uint32)
    for i in range(n):
        nb = [j for j in range(n) if j != i]
        layer0[i, 0] = len(nb)
        layer0[i, 1:1 + len(nb)] = nb
    g = O.OracleGraph(dim, O.L2, 16, m0, vec, np.arange(n, dtype=np.uint64), np.zeros(n, np.uint8),
                      layer0.reshape(-1), np.zeros(n, np.uint64), np.zeros(1, np.uint32), 5, 0,
                      bf16=True, beam_width=width)
    for qi in range(10):
        q = rng.standard_normal(dim).astype(np.float32)
        keys, dists, _, _ = g.search(q, 5, 5, precision=O.PREC_BF16)
        bk, bd = O.brute_search(O.L2, np.arange(n, dtype=np.uint64), vec, q, 5)
        assert np.array_equal(keys, bk) and np.array_equal(dists, bd)


def test_beam_steps_keep_recall_and_add_few_distances():
    """On a 10k gaussian graph built by the crate's sequential insert, beam
    steps of 2 and 4 (bf16 and per-dimension codes) lose no recall against
    the crate's rule (a beam explores more, never less: at most 0.01 below
    it, where a different path may end elsewhere) at ef 32 and 128, with at
    most 15% more distance evaluations."""
    import hnsw_amd as H
    base, queries = O.dataset_gaussian(10000, 200, 64, 42)
    idx = H.HnswIndex(64, H.Metric.L2, H.HnswParams(16))
    idx.insert_batch(np.arange(10000), base, threads=1)
    v = idx.graph_arrays()
    tk, _, tc = O.brute_search_batch(O.L2, v["keys"], v["vectors"], queries, 10)
    mid, scale = O.pd_train(v["vectors"])
    g = O.OracleGraph(v["dim"], v["metric"], v["m"], v["m0"], v["vectors"], v["keys"],
                      v["levels"], v["layer0"], v["upper_off"], v["upper"], v["entry_id"],
                      v["top"], bf16=True, perdim=(mid, scale))
    for prec in (O.PREC_BF16, O.PREC_INT8_PERDIM):
        for ef in (32, 128):
            g.set_beam_width(1)
            ok, _, oc, c1 = g.search_batch(queries, 10, ef, precision=prec)
            r1 = O.recall_at_k(tk, tc, ok, oc, 10)
            for width in (2, 4):
                g.set_beam_width(width)
                ok, _, oc, cw = g.search_batch(queries, 10, ef, precision=prec)
                rw = O.recall_at_k(tk, tc, ok, oc, 10)
                assert rw >= r1 - 0.01, (prec, ef, width, r1, rw)
                assert cw.n_dist <= 1.15 * c1.n_dist, (prec, ef, width)
                assert cw.n_expand >= c1.n_expand  # every step expands at least as much
