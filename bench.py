"""Benchmark of the HNSW query path on one MI355X.

Workload (the GPU configuration of the headline metric): 1,000,000 f32
vectors, dim 128, clustered (1,000 centres uniform in [-1,1]^128, rows =
centre + N(0, 0.1^2)), L2, M=16, efConstruction=200, graph built on the CPU
by the host library and uploaded; 10,000 queries = random centre +
N(0, 0.12^2); k=10; ef swept over {16,32,64,128,256}.  Ground truth is the
exact GPU brute force (bit-identical to the crate's BruteForceIndex),
spot-checked against the FP64 oracle kernel.

A step is one batched search of all queries, inputs resident in HBM, at the
smallest swept ef whose recall@10 reaches 0.95.  The JSON line reports
QPS over the timed steps, the int8 and bf16 traversal modes at that ef, the
search kernel's roofline position (algorithmic gathered bytes over the
kernel time measured with HIP events on the launch stream), and the CPU
baseline: the oracle's C restatement of search_batch on the host cores, on
a bounded sample of the queries.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hnsw-index_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def warm_server(srv, queries, k, ef, precision=0):
    """Untimed requests through a new server before its measured replay:
    the first launch of a kernel instance in a process pays a one-time
    load (milliseconds), which would otherwise land in the measured
    latencies of the requests queued behind it."""
    n = min(256, len(queries))
    srv.replay(queries[:n], np.full(n, k), np.full(n, ef), np.zeros(n),
               precisions=np.full(n, int(precision)))
    return srv.stats()["batches"]


def measure_quantizer(H, idx, reps):
    """The int8 quantizer trained and applied on the device over the
    resident rows (reps timed runs of hh_index_quantize, wall time including
    the parameter read-back), installed in the host graph, then the host
    library's quantizer on the same vectors on one core; both must agree bit
    for bit."""
    h = idx.device_index()
    dim, n = idx.dim(), idx.node_count()
    off = np.zeros(dim, np.float32)
    sc = ctypes.c_float()
    dev_s = []
    for _ in range(reps):
        t0 = time.perf_counter()
        H._check(H.lib.hh_index_quantize(h, off.ctypes.data, ctypes.byref(sc)))
        dev_s.append(time.perf_counter() - t0)
    idx.enable_quantization(on_device=True)
    t0 = time.perf_counter()
    H._check(H.lib.hh_graph_enable_quantization(idx._g))
    host_s = time.perf_counter() - t0
    ga = idx.graph_arrays()
    same = bool(np.array_equal(ga["sq_offset"].view(np.uint32), off.view(np.uint32))
                and np.float32(ga["sq_scale"]) == np.float32(sc.value))
    if not same:
        raise SystemExit("device quantizer differs from the host quantizer")
    log(f"quantizer: device {min(dev_s) * 1e3:.2f} ms, host {host_s * 1e3:.0f} ms "
        f"({n} x {dim}), parameters identical")
    return {"device_s": round(min(dev_s), 5), "host_s": round(host_s, 4), "rows": n, "dim": dim,
            "identical_parameters": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--dim", type=int, default=128)
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--efs", default="16,32,64,66,68,72,96,128,256")
    ap.add_argument("--target", type=float, default=0.95)
    ap.add_argument("--cpu-sample", type=int, default=10000)
    ap.add_argument("--threads", type=int, default=0,
                    help="host threads for the CPU baseline and the device-assisted build "
                         "(default: up to 16)")
    ap.add_argument("--build-threads", type=int, default=1,
                    help="host threads for the benchmark graph's build; 1 (the default) builds "
                         "the same graph every run, so recall at each ef is reproducible (a "
                         "parallel build, like the crate's insert_batch, is not deterministic)")
    ap.add_argument("--graph-cache", default="",
                    help="index file of the benchmark graph: loaded (validated) when it exists, "
                         "else written after the build; with the default one-thread build the "
                         "file holds the same graph every run builds, so profiler runs need not "
                         "build it again")
    ap.add_argument("--f64-check", type=int, default=500)
    ap.add_argument("--quick", action="store_true",
                    help="search measurement only: no precision variants, serving or CPU "
                         "baseline (for profiler runs)")
    ap.add_argument("--quick-precision", choices=["bf16", "int8", "int8_perdim"], default="",
                    help="with --quick: after the f32 measurement, enable this precision and "
                         "time --steps launches of it at the same ef and launch shape (for "
                         "profiler runs of the bf16 and int8 traversals)")
    ap.add_argument("--brute-only", action="store_true",
                    help="time only the exact brute force (MFMA shortlist and streaming "
                         "kernel) over the workload's rows (for profiler runs) and print its "
                         "JSON line")
    ap.add_argument("--quantizer-only", action="store_true",
                    help="time only the device quantizer on the uploaded index (for "
                         "profiler runs) and print its JSON line")
    ap.add_argument("--one-pass", action="store_true",
                    help="search each batch in one pass in query order instead of the "
                         "two-pass longest-first order")
    ap.add_argument("--launch-shape", default="",
                    help="CAP,WPB: visited-set capacity and waves per block instead of the "
                         "autotuner's choice")
    ap.add_argument("--workload", choices=["cosine-1m960", "filtered-2m256", "l2-10m96"],
                    default="",
                    help="measure one of the other workloads instead of the headline: "
                         "cosine-1m960 (1M x 960 cosine, M=32, f32/int8/bf16), filtered-2m256 "
                         "(2M x 256 inner product, M=24, filtered mixed batches), l2-10m96 "
                         "(10M x 96 L2, M=16, f32/int8/bf16 + serving)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (each query is "
              "independent, so N GPUs would each hold a replica and take 1/N of the queries)",
              file=sys.stderr)
        sys.exit(2)

    import hnsw_amd as H
    from hnsw_amd import _lib as L
    from hnsw_amd.device import Buffer, Stream, synchronize

    if H.lib.hh_device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    threads = args.threads or min(16, os.cpu_count() or 1)
    if args.workload:
        return run_config(args, threads)
    efs = [int(x) for x in args.efs.split(",")]
    n, nq, dim, k = args.n, args.nq, args.dim, args.k

    log(f"generating {n} x {dim} clustered rows + {nq} queries")
    data = H.synthetic.Dataset.uniform_clusters(n, nq, dim, 1000, 0.1, 0.12, 42, threads=threads)
    if args.brute_only:
        return run_brute(args, H, data)
    keys = np.arange(n, dtype=np.uint64)
    t0 = time.time()
    if args.graph_cache and os.path.exists(args.graph_cache):
        idx = H.HnswIndex.read_from(args.graph_cache)
        if idx.node_count() != n or idx.dim() != dim:
            raise SystemExit(f"{args.graph_cache} does not hold this workload's graph")
        log(f"graph loaded from {args.graph_cache} ({time.time() - t0:.1f} s)")
    else:
        idx = H.HnswIndex(dim, H.Metric.L2, H.HnswParams(16))
        chunk = 100_000
        for s in range(0, n, chunk):
            e = min(n, s + chunk)
            idx.insert_batch(keys[s:e], data.base()[s:e], threads=args.build_threads)
            log(f"graph build: {e}/{n} linked ({time.time() - t0:.1f} s, "
                f"{args.build_threads} threads)")
        if args.graph_cache:
            idx.write_to(args.graph_cache)
    build_s = time.time() - t0

    # device buffers and the launch stream, on the library's own HIP runtime
    stream = Stream()
    sptr = stream.handle
    q_dev = Buffer.from_array(data.queries())
    out_k, out_d = Buffer(nq * k * 8), Buffer(nq * k * 4)
    out_c, status = Buffer(nq * 4), Buffer(nq * 4)
    log("uploading the index")
    h = idx.device_index()
    if args.one_pass:
        idx.set_query_order(False)
    if args.quantizer_only:
        q = measure_quantizer(H, idx, max(args.steps, 1))
        print(json.dumps({"metric": "int8 quantizer train+encode, rows/s", "unit": "rows/s",
                          "value": round(n / q["device_s"]), "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f64 sums, f32 codes",
                          "data": "synthetic", "config": {"workload": f"{n} x {dim} rows"},
                          "quantizer": q}), flush=True)
        return

    # ground truth: exact GPU brute force, checked against the FP64 oracle
    log("ground truth: GPU brute force over all rows")
    bf = H.BruteForceIndex(dim, H.Metric.L2)
    bf.insert_batch(keys, data.base())
    bf.search_arrays(data.queries()[:256], k)  # warm-up (workspace, code objects)
    t = time.time()
    tk, td, tc = bf.search_arrays(data.queries(), k)
    bf_s = time.time() - t
    mfma_used, mfma_redo = bf.mfma_stats()
    bf.set_mfma(0)
    t = time.time()
    ek, ed, ec = bf.search_arrays(data.queries(), k)
    stream_s = time.time() - t
    bf.set_mfma(1)
    same = bool(np.array_equal(ek, tk) and np.array_equal(ed.view(np.uint32), td.view(np.uint32)))
    if not same:
        raise SystemExit("MFMA brute force differs from the streaming kernel")
    gemm_tflops = 2.0 * nq * n * dim / bf_s / 1e12
    log(f"brute force: MFMA shortlist {bf_s:.3f} s ({mfma_used} queries, {mfma_redo} re-run "
        f"exactly, {gemm_tflops:.1f} TFLOP/s of 2*Q*N*D), streaming kernel {stream_s:.3f} s, "
        f"identical results")
    f64n = min(args.f64_check, nq)
    fk, fd, fc = bf.search_arrays(data.queries()[:f64n], k, f64=True)
    f64_agree = float(np.mean([len(set(tk[i].tolist()) & set(fk[i].tolist())) / k
                               for i in range(f64n)]))
    log(f"brute force {nq} queries in {bf_s:.2f} s; recall of f32 truth vs FP64 oracle "
        f"on {f64n} queries = {f64_agree:.5f}")
    del bf

    def run(ef, precision):
        st = H.lib.hh_search_batch_device(
            h, q_dev.ptr, nq, k, ef, int(precision), None,
            out_k.ptr, out_d.ptr, out_c.ptr, status.ptr, sptr)
        if st:
            raise RuntimeError(L.last_error())

    def recall():
        stream.synchronize()
        gk = out_k.download((nq, k), np.uint64)
        gc = out_c.download(nq, np.uint32)
        return H.lib.hh_recall_at_k(nq, tk.ctypes.data, tc.ctypes.data, k, gk.ctypes.data,
                                    gc.ctypes.data, k, k)

    def timed(ef, precision, steps):
        synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            run(ef, precision)
        synchronize()
        return (time.perf_counter() - t) / steps

    sweep = []
    for ef in efs:
        run(ef, H.Precision.F32)
        r = recall()
        sec = timed(ef, H.Precision.F32, 3)
        sweep.append({"ef": ef, "recall": round(r, 4), "qps": round(nq / sec)})
        log(f"f32 ef={ef}: recall@{k}={r:.4f} qps={nq / sec:,.0f}")
    chosen = next((s for s in sweep if s["recall"] >= args.target), sweep[-1])
    ef = chosen["ef"]

    # launch shape for this ef: given, or autotuned (untimed)
    cap, wpb = ctypes.c_uint32(), ctypes.c_int()
    if args.launch_shape:
        cap.value, wpb.value = (int(x) for x in args.launch_shape.split(","))
        if H.lib.hh_index_set_launch_shape(h, ef, cap.value, wpb.value):
            raise RuntimeError(L.last_error())
        how = "launch shape"
    else:
        st = H.lib.hh_index_autotune(h, q_dev.ptr, nq, k, ef, 0,
                                     out_k.ptr, out_d.ptr, out_c.ptr, status.ptr,
                                     sptr, ctypes.byref(cap), ctypes.byref(wpb))
        if st:
            raise RuntimeError(L.last_error())
        how = "autotune"
    log(f"{how} ef={ef}: visited-set capacity {cap.value}, {wpb.value} waves per block")

    # timed region: the steps issued back to back on one stream (each batch
    # starts when the previous one ends; the host does not wait between
    # them), synchronised on both sides
    for _ in range(args.warmup):
        run(ef, H.Precision.F32)
    synchronize()
    t_all = time.perf_counter()
    for _ in range(args.steps):
        run(ef, H.Precision.F32)
    synchronize()
    total = time.perf_counter() - t_all
    qps = nq * args.steps / total
    rec = recall()
    # latency of one step on its own (the host waits for each)
    step_ms = []
    for _ in range(args.steps):
        t = time.perf_counter()
        run(ef, H.Precision.F32)
        synchronize()
        step_ms.append((time.perf_counter() - t) * 1e3)

    # the same steps issued on two streams with no wait between them: one
    # batch's tail (its last, longest queries) overlaps the next batch's start
    stream2 = Stream()
    outs2 = [Buffer(nq * k * 8), Buffer(nq * k * 4), Buffer(nq * 4), Buffer(nq * 4)]

    def run2(i, ef_, precision):
        if i % 2 == 0:
            return run(ef_, precision)
        st = H.lib.hh_search_batch_device(
            h, q_dev.ptr, nq, k, ef_, int(precision), None,
            outs2[0].ptr, outs2[1].ptr, outs2[2].ptr, outs2[3].ptr, stream2.handle)
        if st:
            raise RuntimeError(L.last_error())

    for i in range(args.warmup):
        run2(i, ef, H.Precision.F32)
    synchronize()
    t_pipe = time.perf_counter()
    for i in range(args.steps):
        run2(i, ef, H.Precision.F32)
    synchronize()
    pipe_qps = nq * args.steps / (time.perf_counter() - t_pipe)
    same_pipe = bool(np.array_equal(outs2[0].download((nq, k), np.uint64),
                                    out_k.download((nq, k), np.uint64)))
    log(f"two streams: {pipe_qps:,.0f} QPS (results identical: {same_pipe})")

    # kernel time (HIP events on the launch stream) and algorithmic bytes
    H.lib.hh_index_set_timing(h, 1)
    kms = []
    for _ in range(5):
        run(ef, H.Precision.F32)
        ms = ctypes.c_float()
        H.lib.hh_index_last_kernel_ms(h, ctypes.byref(ms))
        kms.append(ms.value)
    H.lib.hh_index_set_timing(h, 0)
    idx.search_stats(reset=True)
    run(ef, H.Precision.F32)
    synchronize()
    stats = idx.search_stats()
    n_dist = stats["n_dist"] + stats["n_upper"]
    alg_bytes = n_dist * dim * 4 + stats["n_rerank"] * dim * 4
    adj_bytes = stats["n_expand"] * 32 * 4
    kern_ms = float(np.median(kms))
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    log(f"search kernel {kern_ms:.3f} ms/launch; {n_dist / nq:.0f} distance evaluations/query "
        f"({stats['n_upper'] / nq:.0f} in the upper layers), {stats['n_expand'] / nq:.1f} layer-0 "
        f"expansions/query; {achieved:.0f} GB/s of gathered rows")

    # what the device delivers to this gather pattern on the same rows with
    # no dependent work (random rows, 16 waves per CU, 4 rows per quarter
    # in flight): the practical ceiling the search's gathered rate is read
    # against
    gather_probe = None
    gbps, pms = ctypes.c_double(), ctypes.c_float()
    if H.lib.hh_index_gather_probe(h, 64, ctypes.byref(gbps), ctypes.byref(pms)) == 0:
        gather_probe = {"gbps": round(gbps.value, 1), "ms": round(pms.value, 4),
                        "search_fraction": round(achieved / gbps.value, 3)}
        log(f"gather probe: {gbps.value:.0f} GB/s for random {dim * 4}-byte rows with no "
            f"dependent work; the search gathers at {achieved / gbps.value:.0%} of it")

    variants, serve, cpu = {}, None, None
    if args.quick:
        if args.quick_precision:
            name = args.quick_precision
            prec = {"bf16": H.Precision.BF16, "int8": H.Precision.INT8,
                    "int8_perdim": H.Precision.INT8_PERDIM}[name]
            if name == "bf16":
                idx.enable_bf16()
            elif name == "int8":
                idx.enable_quantization(on_device=True)
            else:
                idx.enable_quantization_perdim()
            h = idx.device_index()
            run(ef, prec)
            r = recall()
            for _ in range(args.warmup):
                run(ef, prec)
            sec = timed(ef, prec, args.steps)
            idx.search_stats(reset=True)
            run(ef, prec)
            synchronize()
            pst = idx.search_stats()
            variants[name] = {"recall": round(r, 4), "qps": round(nq / sec),
                              "vs_f32": round(nq / sec / qps, 3),
                              "dist_evals_per_query": round((pst["n_dist"] + pst["n_upper"]) / nq, 1),
                              "expansions_per_query": round(pst["n_expand"] / nq, 1),
                              "reranked_per_query": round(pst["n_rerank"] / nq, 1)}
            log(f"{name} ef={ef}: recall={r:.4f} qps={nq / sec:,.0f} ({nq / sec / qps:.2f}x "
                f"the timed f32 steps); {variants[name]['dist_evals_per_query']} distance "
                f"evaluations, {variants[name]['expansions_per_query']} layer-0 expansions, "
                f"{variants[name]['reranked_per_query']} f32 reranks per query")
        emit(args, locals())
        return

    # precision variants at the same ef, each timed the same way as f32 is
    # here (vs_f32 is the ratio of these two timings)
    reps = max(3, args.steps // 2)
    f32_sec = timed(ef, H.Precision.F32, reps)
    idx.enable_bf16()
    quant = measure_quantizer(H, idx, 3)
    t = time.time()
    idx.enable_quantization_perdim()
    perdim_s = time.time() - t
    h = idx.device_index()
    for name, prec in (("bf16", H.Precision.BF16), ("int8", H.Precision.INT8),
                       ("int8_perdim", H.Precision.INT8_PERDIM)):
        run(ef, prec)
        r = recall()
        sec = timed(ef, prec, reps)
        variants[name] = {"recall": round(r, 4), "qps": round(nq / sec),
                          "vs_f32": round(f32_sec / sec, 3)}
        log(f"{name} ef={ef}: recall={r:.4f} qps={nq / sec:,.0f} ({f32_sec / sec:.2f}x f32)")
        # equal recall: the smallest ef at which this form reaches the
        # metric's recall target, against f32 at its own smallest such ef
        for e in sorted(set([40, 48, 52, 56, 60, 62, 64, 66, 68, 72, 80, 96, 128]) | {ef}):
            run(e, prec)
            re_ = recall()
            if re_ >= args.target:
                se = timed(e, prec, reps)
                variants[name]["at_target"] = {"ef": e, "recall": round(re_, 4),
                                               "qps": round(nq / se),
                                               "vs_f32": round(f32_sec / se, 3)}
                log(f"{name} at recall >= {args.target}: ef={e} recall={re_:.4f} "
                    f"qps={nq / se:,.0f} ({f32_sec / se:.2f}x f32 at ef={ef})")
                break
    variants["int8"]["quantizer"] = quant
    variants["int8_perdim"]["quantizer_s"] = round(perdim_s, 3)
    variants["f32_same_timing"] = {"qps": round(nq / f32_sec)}
    variants["beam_width"] = idx.beam_width()

    # serving loop: Poisson arrivals replayed by the native load generator
    try:
        rate = min(0.5 * qps, 400_000.0)
        nreq = 50_000
        rng = np.random.default_rng(7)
        arrivals = np.cumsum(rng.exponential(1e6 / rate, nreq))
        qs = data.queries()[np.arange(nreq) % nq]
        with H.Server(idx, k_max=k, max_inflight=2) as srv:
            warm_batches = warm_server(srv, qs, k, ef)
            t = time.perf_counter()
            lat, sk, sd, sc, sst = srv.replay(qs, np.full(nreq, k), np.full(nreq, ef), arrivals)
            wall = time.perf_counter() - t
            sstats = srv.stats()
        sk_, sc_ = np.ascontiguousarray(sk[:nq]), np.ascontiguousarray(sc[:nq])
        served_recall = H.lib.hh_recall_at_k(nq, tk.ctypes.data, tc.ctypes.data, k,
                                             sk_.ctypes.data, sc_.ctypes.data, k, k)
        serve = {"offered_qps": round(rate), "achieved_qps": round(nreq / wall),
                 "p50_ms": round(float(np.percentile(lat, 50)), 3),
                 "p99_ms": round(float(np.percentile(lat, 99)), 3),
                 "requests": nreq, "launches": sstats["batches"] - warm_batches,
                 "recall_at_10": round(served_recall, 4)}
        log(f"serving: {serve}")
    except Exception as e:
        log(f"serving measurement failed: {e}")

    # CPU baseline: the oracle's restatement of search_batch on host cores
    try:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        v = idx.graph_arrays()
        g = O.OracleGraph(v["dim"], v["metric"], v["m"], v["m0"], v["vectors"], v["keys"],
                          v["levels"], v["layer0"], v["upper_off"], v["upper"], v["entry_id"],
                          v["top"])
        import resource
        ns = min(args.cpu_sample, nq)
        sample = data.queries()[:ns]
        load0 = os.getloadavg()[0]
        # all the cores this process may run on, up to the host share of
        # one GPU (16 cores on these hosts), however many the OS lists
        cores = min(len(os.sched_getaffinity(0)), 16)
        # warm-up pass (also the recall of the CPU results), then the median
        # of 3 timed passes on all those cores (wall clock, as the crate's
        # 8-thread figures), then the median of 3 single-thread passes timed
        # per CPU-second of the searching thread (getrusage, as the crate's
        # hnsw-bench reports single-thread throughput)
        ok, od, oc, _ = g.search_batch(sample, k, ef, threads=cores)
        cpu_recall = O.recall_at_k(tk[:ns], tc[:ns], ok, oc, k)
        passes, proc_cpu = [], []
        for _ in range(3):
            r0 = resource.getrusage(resource.RUSAGE_SELF)
            t = time.perf_counter()
            g.search_batch(sample, k, ef, threads=cores)
            passes.append(time.perf_counter() - t)
            r1 = resource.getrusage(resource.RUSAGE_SELF)
            proc_cpu.append((r1.ru_utime + r1.ru_stime) - (r0.ru_utime + r0.ru_stime))
        cpu_s = sorted(passes)[1]
        n1 = min(ns, 2000)
        one_wall, one_cpu = [], []
        for _ in range(3):
            r0 = resource.getrusage(resource.RUSAGE_THREAD)
            t = time.perf_counter()
            g.search_batch(sample[:n1], k, ef, threads=1)  # runs on this thread
            one_wall.append(time.perf_counter() - t)
            r1 = resource.getrusage(resource.RUSAGE_THREAD)
            one_cpu.append((r1.ru_utime + r1.ru_stime) - (r0.ru_utime + r0.ru_stime))
        one_s, one_c = sorted(one_wall)[1], sorted(one_cpu)[1]
        load1 = os.getloadavg()[0]
        cpu = {"value": round(ns / cpu_s), "unit": "queries/s", "cores": cores, "kind": "port",
               "sample": f"{ns} of the {nq} queries, same graph, k={k}, ef={ef}: warm-up pass, "
                         f"then median of 3 passes ({', '.join(f'{p:.2f}' for p in passes)} s) "
                         f"on {cores} threads (the host's {os.cpu_count()} CPUs, of which this "
                         f"process may use {len(os.sched_getaffinity(0))}); oracle C restatement "
                         "of search_batch with AVX2-order kernels",
               "host_cpus": os.cpu_count(), "cpus_allowed": len(os.sched_getaffinity(0)),
               "process_cpu_s_per_pass": round(sorted(proc_cpu)[1], 3),
               "single_thread_qps_per_cpu_s": round(n1 / one_c),
               "single_thread_qps_wall": round(n1 / one_s),
               "single_thread_sample": f"{n1} queries, median of 3 passes",
               "recall_at_10": round(cpu_recall, 4),
               "loadavg_start_end": [round(load0, 2), round(load1, 2)]}
        log(f"cpu baseline {ns / cpu_s:,.0f} qps on {cores} threads, "
            f"{n1 / one_c:,.0f} qps per CPU-second on 1 thread, recall {cpu_recall:.4f}")
    except Exception as e:  # the baseline is reported, never required
        log(f"cpu baseline unavailable: {e}")

    # keeping the resident index in step with the graph: 10,000 inserts (1%)
    # and 1,000 removals, then the incremental sync against a full upload
    mutation = None
    try:
        extra = H.synthetic.Dataset.uniform_clusters(10_000, 1, dim, 1000, 0.1, 0.12, 43,
                                                     threads=threads).base()
        idx.device_index()  # nothing pending
        t0 = time.perf_counter()
        idx.insert_batch(np.arange(n, n + 10_000, dtype=np.uint64), extra, threads=threads)
        for key in range(0, 1_000_000, 1000):
            idx.remove(key)
        host_s = time.perf_counter() - t0
        ids = idx._take_dirty()
        d = idx._view()
        synchronize()
        t0 = time.perf_counter()
        rc = H.lib.hh_index_sync(idx._idx, ctypes.byref(d), ids.ctypes.data if ids.size else None,
                                 ids.size)
        sync_rows_s = time.perf_counter() - t0
        if rc:
            raise RuntimeError(L.last_error())
        idx._stale = False
        pending = int(ids.size)
        t0 = time.perf_counter()
        h2 = ctypes.c_void_p()
        H._check(H.lib.hh_index_upload(ctypes.byref(d), 0, ctypes.byref(h2)))
        synchronize()
        full_s = time.perf_counter() - t0
        H.lib.hh_index_free(h2)
        mutation = {"inserted": 10_000, "removed": 1000, "host_insert_s": round(host_s, 3),
                    "changed_link_lists": pending, "sync_s": round(sync_rows_s, 4),
                    "full_upload_s": round(full_s, 3)}
        log(f"mutation: 10,000 inserts + 1,000 removals on the host in {host_s:.2f} s "
            f"({pending} link lists changed); sync {sync_rows_s * 1e3:.1f} ms vs full upload "
            f"{full_s * 1e3:.0f} ms")
    except Exception as e:
        log(f"mutation measurement unavailable: {e}")

    # the same graph built with INSERT's layer-0 searches on the GPU
    device_build = None
    try:
        t0 = time.perf_counter()
        di = H.HnswIndex(dim, H.Metric.L2, H.HnswParams(16))
        di.insert_batch(keys, data.base(), threads=threads, device=True)
        dev_s = time.perf_counter() - t0
        dk, dd, dc = di.search_arrays(data.queries(), k, ef)
        dr = H.lib.hh_recall_at_k(nq, tk.ctypes.data, tc.ctypes.data, k, dk.ctypes.data,
                                  dc.ctypes.data, k, k)
        device_build = {"seconds": round(dev_s, 2), "host_build_s": round(build_s, 2),
                        "host_build_threads": args.build_threads, "threads": threads,
                        "recall_at_10": round(dr, 4), "ef": ef}
        log(f"device-assisted build: {dev_s:.1f} s on {threads} host threads (the benchmark "
            f"graph's host build: {build_s:.1f} s on {args.build_threads}), recall@{k} "
            f"at ef={ef} {dr:.4f} (host-built graph {rec:.4f})")
        del di
    except Exception as e:
        log(f"device-assisted build unavailable: {e}")

    emit(args, locals())


def run_brute(args, H, data):
    """The exact brute force over the workload: BruteForceIndex::search_batch
    of every query over every row, timed on the MFMA shortlist path (proven
    exact per query, re-run by the streaming kernel otherwise) and on the
    streaming kernel; both must return identical results.  Algorithmic work
    2*Q*N*D flops."""
    from hnsw_amd.device import synchronize
    n, nq, dim, k = data.base_len(), data.query_len(), data.dim(), args.k
    bf = H.BruteForceIndex(dim, H.Metric.L2)
    bf.insert_batch(np.arange(n, dtype=np.uint64), data.base())
    bf.search_arrays(data.queries()[:256], k)
    times = []
    for _ in range(max(args.steps, 1)):
        synchronize()
        t = time.perf_counter()
        tk, td, tc = bf.search_arrays(data.queries(), k)
        times.append(time.perf_counter() - t)
    used, redo = bf.mfma_stats()
    bf.set_mfma(0)
    t = time.perf_counter()
    ek, ed, ec = bf.search_arrays(data.queries(), k)
    stream_s = time.perf_counter() - t
    same = bool(np.array_equal(ek, tk) and np.array_equal(ed.view(np.uint32), td.view(np.uint32)))
    best = min(times)
    line = {"metric": "exact brute-force kNN, 2*Q*N*D TFLOP/s (whole call)", "unit": "TFLOP/s",
            "value": round(2.0 * nq * n * dim / best / 1e12, 2), "n_gpus": 1,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"BruteForceIndex::search_batch, {n} x {dim} L2, {nq} queries, "
                                   f"k={k}"},
            "mfma_call_s": round(best, 4), "streaming_call_s": round(stream_s, 4),
            "mfma_queries": used, "rerun_exactly": redo, "identical_results": same,
            "mfma_counters": measured_mfma()}
    print(json.dumps(line), flush=True)


def build_index(H, data, metric, m, threads):
    n = data.base_len()
    idx = H.HnswIndex(data.dim(), metric, H.HnswParams(m))
    keys = np.arange(n, dtype=np.uint64)
    t0 = time.time()
    chunk = 200_000
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        idx.insert_batch(keys[s:e], data.base()[s:e], threads=threads)
        log(f"graph build: {e}/{n} linked ({time.time() - t0:.1f} s, {threads} threads)")
    return idx, keys, time.time() - t0


class DeviceBatch:
    """Device copies of a query set and result buffers for k_stride results."""

    def __init__(self, H, queries, k_stride):
        from hnsw_amd.device import Buffer, Stream
        self.nq, self.k_stride = queries.shape[0], k_stride
        self.stream = Stream()
        self.q = Buffer.from_array(np.ascontiguousarray(queries, np.float32))
        nq = self.nq
        self.ok, self.od = Buffer(nq * k_stride * 8), Buffer(nq * k_stride * 4)
        self.oc, self.st = Buffer(nq * 4), Buffer(nq * 4)

    def results(self, k=0):
        """Keys, distances and counts of the last launch; a uniform launch
        packs its results at stride k, a mixed one at k_stride."""
        k = k or self.k_stride
        self.stream.synchronize()
        return (self.ok.download((self.nq, k), np.uint64),
                self.od.download((self.nq, k), np.float32),
                self.oc.download(self.nq, np.uint32))


def qps_of(launch, nq, reps=5):
    from hnsw_amd.device import synchronize
    launch()
    synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        launch()
    synchronize()
    return nq * reps / (time.perf_counter() - t)


def run_config(args, threads):
    """The other workloads of the query path, each printed as one JSON line:
    recall against the exact GPU brute force (spot-checked against FP64) and
    throughput with queries resident in HBM."""
    import hnsw_amd as H
    from hnsw_amd import _lib as L
    from hnsw_amd.device import Buffer
    out = {"metric": f"query-path measurements, workload {args.workload}", "n_gpus": 1,
           "data": "synthetic", "higher_is_better": True, "unit": "queries/s"}

    def uniform_search(idx, batch, k, ef, prec):
        h = idx.device_index()

        def launch():
            if H.lib.hh_search_batch_device(h, batch.q.ptr, batch.nq, k, ef, int(prec), None,
                                            batch.ok.ptr, batch.od.ptr, batch.oc.ptr,
                                            batch.st.ptr, batch.stream.handle):
                raise RuntimeError(L.last_error())
        return launch

    def truth_of(data, metric, k):
        bf = H.BruteForceIndex(data.dim(), metric)
        bf.insert_batch(np.arange(data.base_len(), dtype=np.uint64), data.base())
        t = time.time()
        tk, td, tc = bf.search_arrays(data.queries(), k)
        log(f"ground truth: {data.query_len()} queries, k={k}, in {time.time() - t:.2f} s")
        nf = min(200, data.query_len())
        fk, fd, fc = bf.search_arrays(data.queries()[:nf], 10, f64=True)
        agree = float(np.mean([len(set(tk[i, :10].tolist()) & set(fk[i].tolist())) / 10
                               for i in range(nf)]))
        return bf, tk, tc, fk, fd, agree

    if args.workload in ("cosine-1m960", "l2-10m96"):
        if args.workload == "cosine-1m960":
            n, nq, dim, m, metric = args.n, 10_000, 960, 32, H.Metric.Cosine
            log(f"generating {n} x {dim} U(0,1) rows, L2-normalised, + {nq} queries")
            data = H.synthetic.Dataset.unit_uniform(n, nq, dim, 42, threads=threads)
            ks, efs = (10, 100), (64, 256)
        else:
            n = 10_000_000 if args.n == 1_000_000 else args.n
            nq, dim, m, metric = 10_000, 96, 16, H.Metric.L2
            log(f"generating {n} x {dim} rows in 8 clusters (sigma 0.3) + {nq} queries")
            data = H.synthetic.Dataset.gauss_clusters(n, nq, dim, 8, 0.3, 42, threads=threads)
            ks, efs = (10,), (64,)
        idx, keys, build_s = build_index(H, data, metric, m, threads)
        bf, tk, tc, fk, fd, agree = truth_of(data, metric, max(ks))
        del bf
        idx.device_index()
        t = time.time()
        idx.enable_quantization(on_device=True)
        quant_s = time.time() - t
        idx.enable_bf16()
        t = time.time()
        idx.enable_quantization_perdim()
        perdim_s = time.time() - t
        batch = DeviceBatch(H, data.queries(), max(ks))
        rows = []
        for prec in (H.Precision.F32, H.Precision.INT8, H.Precision.INT8_PERDIM, H.Precision.BF16):
            for k in ks:
                for ef in efs:
                    launch = uniform_search(idx, batch, k, ef, prec)
                    idx.search_stats(reset=True)
                    launch()
                    gk, gd, gc = batch.results(k)
                    sts = idx.search_stats()
                    rec = H.lib.hh_recall_at_k(nq, tk.ctypes.data, tc.ctypes.data, tk.shape[1],
                                               gk.ctypes.data, gc.ctypes.data, k, k)
                    # returned distances against the FP64 oracle's, same keys
                    err = 0.0
                    for i in range(fk.shape[0]):
                        want = dict(zip(fk[i].tolist(), fd[i].tolist()))
                        for j in range(min(int(gc[i]), 10)):
                            if int(gk[i, j]) in want:
                                err = max(err, abs(float(gd[i, j]) - want[int(gk[i, j])]))
                    q = qps_of(launch, nq)
                    rows.append({"precision": prec.name.lower(), "k": k, "ef": ef,
                                 "recall": round(rec, 4), "qps": round(q),
                                 "max_abs_dist_err_vs_f64": float(f"{err:.3g}"),
                                 "dist_evals_per_query": round(
                                     (sts["n_dist"] + sts["n_upper"]) / nq, 1),
                                 "expansions_per_query": round(sts["n_expand"] / nq, 1)})
                    log(f"{prec.name.lower()} k={k} ef={ef}: recall@{k}={rec:.4f} "
                        f"qps={q:,.0f} max |d - d_f64| {err:.3g}")
        out.update({"value": max(r["qps"] for r in rows if r["precision"] == "f32"),
                    "dtype": "f32 / int8 (crate codes, per-dimension codes) / bf16",
                    "results": rows, "perdim_quantizer_s": round(perdim_s, 3),
                    "config": {"workload": f"hnsw search_batch, {n} x {dim} "
                                           f"{'cosine' if metric == H.Metric.Cosine else 'L2'}, "
                                           f"M={m}, efC=200, {nq} queries",
                               "graph_build_s": round(build_s, 1),
                               "quantizer_train_encode_s": round(quant_s, 3),
                               "truth_vs_f64_agreement": round(agree, 5)}})
        if args.workload == "l2-10m96":
            serve = []
            qs = data.queries()
            for rate in (50_000.0, 200_000.0, 500_000.0):
                for prec in (H.Precision.BF16, H.Precision.INT8):
                    nreq = 50_000
                    rng = np.random.default_rng(7)
                    arrivals = np.cumsum(rng.exponential(1e6 / rate, nreq))
                    with H.Server(idx, k_max=10, max_inflight=2) as srv:
                        warm_server(srv, qs, 10, 64, prec)
                        t = time.perf_counter()
                        lat, sk, sd, sc, sst = srv.replay(
                            qs[np.arange(nreq) % nq], np.full(nreq, 10), np.full(nreq, 64),
                            arrivals, precisions=np.full(nreq, int(prec)))
                        wall = time.perf_counter() - t
                    sk_, sc_ = np.ascontiguousarray(sk[:nq]), np.ascontiguousarray(sc[:nq])
                    rec = H.lib.hh_recall_at_k(nq, tk.ctypes.data, tc.ctypes.data, tk.shape[1],
                                               sk_.ctypes.data, sc_.ctypes.data, 10, 10)
                    row = {"offered_qps": int(rate), "precision": prec.name.lower(),
                           "achieved_qps": round(nreq / wall),
                           "p50_ms": round(float(np.percentile(lat, 50)), 3),
                           "p99_ms": round(float(np.percentile(lat, 99)), 3),
                           "recall_at_10": round(rec, 4)}
                    serve.append(row)
                    log(f"serving: {row}")
            out["serving"] = serve
        print(json.dumps(out), flush=True)
        return

    # filtered-2m256: filtered mixed batches
    n = 2_000_000 if args.n == 1_000_000 else args.n
    nq, dim, m, cats = 50_000, 256, 24, 20
    log(f"generating {n} x {dim} N(0,1) rows with lognormal(0,0.5) norms, {cats} categories")
    data = H.synthetic.Dataset.lognormal(n, nq, dim, 0.5, cats, 42, threads=threads)
    idx, keys, build_s = build_index(H, data, H.Metric.InnerProduct, m, threads)
    h = idx.device_index()
    rng = np.random.default_rng(11)
    nfilt = 64
    masks, pass_rates = [], []
    for f in range(nfilt):
        chosen = rng.choice(cats, rng.integers(1, 4), replace=False)
        mask = np.isin(data.categories, chosen)
        masks.append(mask)
        pass_rates.append(float(mask.mean()))
    words = (n + 31) // 32
    table = np.zeros((nfilt, words), np.uint32)
    for f, mask in enumerate(masks):
        ids = np.nonzero(mask)[0]
        np.bitwise_or.at(table[f], ids >> 5, (np.uint32(1) << (ids & 31).astype(np.uint32)))
    kq = rng.choice([1, 10, 100], nq).astype(np.uint32)
    efq = np.maximum(rng.integers(32, 513, nq), kq).astype(np.uint32)
    fq = rng.integers(0, nfilt, nq).astype(np.uint32)
    specs = np.stack([kq, efq, np.zeros(nq, np.uint32), fq], axis=1).astype(np.uint32)
    batch = DeviceBatch(H, data.queries(), 100)
    s_dev, f_dev = Buffer.from_array(specs), Buffer.from_array(table)

    def launch():
        if H.lib.hh_search_batch_spec_device(h, batch.q.ptr, nq, s_dev.ptr, 100, 512, f_dev.ptr,
                                             nfilt, batch.ok.ptr, batch.od.ptr, batch.oc.ptr,
                                             batch.st.ptr, batch.stream.handle):
            raise RuntimeError(L.last_error())
    idx.search_stats(reset=True)
    launch()
    gk, gd, gc = batch.results()
    sts = idx.search_stats()
    q_direct = qps_of(launch, nq, 3)
    per_q = {f: round(sts[f] / nq, 1) for f in ("n_dist", "n_upper", "n_expand", "n_overflow",
                                                  "n_visited_spill", "n_cand_spill")}
    log(f"per query: {per_q}; gathered rows at the direct rate "
        f"{(sts['n_dist'] + sts['n_upper']) * dim * 4 * q_direct / nq / 1e9:.0f} GB/s")
    plan = ctypes.c_void_p()
    if H.lib.hh_search_plan_create(h, batch.q.ptr, nq, 100, 512, 0, None, s_dev.ptr, 100,
                                   f_dev.ptr, nfilt, batch.ok.ptr, batch.od.ptr, batch.oc.ptr,
                                   batch.st.ptr, ctypes.byref(plan)):
        raise RuntimeError(L.last_error())

    def replay():
        if H.lib.hh_search_plan_launch(plan, batch.stream.handle):
            raise RuntimeError(L.last_error())
    q_plan = qps_of(replay, nq, 3)
    pk, pd, pc = batch.results()
    H.lib.hh_search_plan_free(plan)
    same = bool(np.array_equal(pk, gk) and np.array_equal(pc, gc))
    log(f"filtered mixed batch: {q_direct:,.0f} QPS direct, {q_plan:,.0f} QPS graph replay "
        f"(identical: {same})")
    # recall against the filtered exact brute force, query by query at its own k
    bf = H.BruteForceIndex(dim, H.Metric.InnerProduct)
    bf.insert_batch(keys, data.base())
    hits = expected = 0
    t = time.time()
    for f in range(nfilt):
        sel = np.nonzero(fq == f)[0]
        if sel.size == 0:
            continue
        tk, td, tc = bf.search_arrays(data.queries()[sel], 100, filter=masks[f])
        for r, i in enumerate(sel):
            kk = int(kq[i])
            tn, fn = min(int(tc[r]), kk), min(int(gc[i]), kk)
            expected += tn
            hits += len(set(tk[r, :tn].tolist()) & set(gk[i, :fn].tolist()))
    rec = hits / max(expected, 1)
    log(f"recall at each query's k vs filtered brute force: {rec:.4f} ({time.time() - t:.1f} s)")
    out.update({"value": round(q_plan), "dtype": "f32",
                "results": {"direct_launch_qps": round(q_direct), "graph_replay_qps": round(q_plan),
                            "recall_at_k": round(rec, 4), "replay_identical": same,
                            "per_query": per_q,
                            "filter_pass_rate_range": [round(min(pass_rates), 3),
                                                       round(max(pass_rates), 3)]},
                "config": {"workload": f"hnsw mixed filtered batch, {n} x {dim} inner product, "
                                       f"M={m}, efC=200, {nq} queries, k in {{1,10,100}}, "
                                       "ef in [32,512], 64 bitmaps of 1-3 of 20 categories",
                           "graph_build_s": round(build_s, 1)}})
    print(json.dumps(out), flush=True)


TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "search_1m128_l2_traffic.md")
MFMA_SUMMARY = os.path.join(ROOT, "profiles", "brute_mfma_1m128_counters.md")


def measured_mfma():
    """MFMA utilisation and flops of brute_mfma_kernel from the committed
    rocprofv3 counter summary of this workload (keys mfma_util_percent,
    mfma_tflops_counted), or None when it is absent."""
    out = {}
    try:
        with open(MFMA_SUMMARY) as f:
            for line in f:
                for key in ("mfma_util_percent:", "mfma_flops_per_launch:", "mfma_tflops_counted:",
                            "main_pass_ms:"):
                    if line.startswith(key):
                        out[key[:-1]] = float(line.split(":", 1)[1])
    except (OSError, ValueError):
        return None
    if out:
        out["source"] = os.path.relpath(MFMA_SUMMARY, ROOT)
    return out or None


def measured_traffic():
    """HBM bytes per search launch (FETCH_SIZE + WRITE_SIZE counter passes,
    calibrated on a copy kernel) from the committed counter summary of this
    workload, or None when it is absent."""
    try:
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def emit(args, v):
    """Prints the one JSON line of the run from main()'s values."""
    qps, total, nq, k, ef, rec = v["qps"], v["total"], v["nq"], v["k"], v["ef"], v["rec"]
    n, dim, achieved, cpu, step_ms = v["n"], v["dim"], v["achieved"], v["cpu"], v["step_ms"]
    kern_ms, n_dist, alg_bytes, adj_bytes = v["kern_ms"], v["n_dist"], v["alg_bytes"], v["adj_bytes"]
    sweep, variants, serve = v["sweep"], v["variants"], v["serve"]
    bf_s, f64_agree, mfma_used, mfma_redo = v["bf_s"], v["f64_agree"], v["mfma_used"], v["mfma_redo"]
    stream_s, gemm_tflops, build_s, cap, wpb = (v["stream_s"], v["gemm_tflops"], v["build_s"],
                                                v["cap"], v["wpb"])
    line = {
        "metric": "QPS at recall@10>=0.95, 1Mx128 L2 (GPU FP32 search_batch)",
        "value": round(qps, 1),
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "hnsw search_batch, 1M x 128 clustered L2, M=16, efC=200, "
                               f"{nq} queries, k={k}",
                   "ef": ef, "recall_at_10": round(rec, 4), "n": n, "dim": dim,
                   "queries": nq, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic()},
        "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT) if measured_traffic() else None,
        "cpu_baseline": cpu,
        "step_ms_p50": round(float(np.percentile(step_ms, 50)), 3),
        "step_ms_p99": round(float(np.percentile(step_ms, 99)), 3),
        "search_kernel_ms": round(kern_ms, 4),
        "two_stream_qps": round(v["pipe_qps"]),
        "gather_probe": v.get("gather_probe"),
        "index_sync": v.get("mutation"),
        "device_assisted_build": v.get("device_build"),
        "dist_evals_per_query": round(n_dist / nq, 1),
        "upper_evals_per_query": round(v["stats"]["n_upper"] / nq, 1),
        "expansions_per_query": round(v["stats"]["n_expand"] / nq, 1),
        "gathered_bytes_per_launch": alg_bytes,
        "adjacency_bytes_per_launch": adj_bytes,
        "sweep_f32": sweep,
        "variants": variants,
        "serving": serve,
        "ground_truth": {"brute_force_s": round(bf_s, 3), "f64_agreement": round(f64_agree, 5),
                         "mfma_queries": mfma_used, "mfma_rerun_exact": mfma_redo,
                         "streaming_kernel_s": round(stream_s, 3),
                         "gemm_tflops_2QND": round(gemm_tflops, 2),
                         "mfma_counters": measured_mfma()},
        "graph_build_s": round(build_s, 1),
        "graph_build_threads": args.build_threads,
        "launch_shape": {"visited_set_capacity": cap.value, "waves_per_block": wpb.value},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
