"""Benchmark of the hot path: stable key-value radix sort (satradix) of 2^28
pairs, uint64 keys uniform over 64 bits (seed 0xC3A), uint32 payload = index.

One step = one clo_sort_kv_with_device_data call, out of place (the unsorted
input stays resident in HBM and is re-sorted every step).  Prints one JSON
line.  `roofline` is for the dominant kernel (satradix_downsweep, one per
digit pass): algorithmic bytes per launch = n * 2 * (8 + 4) divided by its
average launch duration measured with HIP events on the exec stream over the
timed region.  `cpu_baseline` times the oracle's reference-structured satradix
(4-bit digits, 256-element blocks, blocks spread over host threads) on a
bounded sample of the same workload.

`dispatch` reports the learned dispatcher on the mixed sort/scan request
stream: regret against the exhaustive tuner and speedup over the fixed
configuration (see dispatch_eval).

Other modes (not the headline line):
  --copy-probe     runs only the calibration copy kernel (for memory counters)
  --suite          also measures the other configurations, one JSON line each
                   on stderr
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "cl-ops_amd")]

import clops  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "radix sort Gkeys/s (2^28 u64 KV), % of HBM roofline; learned-dispatch regret"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_steps(q, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    q.finish()
    q.prof_reset()
    e0 = q.marker()
    for _ in range(steps):
        fn()
    e1 = q.marker()
    q.finish()
    return q.elapsed_ms(e0, e1) / steps


def cpu_baseline_kv(n_sample, threads):
    """Reference-structured satradix on the host (the oracle), KV u64/u32."""
    sys.path[:0] = [os.path.join(ROOT, "oracle")]
    import oracle
    keys = oracle.generate(oracle.GEN_U64, 0xC3A, n_sample)
    vals = np.arange(n_sample, dtype=np.uint32)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        ko, vo = oracle.satradix(keys, vals, bits=4, lws=256, threads=threads)
        times.append(time.perf_counter() - t0)
    t = min(times)
    ok = bool(np.all(ko[:-1] <= ko[1:]))
    del ko, vo
    # Context: the same decomposition on one thread (a smaller sample), and a
    # plain serial LSD radix sort (8-bit digits) of the same pairs.
    n1 = min(n_sample, 1 << 22)
    t1 = _best_time(lambda: oracle.satradix(keys[:n1], vals[:n1], bits=4, lws=256, threads=1), 1)
    t_std = _best_time(lambda: oracle.stable_sort(keys[:n1], vals[:n1]), 1)
    return {
        "value": n_sample / t / 1e9,
        "unit": "Gkeys/s",
        "cores": threads,
        "kind": "port",
        "sample": "%d KV pairs (u64 key U(0xC3A,i), u32 index), oracle satradix restating "
                  "the reference decomposition (4-bit digits, 16 passes, 256-element blocks: "
                  "split-sort, histogram, counter scan, scatter) on %d threads of %s, best of "
                  "2, %.2f s, sorted=%s" % (n_sample, threads, cpu_model(), t, ok),
        "value_1_thread": n1 / t1 / 1e9,
        "serial_lsd_1_thread": n1 / t_std / 1e9,
        "sample_1_thread": "%d KV pairs, the same decomposition on 1 thread (%.2f s); a "
                           "plain serial LSD radix sort, 8-bit digits (%.2f s)"
                           % (n1, t1, t_std),
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def traffic_from_profiles():
    """HBM bytes per satradix_downsweep launch from the committed counter
    summary (FETCH_SIZE and WRITE_SIZE passes scaled by a copy kernel's)."""
    import re
    p = os.path.join(ROOT, "profiles", "satradix_downsweep_kv_u64_2p28_traffic.md")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        m = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", f.read(), re.M)
    return int(m.group(1)) if m else None


def verify_kv(q, kin, kout, vout, n, width):
    """Device checks of a KV sort whose input values were 0..n-1: keys
    non-decreasing with values increasing inside runs of equal keys, values a
    permutation of 0..n-1, and each output key the input key its value names.
    Zero violations of both means the output is exactly the stable sort."""
    return (clops.check_sorted(q, kout, n, width, vout),
            clops.check_pairs(q, kin, kout, vout, n, width))


def headline(args, ctx):
    n = 1 << args.log2n
    q = clops.Queue(ctx, profiling=True)
    kin = clops.Buffer(ctx, n * 8)
    vin = clops.Buffer(ctx, n * 4)
    kout = clops.Buffer(ctx, n * 8)
    vout = clops.Buffer(ctx, n * 4)
    clops.generate(q, kin, n, clops.GEN_U64, 0xC3A)
    clops.generate(q, vin, n, clops.GEN_IOTA_U32, 0)
    q.finish()
    s = clops.Sorter(ctx, "satradix", np.uint64)

    def step(qq):
        s.sort_kv_device(qq, kin, vin, kout, vout, n)

    # Timed region 1 (value): the steps with no per-kernel events, only the
    # markers around the region.  Timed region 2 (roofline): the same steps
    # with HIP events around the dominant kernel's launches on the exec
    # stream (8 of the ~40 launches of a sort).  Events around kernels cost
    # time of their own (0.2-0.3 ms per step here), so the un-instrumented
    # region gives the throughput.  A third pass with events around every
    # kernel gives the per-kernel breakdown for the log.
    q_plain = clops.Queue(ctx)
    ms = time_steps(q_plain, lambda: step(q_plain), args.steps, args.warmup)
    q.prof_filter("satradix_downsweep")
    ms_ds = time_steps(q, lambda: step(q), args.steps, 1)
    cnt, tot = q.prof("satradix_downsweep")
    q.prof_filter("")
    ms_all = time_steps(q, lambda: step(q), args.steps, 1)
    log("  same steps with events around the downsweeps: %.3f ms/step; around every kernel: "
        "%.3f ms/step; without events: %.3f ms/step" % (ms_ds, ms_all, ms))
    for name in q.prof_names():
        c, t = q.prof(name)
        log("  kernel %-22s launches/step %5.1f  avg %.4f ms  total/step %.4f ms"
            % (name, c / args.steps, t / max(c, 1), t / args.steps))
    hcnt, htot = q.prof("satradix_upsweep")
    passes = s.last_passes()
    scatter_ms = tot / max(cnt, 1)
    bytes_per_launch = n * 2 * (8 + 4)
    achieved = bytes_per_launch / (scatter_ms * 1e-3) / 1e9
    # verification of the last step's output on the device
    bad, pairs = verify_kv(q, kin, kout, vout, n, 8)
    same = clops.multiset_hash(q, kin, n, 8) == clops.multiset_hash(q, kout, n, 8)
    total_bytes = n * 8 + passes * bytes_per_launch
    log("headline: %.3f ms/step, downsweep %.3f ms x %d, upsweep %.3f ms, passes %d, "
        "sorted/stability violations %d, pairing/permutation violations %d, multiset-equal %s, "
        "whole-sort HBM fraction %.3f"
        % (ms, scatter_ms, cnt // max(args.steps, 1), htot / max(hcnt, 1), passes, bad, pairs,
           same, total_bytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS))
    if bad != 0 or pairs != 0 or not same:
        raise SystemExit("benchmark output failed verification")
    return {
        "ms": ms, "ms_downsweep_events": ms_ds, "n": n, "passes": passes,
        "scatter_ms": scatter_ms, "achieved": achieved, "bytes_per_launch": bytes_per_launch,
    }


def suite(ctx, cpu_threads=0):
    """Other configurations, one stderr JSON line each (with the CPU
    restatement timed beside them when cpu_threads > 0)."""
    q = clops.Queue(ctx, profiling=True)
    out = []
    # C1: 2^20 uint32 keys only.
    n = 1 << 20
    kin = clops.Buffer(ctx, n * 4)
    kout = clops.Buffer(ctx, n * 4)
    clops.generate(q, kin, n, clops.GEN_U32_HI, 0xC1)
    s = clops.Sorter(ctx, "satradix", np.uint32)
    qp = q
    q = clops.Queue(ctx)  # no per-kernel events: this size is launch-bound
    ms = time_steps(q, lambda: s.sort_device(q, kin, kout, n), 20, 5)
    ms_flush = None
    flush = clops.Buffer(ctx, 512 << 20)
    t = 0.0
    for _ in range(5):  # each sort after a 512 MiB write that evicts the caches
        clops.generate(q, flush, (512 << 20) // 4, clops.GEN_U32_HI, 1)
        e0 = q.marker()
        s.sort_device(q, kin, kout, n)
        e1 = q.marker()
        q.finish()
        t += q.elapsed_ms(e0, e1)
    ms_flush = t / 5
    g = clops.Graph(q)
    g.begin()
    try:
        s.sort_device(q, kin, kout, n)
    finally:
        g.end()
    ms_graph = time_steps(q, g.launch, 20, 5)
    bad = clops.check_sorted(q, kout, n, 4, None)
    g.close()
    out.append({"config": "satradix u32 2^20 keys", "ms": ms, "Gkeys/s": n / ms / 1e6,
                "ms_after_cache_flush": ms_flush, "ms_graph_replay": ms_graph,
                "unsorted_pairs": bad})
    q = qp
    # C3 variants: [0, 2^20) keys and Zipf.
    n = 1 << 28
    kin = clops.Buffer(ctx, n * 8)
    vin = clops.Buffer(ctx, n * 4)
    kout = clops.Buffer(ctx, n * 8)
    vout = clops.Buffer(ctx, n * 4)
    clops.generate(q, vin, n, clops.GEN_IOTA_U32, 0)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    clops.generate(q, kin, n, clops.GEN_U64_MASK, 0xC3B, (1 << 20) - 1)
    ms = time_steps(q, lambda: s.sort_kv_device(q, kin, vin, kout, vout, n), 5, 2)
    p = s.last_passes()
    bad, pairs = verify_kv(q, kin, kout, vout, n, 8)
    out.append({"config": "satradix KV u64 2^28 keys<2^20", "ms": ms, "Gkeys/s": n / ms / 1e6,
                "passes": p, "hbm_frac": n * (8 + 24 * p) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "order_violations": bad, "pair_violations": pairs})
    if bad or pairs:
        raise SystemExit("keys<2^20 sort failed verification")
    cdfh = np.empty(1 << 24, dtype=np.float64)
    import ctypes
    clops.lib().clo_util_zipf_cdf(cdfh.ctypes.data_as(ctypes.c_void_p), cdfh.size, 1.1)
    cdf = clops.Buffer.from_array(ctx, cdfh)
    clops._call(clops.lib().clo_util_generate_zipf, q.ptr, kin.ptr, n, cdf.ptr, cdfh.size, 0xC3C)
    ms = time_steps(q, lambda: s.sort_kv_device(q, kin, vin, kout, vout, n), 5, 2)
    p = s.last_passes()
    bad, pairs = verify_kv(q, kin, kout, vout, n, 8)
    out.append({"config": "satradix KV u64 2^28 zipf1.1", "ms": ms, "Gkeys/s": n / ms / 1e6,
                "passes": p, "hbm_frac": n * (8 + 24 * p) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                "order_violations": bad, "pair_violations": pairs})
    if bad or pairs:
        raise SystemExit("Zipf sort failed verification")
    # Scans (C2 shapes).
    for elem, sumt, n in ((np.uint32, np.uint32, 1 << 28), (np.uint32, np.uint64, 1 << 28),
                          (np.uint64, np.uint64, 1 << 27)):
        bi = clops.Buffer(ctx, n * np.dtype(elem).itemsize)
        bo = clops.Buffer(ctx, n * np.dtype(sumt).itemsize)
        clops.generate(q, bi, n, clops.GEN_U32_MASK if elem == np.uint32 else clops.GEN_U64,
                       0x2C, 0xFF)
        sc = clops.Scanner(ctx, elem, sumt)
        ms = time_steps(q, lambda: sc.scan_device(q, bi, bo, n), 10, 3)
        by = n * (np.dtype(elem).itemsize + np.dtype(sumt).itemsize)
        kern = {nm: round(q.prof(nm)[1] / 10, 4) for nm in q.prof_names()}
        out.append({"config": "scan %s->%s n=%d" % (np.dtype(elem).name, np.dtype(sumt).name, n),
                    "ms": ms, "GB/s": by / ms / 1e6, "hbm_frac": by / ms / 1e6 / HBM_PEAK_GBS,
                    "kernel_ms_per_scan": kern})
    # RNG fills (C2).
    G, draws = 1 << 26, 16
    ob = clops.Buffer(ctx, G * draws * 4)
    for gen, ss in (("xorshift128", 16), ("mwc64x", 8)):
        # Seeding (clo_rng_new) in both of the reference bench's modes, main
        # seed 0: HOST_MT draws the seed words on the host and copies them in;
        # DEV_GID derives them on the device (one launch).  Wall time to a
        # finished queue.
        init_ms = {}
        for mode, key in ((clops.SEED_HOST_MT, "host_mt"), (clops.SEED_DEV_GID, "dev_gid")):
            q.finish()
            t0 = time.perf_counter()
            r = clops.Rng(ctx, gen, mode, G, 0, q=q)
            q.finish()
            init_ms[key] = round((time.perf_counter() - t0) * 1e3, 3)
        ms = time_steps(q, lambda: r.fill(q, ob, draws), 10, 3)
        by = G * (2 * ss + 4 * draws)
        out.append({"config": "rng fill %s 2^26x16" % gen, "ms": ms, "GB/s": by / ms / 1e6,
                    "hbm_frac": by / ms / 1e6 / HBM_PEAK_GBS, "seed_init_ms": init_ms})
    # Bitonic 2^24 u32.
    n = 1 << 24
    kin = clops.Buffer(ctx, n * 4)
    kout = clops.Buffer(ctx, n * 4)
    clops.generate(q, kin, n, clops.GEN_U32_HI, 0xC3D)
    for algo in ("abitonic", "sbitonic"):
        s = clops.Sorter(ctx, algo, np.uint32)
        ms = time_steps(q, lambda: s.sort_device(q, kin, kout, n), 5, 2)
        sw = s.last_sweeps()
        kern = {nm: round(q.prof(nm)[1] / 5, 4) for nm in q.prof_names()}
        out.append({"config": "%s u32 2^24" % algo, "ms": ms, "Gkeys/s": n / ms / 1e6,
                    "sweeps": sw, "hbm_frac": sw * 2 * n * 4 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                    "kernel_ms_per_sort": kern})
    out.extend(suite_segmented(ctx, q))
    out.extend(suite_snippets(ctx, q))
    if cpu_threads:
        cpu = cpu_suite(cpu_threads)
        for o in out:
            for key, rec in cpu.items():
                if o["config"].startswith(key):
                    o["cpu_baseline"] = rec
    for o in out:
        log(json.dumps(o))


FLOAT_KEY = "(as_uint(x) ^ ((uint)(as_int(x) >> 31) | 0x80000000u))"


def suite_snippets(ctx, q):
    """Sorters compiled at run time from compare / key snippets, each timed
    beside the precompiled sorter of the same size and element width."""
    out = []
    t0 = time.perf_counter()
    clops.Sorter(ctx, "sbitonic", np.uint64, key_dtype=np.uint32,
                 get_key="((uint)((x) >> SHIFT))", compiler_opts="-DSHIFT=%d" % (int(t0) % 7 + 30))
    compile_s = time.perf_counter() - t0
    cases = [
        # (label, n, elem width, snippet sorter, precompiled sorter)
        ("satradix float32 2^28 via key snippet", 1 << 28, 4,
         clops.Sorter(ctx, "satradix", np.float32, key_dtype=np.uint32, get_key=FLOAT_KEY),
         clops.Sorter(ctx, "satradix", np.uint32)),
        ("abitonic u64 2^24 keyed by high 32 bits", 1 << 24, 8,
         clops.Sorter(ctx, "abitonic", np.uint64, key_dtype=np.uint32,
                      get_key="((uint)((x) >> 32))"),
         clops.Sorter(ctx, "abitonic", np.uint64)),
        ("gselect float32 2^14", 1 << 14, 4, clops.Sorter(ctx, "gselect", np.float32),
         clops.Sorter(ctx, "gselect", np.uint32)),
    ]
    for label, n, w, snip, pre in cases:
        kin = clops.Buffer(ctx, n * w)
        kout = clops.Buffer(ctx, n * w)
        # Finite floats of both signs for the float sorters: uniform bits with
        # the lowest exponent bit cleared, so the exponent is never 0xFF.
        clops.generate(q, kin, n, clops.GEN_U32_MASK if w == 4 else clops.GEN_U64, 0x5A,
                       0xFF7FFFFF)
        reps = 3 if n > (1 << 20) else 10
        ms = time_steps(q, lambda: snip.sort_device(q, kin, kout, n), reps, 1)
        kern = {nm: round(q.prof(nm)[1] / reps, 4) for nm in q.prof_names()}
        ms_pre = time_steps(q, lambda: pre.sort_device(q, kin, kout, n), reps, 1)
        out.append({"config": "snippet " + label, "ms": ms, "Gkeys/s": n / ms / 1e6,
                    "precompiled_ms": ms_pre, "slowdown": ms / ms_pre,
                    "kernel_ms_per_sort": kern, "first_compile_s": compile_s})
        kin.close()
        kout.close()
    return out


def _best_time(fn, reps=2):
    best = None
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t = time.perf_counter() - t0
        best = t if best is None else min(best, t)
    return best


def cpu_suite(threads):
    """The oracle's reference-structured CPU restatements, timed on bounded
    samples of each suite configuration (rates in the GPU line's units)."""
    sys.path[:0] = [os.path.join(ROOT, "oracle")]
    import oracle
    res = {}
    n = 1 << 20
    k = oracle.generate(oracle.GEN_U32_HI, 0xC1, n)
    t = _best_time(lambda: oracle.satradix(k, bits=4, lws=256, threads=threads))
    res["satradix u32 2^20"] = {"value": n / t / 1e9, "unit": "Gkeys/s", "cores": threads,
                                "kind": "port", "sample": "the full 2^20 keys, 4-bit digits, "
                                "256-element blocks, %.3f s" % t}
    m = 1 << 24
    a = oracle.generate(oracle.GEN_U32_MASK, 0x2C, m, 0xFF)
    for sumt, key in ((np.uint32, "scan uint32->uint32"), (np.uint64, "scan uint32->uint64")):
        t = _best_time(lambda: oracle.scan_blelloch(a, sumt, lws=256, threads=threads))
        by = m * (4 + np.dtype(sumt).itemsize)
        res[key] = {"value": by / t / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
                    "sample": "2^24 elements, 3-phase Blelloch with 2*256-element blocks, %.3f s" % t}
    G, draws = 1 << 20, 16
    for gen, ss in (("xorshift128", 16), ("mwc64x", 8)):
        st = oracle.rng_init_dev_gid(gen, G, 0, 0)
        t = _best_time(lambda: oracle.rng_fill(gen, st.copy(), draws, threads=threads))
        res["rng fill " + gen] = {"value": G * (2 * ss + 4 * draws) / t / 1e9, "unit": "GB/s",
                                  "cores": threads, "kind": "port",
                                  "sample": "2^20 work-items x 16 draws, %.3f s" % t}
    b = oracle.generate(oracle.GEN_U32_HI, 0xC3D, 1 << 20)
    t = _best_time(lambda: oracle.sbitonic(b), 1)
    res["sbitonic u32"] = res["abitonic u32"] = {
        "value": (1 << 20) / t / 1e9, "unit": "Gkeys/s", "cores": 1, "kind": "port",
        "sample": "2^20 keys through the full network, one thread, %.2f s" % t}
    lens = np.clip(np.floor(np.exp(np.random.default_rng(0xC4).uniform(0, np.log(65536), 512))),
                   1, 65536).astype(np.int64)
    keys = oracle.generate(oracle.GEN_U32_HI, 0xC4B, int(lens.sum()))
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]])

    def segs():
        for o, ln in zip(offs, lens):
            oracle.stable_sort(keys[o:o + ln])

    t = _best_time(segs, 1)
    res["segmented sort u32"] = {"value": int(lens.sum()) / t / 1e9, "unit": "Gkeys/s",
                                 "cores": 1, "kind": "port",
                                 "sample": "512 segments of the same length distribution "
                                           "(%d keys), stable LSD per segment, %.2f s"
                                           % (int(lens.sum()), t)}
    return res


def c4_offsets(ctx, q, nseg=32768, max_len=65536, seed=0xC4A):
    """Segment lengths floor(exp(u ln 65536)) clamped to [1, 65536], u from the
    seeded generator (U(seed, j) >> 11) * 2^-53; offsets = exclusive sum."""
    ub = clops.Buffer(ctx, nseg * 8)
    clops.generate(q, ub, nseg, clops.GEN_U64, seed)
    u = ub.read(q, np.uint64, nseg)
    x = (u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    lens = np.clip(np.floor(np.exp(x * np.log(max_len))), 1, max_len).astype(np.uint32)
    off = np.zeros(nseg, dtype=np.uint32)
    off[1:] = np.cumsum(lens[:-1], dtype=np.uint64).astype(np.uint32)
    return lens, off


def suite_segmented(ctx, q, baseline_stride=16):
    """Batched segmented sort and scan (32768 segments, log-uniform lengths),
    eager and as one captured graph, against one library call per segment
    (timed on every baseline_stride-th segment and scaled by the segment
    count)."""
    res = []
    lens, off = c4_offsets(ctx, q)
    nseg = lens.size
    n = int(lens.sum(dtype=np.uint64))
    kin = clops.Buffer(ctx, n * 4)
    kout = clops.Buffer(ctx, n * 4)
    sout = clops.Buffer(ctx, n * 4)
    clops.generate(q, kin, n, clops.GEN_U32_HI, 0xC4B)
    offb = clops.Buffer.from_array(ctx, off)
    s = clops.Sorter(ctx, "satradix", np.uint32)
    sc = clops.Scanner(ctx, np.uint32, np.uint32)
    ms_sort = time_steps(q, lambda: s.sort_segmented_device(q, kin, offb, nseg, n, kout=kout), 5, 2)
    ko = kout.read(q, np.uint32, n)
    inner = np.ones(n - 1, dtype=bool)
    inner[off[1:] - 1] = False  # pairs that straddle a segment start
    bad = int(np.count_nonzero((ko[1:] < ko[:-1]) & inner))
    del ko
    ms_scan = time_steps(q, lambda: sc.scan_segmented_device(q, kout, sout, offb, nseg, n), 10, 3)

    def both():
        s.sort_segmented_device(q, kin, offb, nseg, n, kout=kout)
        sc.scan_segmented_device(q, kout, sout, offb, nseg, n)

    both()
    q.finish()
    g = clops.Graph(q)
    g.begin()
    try:
        both()
    finally:
        g.end()
    ms_graph = time_steps(q, g.launch, 10, 3)
    by = n * 8
    res.append({"config": "segmented sort u32, %d segments, %d keys" % (nseg, n), "ms": ms_sort,
                "Gkeys/s": n / ms_sort / 1e6, "hbm_frac": by / ms_sort / 1e6 / HBM_PEAK_GBS,
                "unsorted_pairs": bad})
    res.append({"config": "segmented scan u32->u32, same offsets", "ms": ms_scan,
                "hbm_frac": by / ms_scan / 1e6 / HBM_PEAK_GBS})
    res.append({"config": "segmented sort + scan as one graph (%d nodes)" % g.num_nodes(),
                "ms": ms_graph, "eager_ms": ms_sort + ms_scan})
    # Baseline: one sort call and one scan call per segment (sampled).
    idx = np.arange(0, nseg, baseline_stride)
    views = [(kin.view(int(off[j]) * 4, int(lens[j]) * 4), kout.view(int(off[j]) * 4,
              int(lens[j]) * 4), sout.view(int(off[j]) * 4, int(lens[j]) * 4), int(lens[j]))
             for j in idx]

    def per_segment():
        for a, b, c, m in views:
            s.sort_device(q, a, b, m)
            sc.scan_device(q, b, c, m)

    ms_base = time_steps(q, per_segment, 2, 1) * nseg / idx.size
    res.append({"config": "baseline: one sort + one scan call per segment (every %dth segment "
                          "timed, scaled by %d/%d)" % (baseline_stride, nseg, idx.size),
                "ms": ms_base, "speedup_graph": ms_base / ms_graph})
    g.close()
    return res


M64 = (1 << 64) - 1
DISTS = ("uniform", "16 unique", "presorted+1%", "reversed", "zipf1.1")


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def _U(seed, i):
    return _mix64((seed + (i + 1) * 0x9E3779B97F4A7C15) & M64)


def request_stream(seed, count, max_log2=28):
    """Request j of the mixed stream: size floor(2^(10 + 18 u)) with
    u = (U(seed, 4j) >> 11) 2^-53 (capped at 2^max_log2), key width 4 or 8,
    sort or scan, one of five distributions; all from U(seed, 4j..4j+3)."""
    out = []
    for j in range(count):
        u = (_U(seed, 4 * j) >> 11) * 2.0 ** -53
        n = min(int(2.0 ** (10 + 18 * u)), 1 << max_log2)
        width = 8 if _U(seed, 4 * j + 1) & 1 else 4
        op = clops.OP_SCAN if _U(seed, 4 * j + 2) & 1 else clops.OP_SORT
        dist = _U(seed, 4 * j + 3) % 5
        out.append((op, n, width, dist, _U(seed, 4 * j + 3) >> 8))
    return out


class RequestData:
    """Device buffers big enough for any request, filled per request."""

    def __init__(self, ctx, q, max_n):
        self.q = q
        self.inb = clops.Buffer(ctx, max_n * 8)
        self.outb = clops.Buffer(ctx, max_n * 8)
        self.out2 = clops.Buffer(ctx, max_n * 8)
        import ctypes
        cdfh = np.empty(1 << 24, dtype=np.float64)
        clops.lib().clo_util_zipf_cdf(cdfh.ctypes.data_as(ctypes.c_void_p), cdfh.size, 1.1)
        self.cdf = clops.Buffer.from_array(ctx, cdfh)
        self.ranks = cdfh.size

    def fill(self, n, width, dist, seed):
        q, b = self.q, self.inb
        w8 = width == 8
        step = max(((1 << (8 * width)) - 1) // max(n, 1), 1)
        if dist == 0:
            clops.generate(q, b, n, clops.GEN_U64 if w8 else clops.GEN_U32_HI, seed)
        elif dist == 1:
            clops.generate(q, b, n, clops.GEN_U64_MOD if w8 else clops.GEN_U32_MOD, seed, 16)
        elif dist == 2:
            clops.generate(q, b, n, clops.GEN_NEARLY_SORTED_U64 if w8 else
                           clops.GEN_NEARLY_SORTED_U32, seed, step)
        elif dist == 3:
            clops.generate(q, b, n, clops.GEN_REVERSED_U64 if w8 else clops.GEN_REVERSED_U32,
                           seed, step)
        else:
            fn = clops.lib().clo_util_generate_zipf if w8 else \
                clops.lib().clo_util_generate_zipf_u32
            clops._call(fn, q.ptr, b.ptr, n, self.cdf.ptr, self.ranks, seed)


def dispatch_eval(ctx, n_train=2700, n_eval=300, train_max_log2=28, steps=8000, lr=2e-3,
                  margin=0.02):
    """Learned dispatch on the mixed request stream.

    Training: n_train requests (seed 0xC50, sizes up to 2^train_max_log2),
    every feasible configuration timed by the exhaustive tuner (min of 2
    runs); the cost model is trained on the GPU on those timings, as log2 of
    each configuration's time over the fixed configuration's time for the
    same request (the choice only needs one request's configurations in
    order).  The model policy leaves the fixed configuration only for a
    predicted gain above `margin` (log2).
    Evaluation: the first n_eval requests of the stream of seed 0xC5 (sizes
    2^10..2^28); each is tuned exhaustively (ground truth), and the model's,
    the tuner table's and the fixed configuration's choices are priced with
    those same timings.  Every chosen output is compared word for word with
    the fixed configuration's output."""
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    data = RequestData(ctx, q, 1 << 28)
    t0 = time.perf_counter()
    X, y = [], []
    for (op, n, width, dist, s) in request_stream(0xC50, n_train, train_max_log2):
        data.fill(n, width, dist, s)
        f = d.features(q, op, data.inb, n, width)
        t = d.tune(q, op, data.inb, data.outb, n, width, reps=2)
        d.table_add(op, n, width, f, t)
        t_fixed = t[clops.Dispatch.fixed(op, n, width)]
        for c in range(clops.N_CONFIGS):
            if t[c] > 0:
                X.append(clops.Dispatch.model_row(f, c))
                y.append(np.log2(t[c] / t_fixed))
    t_collect = time.perf_counter() - t0
    X, y = np.array(X, dtype=np.float32), np.array(y, dtype=np.float32)
    m = clops.Mlp(ctx, clops.MODEL_INPUTS, seed=0xC5)
    t0 = time.perf_counter()
    loss = m.train(q, X, y, steps, lr)
    t_train = time.perf_counter() - t0
    d.set_model(m)
    d.set_margin(margin)
    reqs = request_stream(0xC5, n_eval, 28)
    feats, times = [], []
    for (op, n, width, dist, s) in reqs:
        data.fill(n, width, dist, s)
        feats.append(d.features(q, op, data.inb, n, width))
        times.append(d.tune(q, op, data.inb, data.outb, n, width, reps=3))
    feats, times = np.array(feats), np.array(times)
    ops = [r[0] for r in reqs]
    ns = [r[1] for r in reqs]
    ws = [r[2] for r in reqs]
    t0 = time.perf_counter()
    model = d.choose(q, clops.POLICY_MODEL, ops, ns, ws, feats)
    t_choose = (time.perf_counter() - t0) / len(reqs)
    table = d.choose(q, clops.POLICY_TABLE, ops, ns, ws, feats)
    fixed = d.choose(q, clops.POLICY_FIXED, ops, ns, ws, feats)
    best = np.array([int(np.argmin(np.where(t > 0, t, np.inf))) for t in times])
    rows = np.arange(len(reqs))
    tot = {k: float(times[rows, v].sum()) for k, v in
           (("model", model), ("table", table), ("fixed", fixed), ("best", best))}
    # bit-exact: chosen configuration vs the fixed one, word for word
    diffs = 0
    for r, (op, n, width, dist, s) in enumerate(reqs):
        if model[r] == fixed[r]:
            continue
        data.fill(n, width, dist, s)
        d.run(q, int(model[r]), op, data.inb, data.outb, n, width)
        d.run(q, int(fixed[r]), op, data.inb, data.out2, n, width)
        diffs += clops.count_diff(q, data.outb, data.out2, n * width // 4)
    return {
        "train_requests": n_train, "train_timings": int(y.size),
        "train_collect_s": round(t_collect, 1), "train_s": round(t_train, 2),
        "train_loss": round(loss, 4), "train_target": "log2(time / fixed-config time)",
        "margin_log2": margin, "eval_requests": n_eval,
        "eval_keys": int(sum(ns)),
        "regret_model": round(tot["model"] / tot["best"] - 1, 4),
        "regret_table": round(tot["table"] / tot["best"] - 1, 4),
        "regret_fixed": round(tot["fixed"] / tot["best"] - 1, 4),
        "speedup_model_vs_fixed": round(tot["fixed"] / tot["model"], 4),
        "model_matches_best": round(float(np.mean(model == best)), 3),
        "choose_us_per_request": round(t_choose * 1e6, 1),
        "eval_ms": {k: round(v, 3) for k, v in tot.items()},
        "differing_words_vs_fixed": int(diffs),
        "sample": "first %d requests of the seed-0xC5 stream (sizes 2^10..2^28), exhaustive "
                  "tuner ground truth (min of 3 timed runs per configuration)" % n_eval,
    }


def copy_probe(ctx, nbytes):
    q = clops.Queue(ctx)
    a = clops.Buffer(ctx, nbytes)
    b = clops.Buffer(ctx, nbytes)
    for _ in range(3):
        clops._call(clops.lib().clo_util_copy, q.ptr, b.ptr, a.ptr, nbytes)
    q.finish()
    log("copy probe: 3 x util_copy of %d bytes" % nbytes)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--log2n", type=int, default=28)
    ap.add_argument("--cpu-sample-log2", type=int, default=27)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--suite", action="store_true")
    ap.add_argument("--no-dispatch", action="store_true")
    ap.add_argument("--dispatch-train", type=int, default=2700)
    ap.add_argument("--dispatch-eval", type=int, default=300)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    ctx = clops.Context(0)
    if args.copy_probe:
        copy_probe(ctx, 1 << 31)
        return
    h = headline(args, ctx)
    if args.suite:
        suite(ctx, 0 if args.no_cpu_baseline else (args.cpu_threads or min(os.cpu_count() or 1, 16)))
    disp = None
    if not args.no_dispatch:
        disp = dispatch_eval(ctx, args.dispatch_train, args.dispatch_eval)
        log("dispatch: " + json.dumps(disp))
    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        cpu = cpu_baseline_kv(1 << args.cpu_sample_log2, threads)
    n = h["n"]
    value = n / (h["ms"] * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "satradix stable KV sort, 2^%d pairs, u64 key U(0xC3A,i) uniform over "
                        "64 bits, u32 value = index, out of place" % args.log2n,
            "n": n, "key": "u64", "value": "u32", "digit_bits": 8,
            "passes_executed": h["passes"],
            "whole_sort_hbm_frac": round((n * 8 + h["passes"] * h["bytes_per_launch"]) /
                                         (h["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(h["achieved"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(h["achieved"] / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(),
            "kernel": "satradix_downsweep",
            "kernel_avg_ms": round(h["scatter_ms"], 4),
            "timed_region": "%d steps with HIP events around every satradix_downsweep launch "
                            "(%.4f ms/step; value comes from the same steps without events)"
                            % (args.steps, h["ms_downsweep_events"]),
            "bytes_per_launch": h["bytes_per_launch"],
        },
        "cpu_baseline": cpu,
        "dispatch": disp,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
