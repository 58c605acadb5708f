"""Benchmark of the hot path: stable key-value radix sort (satradix) of 2^28
pairs, uint64 keys uniform over 64 bits (seed 0xC3A), uint32 payload = index.

One step = one clo_sort_kv_with_device_data call, out of place (the unsorted
input stays resident in HBM and is re-sorted every step).  Prints one JSON
line.  `roofline` is for the dominant kernel (satradix_downsweep, one per
digit pass): algorithmic bytes per launch = n * 2 * (8 + 4) divided by its
average launch duration measured with HIP events on the exec stream over the
timed region.  `cpu_baseline` times the oracle's reference-structured satradix
(4-bit digits, 256-element blocks, blocks spread over host threads) on a
bounded sample of the same workload.  `config.copy_ceiling_gbs` is a
streaming copy of the 2 GiB key array timed in the same run; the whole sort
and the downsweep are also given as fractions of it.

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
    # Measured copy ceiling of this GPU: util_stream_copy of the 2 GiB key
    # array into the (already verified) output buffer, 3 untimed + 10 timed.
    nb = n * 8 // 16384 * 16384
    L = clops.lib()

    def copy():
        clops._call(L.clo_util_stream_copy, q_plain.ptr, kout.ptr, kin.ptr, nb)

    copy_gbs = None
    if nb:
        copy_ms = time_steps(q_plain, copy, 10, 3)
        copy_gbs = 2 * nb / (copy_ms * 1e-3) / 1e9
        log("stream copy %d bytes: %.4f ms, %.1f GB/s" % (nb, copy_ms, copy_gbs))
    return {
        "ms": ms, "ms_downsweep_events": ms_ds, "n": n, "passes": passes,
        "scatter_ms": scatter_ms, "achieved": achieved, "bytes_per_launch": bytes_per_launch,
        "copy_gbs": copy_gbs,
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
        # The first fill of fresh DEV_GID seeds computes the states instead of
        # reading them (the seeding is deferred into it): states written once.
        r2 = clops.Rng(ctx, gen, clops.SEED_DEV_GID, G, 0, q=q)
        q.finish()
        e0 = q.marker()
        r2.fill(q, ob, draws)
        e1 = q.marker()
        q.finish()
        first_ms = q.elapsed_ms(e0, e1)
        by_first = G * (ss + 4 * draws)
        out.append({"config": "rng fill %s 2^26x16" % gen, "ms": ms, "GB/s": by / ms / 1e6,
                    "hbm_frac": by / ms / 1e6 / HBM_PEAK_GBS, "seed_init_ms": init_ms,
                    "dev_gid_first_fill_ms": round(first_ms, 4),
                    "dev_gid_first_fill_hbm_frac": round(by_first / first_ms / 1e6 /
                                                         HBM_PEAK_GBS, 4)})
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
    res.extend(segmented_tuning(ctx, q, kin, kout, sout, offb, nseg, n))
    return res


def segmented_only(ctx, width, steps, warmup, options=None):
    """The configuration-4 segmented sort alone (keys of `width` bytes, out of
    place, library defaults or the given sorter options), timed and checked
    for per-segment order."""
    q = clops.Queue(ctx)
    lens, off = c4_offsets(ctx, q)
    nseg = lens.size
    n = int(lens.sum(dtype=np.uint64))
    dt = np.uint32 if width == 4 else np.uint64
    kin = clops.Buffer(ctx, n * width)
    kout = clops.Buffer(ctx, n * width)
    clops.generate(q, kin, n, clops.GEN_U32_HI if width == 4 else clops.GEN_U64, 0xC4B)
    offb = clops.Buffer.from_array(ctx, off)
    s = clops.Sorter(ctx, "satradix", dt, options=options)
    ms = time_steps(q, lambda: s.sort_segmented_device(q, kin, offb, nseg, n, kout=kout),
                    steps, warmup)
    ko = kout.read(q, dt, n)
    inner = np.ones(n - 1, dtype=bool)
    inner[off[1:] - 1] = False
    bad = int(np.count_nonzero((ko[1:] < ko[:-1]) & inner))
    return {"config": "segmented sort, %d segments, %d keys of %d bytes%s" % (
                nseg, n, width, ", options " + options if options else ""),
            "ms": ms, "Gkeys/s": n / ms / 1e6, "unsorted_pairs": bad,
            "multiset_equal": clops.multiset_hash(q, kin, n, width) ==
            clops.multiset_hash(q, kout, n, width)}


def segmented_summary(ctx, steps=10, warmup=3):
    """The 32768-segment workload (log-uniform lengths up to 65536, 1.94e8
    u32 keys) as the default run reports it: the keys-alone segmented sort
    (out of place, library defaults, every segment checked for order and the
    multiset kept) and the segmented scan over the same offsets."""
    sort = segmented_only(ctx, 4, steps, warmup)
    q = clops.Queue(ctx)
    lens, off = c4_offsets(ctx, q)
    nseg, n = lens.size, int(lens.sum(dtype=np.uint64))
    kin = clops.Buffer(ctx, n * 4)
    sout = clops.Buffer(ctx, n * 4)
    clops.generate(q, kin, n, clops.GEN_U32_HI, 0xC4B)
    offb = clops.Buffer.from_array(ctx, off)
    sc = clops.Scanner(ctx, np.uint32, np.uint32)
    ms_scan = time_steps(q, lambda: sc.scan_segmented_device(q, kin, sout, offb, nseg, n),
                         steps, warmup)
    return {"workload": "32768 segments, log-uniform lengths in [1, 65536] from U(0xC4A, j), "
                        "%d u32 keys U(0xC4B, i) >> 32" % n,
            "sort_ms": round(sort["ms"], 4), "sort_gkeys_per_s": round(sort["Gkeys/s"], 2),
            "sort_unsorted_pairs": sort["unsorted_pairs"],
            "sort_multiset_equal": sort["multiset_equal"],
            "sort_hbm_frac": round(n * 8 / sort["ms"] / 1e6 / HBM_PEAK_GBS, 4),
            "scan_ms": round(ms_scan, 4),
            "scan_hbm_frac": round(n * 8 / ms_scan / 1e6 / HBM_PEAK_GBS, 4)}


def seg_tile(cfg):
    """Items per thread of segmented configuration cfg (12 + 4 t + g - 2)."""
    return (8, 16, 32)[(cfg - 12) // 4]


def seg_tune(ctx):
    """The segmented tuner and the learned work classes on the 32768
    log-uniform segments, for 32- and 64-bit keys."""
    q = clops.Queue(ctx)
    lens, off = c4_offsets(ctx, q)
    nseg = lens.size
    n = int(lens.sum(dtype=np.uint64))
    offb = clops.Buffer.from_array(ctx, off)
    res = []
    for width in (4, 8):
        bufs = [clops.Buffer(ctx, n * width) for _ in range(3)]
        clops.generate(q, bufs[0], n, clops.GEN_U32_HI if width == 4 else clops.GEN_U64, 0xC4B)
        res.extend(segmented_tuning(ctx, q, *bufs, offb, nseg, n, width))
        for b in bufs:
            b.close()
    return res


def segmented_tuning(ctx, q, kin, kout, kout2, offb, nseg, n, width=4):
    """The exhaustive tuner over the segmented tile x grid configurations on
    this workload (the library's defaults are its choice), then the learned
    per-segment work classes: the segment cost model is trained on work-class
    timings (every length bin up to the tile's CAP, each feasible class) and
    its class table is timed against the length thresholds alone, outputs
    compared word for word."""
    d = clops.Dispatch(ctx)
    t = d.tune_segmented(q, kin, kout, offb, nseg, n, width, reps=3)
    fixed = clops.Dispatch.fixed(clops.OP_SEGSORT, n, width)
    best = int(np.argmin(np.where(t > 0, t, np.inf)))
    res = [{"config": "segmented tuner over tile x grid (%d keys of %d bytes)" % (n, width),
            "times_ms": {d.config_name(c): round(float(t[c]), 4) for c in range(clops.N_CONFIGS)
                         if t[c] > 0},
            "tuner_best": d.config_name(best), "default": d.config_name(fixed),
            "default_vs_best": round(float(t[fixed] / t[best]), 4)}]
    X, y = [], []
    for b in range(clops.SEG_BINS):
        L = clops.Dispatch.seg_bin_length(b)
        if L == 0 or L > 256 * seg_tile(fixed):
            continue
        tc = d.tune_segment_class(q, fixed, width, L, reps=2)
        for c in range(3):
            if tc[c] > 0:
                X.append(clops.Dispatch.segment_row(L, width, fixed, c))
                y.append(np.log2(tc[c]))
    m = clops.Mlp(ctx, clops.SEG_INPUTS, seed=0x5E)
    loss = m.train(q, np.array(X, dtype=np.float32), np.array(y, dtype=np.float32), 3000, 3e-3)
    d.set_segment_model(m)
    classes = d.segment_classes(q, fixed, width, 0.05)
    # Interleaved, least of 7 each: a run measured first would otherwise
    # carry the warm-up of both.
    run_f = lambda: d.run_segmented(q, fixed, kin, kout, offb, nseg, n, width)
    run_c = lambda: d.run_segmented(q, fixed, kin, kout2, offb, nseg, n, width, classes=classes)
    run_f()
    run_c()
    ms_fixed = ms_cls = float("inf")
    for _ in range(7):
        ms_fixed = min(ms_fixed, time_steps(q, run_f, 1, 0))
        ms_cls = min(ms_cls, time_steps(q, run_c, 1, 0))
    diff = clops.count_diff(q, kout, kout2, n * width // 4)
    res.append({"config": "learned per-segment work classes (%s)" % d.config_name(fixed),
                "classes": "".join(map(str, classes)), "train_rows": len(y),
                "train_loss": round(loss, 4), "ms_thresholds": ms_fixed, "ms_learned": ms_cls,
                "differing_words": int(diff)})
    q.gc()
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

    def fill(self, n, width, dist, seed, buf=None, q=None):
        q = q or self.q
        b = buf or self.inb
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


def _collect(d, q, data, reqs, reps):
    """Features and exhaustive tuner timings (min of reps timed runs) of
    every request."""
    feats, times = [], []
    for (op, n, width, dist, s) in reqs:
        data.fill(n, width, dist, s)
        feats.append(d.features(q, op, data.inb, n, width))
        times.append(d.tune(q, op, data.inb, data.outb, n, width, reps=reps))
    q.gc()
    return np.array(feats), np.array(times)


def _rows(reqs, feats, times):
    """Model rows and targets: log2 of each timed configuration's time over
    the fixed configuration's time for the same request."""
    X, y = [], []
    for (op, n, width, _, _), f, t in zip(reqs, feats, times):
        t_fixed = t[clops.Dispatch.fixed(op, n, width)]
        for c in range(clops.N_CONFIGS):
            if t[c] > 0:
                X.append(clops.Dispatch.model_row(f, c))
                y.append(np.log2(t[c] / t_fixed))
    return np.array(X, dtype=np.float32), np.array(y, dtype=np.float32)


def _policy(ctx, q, reqs, feats, times, steps, lr, margin, seed):
    """A dispatcher whose table and model learn from (reqs, feats, times)."""
    d = clops.Dispatch(ctx)
    for (op, n, width, _, _), f, t in zip(reqs, feats, times):
        d.table_add(op, n, width, f, t)
    X, y = _rows(reqs, feats, times)
    m = clops.Mlp(ctx, clops.MODEL_INPUTS, seed=seed)
    t0 = time.perf_counter()
    loss = m.train(q, X, y, steps, lr)
    d.set_model(m)
    d.set_margin(margin)
    return d, m, loss, int(y.size), time.perf_counter() - t0


def _windows(reqs, pool_bytes, max_count=256):
    """Consecutive requests packed into windows whose inputs fit one pool
    (256-byte aligned offsets)."""
    out, cur, used = [], [], 0
    for j, r in enumerate(reqs):
        nb = (r[1] * r[2] + 255) // 256 * 256
        if cur and (used + nb > pool_bytes or len(cur) == max_count):
            out.append(cur)
            cur, used = [], 0
        cur.append((j, used))
        used += nb
    if cur:
        out.append(cur)
    return out


def run_stream(d, data, reqs, pools, windows, mode, q_exec, q_side, choices=None):
    """Runs the request stream through two pools of device memory: while the
    exec queue sorts or scans window w from one pool, the side queue fills
    window w+1 into the other and, under the model policy, computes its
    features (one batched launch) and the model's choices (one inference
    launch).  The exec queue waits for a window's fill with an event, the side
    queue for the exec queue to be done with a pool before refilling it.
    mode: "fixed", "model", or "verify" (runs choices[j] and the fixed
    configuration where they differ and counts differing words).  Returns
    (host wall s, exec-queue ms from first to last work without the drains
    that release events every 64 windows, choices, diffs)."""
    chosen = np.zeros(len(reqs), dtype=np.int32)
    done = [None, None]
    diffs = 0

    def prepare(w):
        slot = w % 2
        if done[slot] is not None:
            q_side.wait_for(done[slot])
        views, ops, ns, ws = [], [], [], []
        for j, offset in windows[w]:
            op, n, width, dist, seed = reqs[j]
            v = pools[slot].view(offset, n * width)
            data.fill(n, width, dist, seed, buf=v, q=q_side)
            views.append(v)
            ops.append(op)
            ns.append(n)
            ws.append(width)
        if mode == "model":
            f = d.features_batch(q_side, ops, views, ns, ws)
            cfgs = d.choose(q_side, clops.POLICY_MODEL, ops, ns, ws, f)
        elif mode == "fixed":
            cfgs = [clops.Dispatch.fixed(o, n, w_) for o, n, w_ in zip(ops, ns, ws)]
        else:
            cfgs = [choices[j] for j, _ in windows[w]]
        return views, cfgs, q_side.marker()

    q_exec.finish()
    q_side.finish()
    t0 = time.perf_counter()
    ms = 0.0
    start = q_exec.marker()
    prep = prepare(0)
    for w in range(len(windows)):
        views, cfgs, ready = prep
        q_exec.wait_for(ready)
        for (j, _), v, c in zip(windows[w], views, cfgs):
            op, n, width = reqs[j][:3]
            chosen[j] = c
            if mode == "verify":
                fx = clops.Dispatch.fixed(op, n, width)
                if c != fx:
                    d.run(q_exec, int(c), op, v, data.outb, n, width)
                    d.run(q_exec, fx, op, v, data.out2, n, width)
                    diffs += clops.count_diff(q_exec, data.outb, data.out2, n * width // 4)
            else:
                d.run(q_exec, int(c), op, v, data.outb, n, width)
        done[w % 2] = q_exec.marker()
        if w % 64 == 63:
            # Release the queues' completed events.  Waiting first means no
            # event still to be waited on is released: the pools are free.
            # The exec time of the windows so far is taken before, and the
            # drain itself is left out of it.
            q_exec.finish()
            q_side.finish()
            ms += q_exec.elapsed_ms(start, done[w % 2])
            q_exec.gc()
            q_side.gc()
            done = [None, None]
            start = q_exec.marker()
        if w + 1 < len(windows):
            prep = prepare(w + 1)
        if w % 200 == 199:
            log("  %s stream: window %d/%d, %.0f s" % (mode, w + 1, len(windows),
                                                     time.perf_counter() - t0))
    end = q_exec.marker()
    q_exec.finish()
    wall = time.perf_counter() - t0
    ms += q_exec.elapsed_ms(start, end)
    q_exec.gc()
    q_side.gc()
    return wall, ms, chosen, diffs


def dispatch_full(ctx, d, data, count, seed=0xC5, pool_bytes=8 << 30):
    """The whole request stream, no exhaustive tuning: once under the fixed
    configuration and once under the model policy (both timed end to end,
    fills and the model's batched features and inference included, see
    run_stream), then an untimed pass that runs the model's choice and the
    fixed configuration wherever they differ and compares the outputs word
    for word.  The stream's first half runs fixed then model, its second
    half model then fixed, so a drift of the GPU's speed over the run
    cancels."""
    reqs = request_stream(seed, count, 28)
    pools = [clops.Buffer(ctx, pool_bytes) for _ in range(2)]
    windows = _windows(reqs, pool_bytes)
    q_exec, q_side = clops.Queue(ctx), clops.Queue(ctx)
    res = {"fixed": [0.0, 0.0], "model": [0.0, 0.0]}
    choices = np.zeros(len(reqs), dtype=np.int32)
    half = len(windows) // 2
    for part, order in ((windows[:half], ("fixed", "model")), (windows[half:], ("model", "fixed"))):
        for mode in order:
            wall, ms, chosen, _ = run_stream(d, data, reqs, pools, part, mode, q_exec, q_side)
            res[mode][0] += wall
            res[mode][1] += ms
            if mode == "model":
                for w in part:
                    for j, _ in w:
                        choices[j] = chosen[j]
            log("  full stream %s, %d windows: %.1f s wall, %.1f ms exec queue"
                % (mode, len(part), wall, ms))
    fixed = np.array([clops.Dispatch.fixed(r[0], r[1], r[2]) for r in reqs])
    wall_v, _, _, diffs = run_stream(d, data, reqs, pools, windows, "verify", q_exec, q_side,
                                     choices=choices)
    for p in pools:
        p.close()
    return {
        "full_stream_requests": count,
        "full_stream_keys": int(sum(r[1] for r in reqs)),
        "full_stream_windows": len(windows),
        "full_stream_wall_s": {"fixed": round(res["fixed"][0], 2),
                               "model": round(res["model"][0], 2)},
        "full_stream_exec_ms": {"fixed": round(res["fixed"][1], 1),
                                "model": round(res["model"][1], 1)},
        "full_stream_speedup_vs_fixed": round(res["fixed"][1] / res["model"][1], 4),
        "full_stream_speedup_vs_fixed_wall": round(res["fixed"][0] / res["model"][0], 4),
        "full_stream_model_changed": int(np.count_nonzero(choices != fixed)),
        "full_stream_differing_words_vs_fixed": int(diffs),
        "full_stream_verify_s": round(wall_v, 1),
        "full_stream_timing": "end to end per policy: windows of up to 256 requests in two "
                              "%d GiB pools; fills and the model policy's batched features "
                              "and inference run on a side queue overlapped with the exec "
                              "queue's sorts and scans; exec ms from the first to the last "
                              "work on the exec queue; first half fixed then model, second "
                              "half model then fixed" % (pool_bytes >> 30),
    }


def _warm_configs(d, q, data):
    """Every configuration once at the largest request size of each width,
    so no timed run later grows a sorter's scratch."""
    for width in (4, 8):
        n = 1 << 28
        data.fill(n, width, 0, 1)
        for c in range(clops.N_CONFIGS):
            for op in (clops.OP_SORT, clops.OP_SCAN):
                if d.feasible(c, op, n, width):
                    d.run(q, c, op, data.inb, data.outb, n, width)
                    d.run(q, c, op, data.inb, data.out2, n, width)
        q.finish()
        q.gc()


def dispatch_eval(ctx, n_train=2000, n_eval=1500, n_full=100000, n_sparse=200,
                  train_max_log2=28, steps=8000, lr=2e-3, margin=0.02):
    """Learned dispatch on the mixed request stream.

    Training: n_train requests (seed 0xC50, sizes up to 2^train_max_log2),
    every feasible configuration timed by the exhaustive tuner (min of 2
    runs); the cost model is trained on the GPU on those timings, as log2 of
    each configuration's time over the fixed configuration's time for the
    same request (the choice only needs one request's configurations in
    order), and the tuner table keeps their per-bucket means.  The model
    policy leaves the fixed configuration only for a predicted gain above
    `margin` (log2).
    Regret: the first n_eval requests of the stream of seed 0xC5 (sizes
    2^10..2^28), each tuned exhaustively (min of 3 timed runs: the ground
    truth); the model's, the table's and the fixed configuration's choices
    are priced with those timings.  Every chosen output is compared word for
    word with the fixed configuration's output.
    Sparse ablation: a table and a model that learn from the first n_sparse
    training requests only (most buckets then have no entry, and the table
    falls back to the fixed configuration there), priced the same way.
    Full stream: all n_full requests of the seed-0xC5 stream run under the
    fixed configuration and under the model policy (dispatch_full)."""
    q = clops.Queue(ctx)
    data = RequestData(ctx, q, 1 << 28)
    t0 = time.perf_counter()
    train = request_stream(0xC50, n_train, train_max_log2)
    tf, tt = _collect(clops.Dispatch(ctx), q, data, train, 2)
    t_collect = time.perf_counter() - t0
    d, m, loss, n_rows, t_train = _policy(ctx, q, train, tf, tt, steps, lr, margin, 0xC5)
    ds, ms_, loss_s, n_rows_s, _ = _policy(ctx, q, train[:n_sparse], tf[:n_sparse],
                                           tt[:n_sparse], steps, lr, margin, 0xC6)
    log("dispatch: trained on %d timings (%.1f s collect, %.2f s train), sparse on %d"
        % (n_rows, t_collect, t_train, n_rows_s))
    reqs = request_stream(0xC5, n_eval, 28)
    t0 = time.perf_counter()
    feats, times = _collect(d, q, data, reqs, 3)
    t_eval = time.perf_counter() - t0
    ops = [r[0] for r in reqs]
    ns = [r[1] for r in reqs]
    ws = [r[2] for r in reqs]
    t0 = time.perf_counter()
    model = d.choose(q, clops.POLICY_MODEL, ops, ns, ws, feats)
    t_choose = (time.perf_counter() - t0) / len(reqs)
    table = d.choose(q, clops.POLICY_TABLE, ops, ns, ws, feats)
    fixed = d.choose(q, clops.POLICY_FIXED, ops, ns, ws, feats)
    model_s = ds.choose(q, clops.POLICY_MODEL, ops, ns, ws, feats)
    table_s = ds.choose(q, clops.POLICY_TABLE, ops, ns, ws, feats)
    best = np.array([int(np.argmin(np.where(t > 0, t, np.inf))) for t in times])
    rows = np.arange(len(reqs))
    tot = {k: float(times[rows, v].sum()) for k, v in
           (("model", model), ("table", table), ("fixed", fixed), ("best", best),
            ("sparse_model", model_s), ("sparse_table", table_s))}
    hit = np.array([ds.table_choose(op, n, w, f) >= 0 for op, n, w, f in zip(ops, ns, ws, feats)])
    sparse_hits = int(hit.sum())
    # Held-out buckets: the evaluation requests whose bucket the sparse table
    # never saw (the table gives the fixed configuration there).
    ho = ~hit
    ho_tot = {k: float(times[rows[ho], v[ho]].sum()) for k, v in
              (("table", table_s), ("model", model_s), ("best", best))}
    # bit-exact: chosen configuration vs the fixed one, word for word
    diffs = 0
    for r, (op, n, width, dist, s) in enumerate(reqs):
        if model[r] == fixed[r]:
            continue
        data.fill(n, width, dist, s)
        d.run(q, int(model[r]), op, data.inb, data.outb, n, width)
        d.run(q, int(fixed[r]), op, data.inb, data.out2, n, width)
        diffs += clops.count_diff(q, data.outb, data.out2, n * width // 4)
    q.gc()
    out = {
        "train_requests": n_train, "train_timings": n_rows,
        "train_collect_s": round(t_collect, 1), "train_s": round(t_train, 2),
        "train_loss": round(loss, 4), "train_target": "log2(time / fixed-config time)",
        "margin_log2": margin, "eval_requests": n_eval, "eval_tune_s": round(t_eval, 1),
        "eval_keys": int(sum(ns)),
        "regret_model": round(tot["model"] / tot["best"] - 1, 4),
        "regret_table": round(tot["table"] / tot["best"] - 1, 4),
        "regret_fixed": round(tot["fixed"] / tot["best"] - 1, 4),
        "speedup_model_vs_fixed": round(tot["fixed"] / tot["model"], 4),
        "model_matches_best": round(float(np.mean(model == best)), 3),
        "regret_sparse_table": round(tot["sparse_table"] / tot["best"] - 1, 4),
        "regret_sparse_model": round(tot["sparse_model"] / tot["best"] - 1, 4),
        "regret_heldout_buckets": {
            "requests": int(ho.sum()),
            "table": round(ho_tot["table"] / ho_tot["best"] - 1, 4) if ho.any() else None,
            "model": round(ho_tot["model"] / ho_tot["best"] - 1, 4) if ho.any() else None},
        "sparse": "table and model from the first %d training requests (%d timings); the sparse "
                  "table has an entry for %d of the %d evaluation requests' buckets"
                  % (n_sparse, n_rows_s, sparse_hits, n_eval),
        "choose_us_per_request": round(t_choose * 1e6, 1),
        "eval_ms": {k: round(v, 3) for k, v in tot.items()},
        "differing_words_vs_fixed": int(diffs),
        "sample": "first %d requests of the seed-0xC5 stream (sizes 2^10..2^28), exhaustive "
                  "tuner ground truth (min of 3 timed runs per configuration)" % n_eval,
    }
    if n_full:
        _warm_configs(d, q, data)
        out.update(dispatch_full(ctx, d, data, n_full))
    return out


def copy_probe(ctx, nbytes):
    """Bandwidth ceilings: util_copy (the traffic counters' calibration
    kernel), util_stream_copy (16 KiB tiles, four 16-byte loads in flight per
    thread, XCD-contiguous) over `nbytes`, and util_scatter_runs, the write
    pattern of one uniform radix pass over 2^28 (u64, u32) pairs in runs of
    4 to 128 pairs (32: 8-bit digits on 8192-pair tiles; 4: 11-bit digits).  Each: 3 untimed launches, then 10 between two
    markers; one JSON line per probe on stdout."""
    q = clops.Queue(ctx)
    L = clops.lib()
    a = clops.Buffer(ctx, nbytes)
    b = clops.Buffer(ctx, nbytes)
    clops.generate(q, a, nbytes // 8, clops.GEN_U64, 1)
    res = []

    def timed(fn, name, nb):
        for _ in range(3):
            fn()
        q.finish()
        e0 = q.marker()
        for _ in range(10):
            fn()
        e1 = q.marker()
        q.finish()
        ms = q.elapsed_ms(e0, e1) / 10
        res.append({"probe": name, "ms": round(ms, 4), "bytes": nb,
                    "GB/s": round(nb / ms / 1e6, 1), "frac_of_8TBs": round(nb / ms / 1e6 / 8000, 4)})

    timed(lambda: clops._call(L.clo_util_copy, q.ptr, b.ptr, a.ptr, nbytes),
          "util_copy %d bytes" % nbytes, 2 * nbytes)
    timed(lambda: clops._call(L.clo_util_stream_copy, q.ptr, b.ptr, a.ptr, nbytes),
          "util_stream_copy %d bytes" % nbytes, 2 * nbytes)
    a.close()
    b.close()
    n = 1 << 28
    ks, vs = clops.Buffer(ctx, n * 8), clops.Buffer(ctx, n * 4)
    kd, vd = clops.Buffer(ctx, n * 8), clops.Buffer(ctx, n * 4)
    clops.generate(q, ks, n, clops.GEN_U64, 2)
    clops.generate(q, vs, n, clops.GEN_IOTA_U32, 0)
    for run in (4, 8, 16, 32, 64, 128):
        timed(lambda: clops._call(L.clo_util_scatter_runs, q.ptr, kd.ptr, vd.ptr, ks.ptr, vs.ptr,
                                  n, run),
              "util_scatter_runs run=%d, 2^28 u64/u32 pairs" % run, n * 24)
    # The scatter is a permutation: every pair lands once.
    assert clops.multiset_hash(q, kd, n, 8) == clops.multiset_hash(q, ks, n, 8)
    for r in res:
        print(json.dumps(r), flush=True)


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
    ap.add_argument("--seg-tune", action="store_true",
                    help="only the segmented tuner and learned work classes, 32768 segments")
    ap.add_argument("--segmented", type=int, default=0, metavar="KEY_BYTES",
                    help="only the configuration-4 segmented sort, keys alone of 4 or 8 bytes "
                         "(for profilers)")
    ap.add_argument("--seg-options", default=None,
                    help="sorter options for --segmented (e.g. segmsd=0)")
    ap.add_argument("--no-dispatch", action="store_true")
    ap.add_argument("--no-segmented", action="store_true",
                    help="leave the 32768-segment sort and scan out of the default line")
    ap.add_argument("--dispatch-train", type=int, default=2000)
    ap.add_argument("--dispatch-eval", type=int, default=1500)
    ap.add_argument("--dispatch-full", type=int, default=100000,
                    help="requests of the stream run under fixed and model policies (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    ctx = clops.Context(0)
    if args.copy_probe:
        copy_probe(ctx, 1 << 31)
        return
    if args.seg_tune:
        for r in seg_tune(ctx):
            print(json.dumps(r), flush=True)
        return
    if args.segmented:
        print(json.dumps(segmented_only(ctx, args.segmented, args.steps, args.warmup,
                                        args.seg_options)), flush=True)
        return
    h = headline(args, ctx)
    if args.suite:
        suite(ctx, 0 if args.no_cpu_baseline else (args.cpu_threads or min(os.cpu_count() or 1, 16)))
    disp = None
    if not args.no_dispatch:
        disp = dispatch_eval(ctx, args.dispatch_train, args.dispatch_eval, args.dispatch_full)
        log("dispatch: " + json.dumps(disp))
    seg = None
    if not args.no_segmented:
        seg = segmented_summary(ctx)
        log("segmented: " + json.dumps(seg))
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
            "copy_ceiling_gbs": h["copy_gbs"] and round(h["copy_gbs"], 1),
            "whole_sort_of_copy_ceiling": h["copy_gbs"] and round(
                (n * 8 + h["passes"] * h["bytes_per_launch"]) / (h["ms"] * 1e-3) / 1e9 /
                h["copy_gbs"], 4),
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
            "frac_of_copy_ceiling": h["copy_gbs"] and round(h["achieved"] / h["copy_gbs"], 4),
        },
        "cpu_baseline": cpu,
        "dispatch": disp,
        "segmented": seg,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
