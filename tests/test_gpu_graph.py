"""Graph capture (clo_graph_*) of library call sequences on the GPU.

Every replay is checked against the eager result of the same calls.  Between
replays the program does its own eager work (overwriting the outputs, other
kernels, queue synchronisation), so a replay that skipped a node or reused
stale scratch would show up."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lengths(oracle, nseg, max_len, seed):
    u = oracle.generate(oracle.GEN_U64, seed, nseg)
    x = (u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.clip(np.floor(np.exp(x * np.log(max_len))), 1, max_len).astype(np.uint32)


def _offsets(lens):
    off = np.zeros(lens.size, dtype=np.uint32)
    off[1:] = np.cumsum(lens[:-1], dtype=np.uint64).astype(np.uint32)
    return off


def test_graph_replays_match_eager(clops, oracle, ctx):
    q = clops.Queue(ctx)
    # 1) KV radix sort, u64 keys, out of place
    n = 300007
    keys = oracle.generate(oracle.GEN_U64, 0x61, n)
    kin = clops.Buffer.from_array(ctx, keys)
    vin = clops.Buffer.from_array(ctx, np.arange(n, dtype=np.uint32))
    kout, vout = clops.Buffer(ctx, n * 8), clops.Buffer(ctx, n * 4)
    srt = clops.Sorter(ctx, "satradix", np.uint64)
    # 2) segmented sort + segmented scan over the same offsets (u32)
    lens = _lengths(oracle, 700, 20000, 0x62)
    off = _offsets(lens)
    m = int(lens.sum(dtype=np.uint64))
    skeys = oracle.generate(oracle.GEN_U32_HI, 0x63, m)
    sk_in = clops.Buffer.from_array(ctx, skeys)
    sk_out = clops.Buffer(ctx, m * 4)
    offb = clops.Buffer.from_array(ctx, off)
    seg = clops.Sorter(ctx, "satradix", np.uint32)
    sc_out = clops.Buffer(ctx, m * 4)
    scn = clops.Scanner(ctx, np.uint32, np.uint32)

    def work():
        srt.sort_kv_device(q, kin, vin, kout, vout, n)
        seg.sort_segmented_device(q, sk_in, offb, off.size, m, kout=sk_out)
        scn.scan_segmented_device(q, sk_out, sc_out, offb, off.size, m)

    def outputs():
        return (kout.read(q, np.uint64, n), vout.read(q, np.uint32, n),
                sk_out.read(q, np.uint32, m), sc_out.read(q, np.uint32, m))

    work()  # eager: sizes the scratch, gives the expected results
    q.finish()
    want = outputs()
    rk, rv = oracle.stable_sort(keys, np.arange(n, dtype=np.uint32))
    assert np.array_equal(want[0], rk) and np.array_equal(want[1], rv)

    g = clops.Graph(q)
    g.begin()
    try:
        work()
    finally:
        g.end()
    assert g.num_nodes() > 10
    for rep in range(4):
        # eager work between replays: clobber every output
        for b, nb in ((kout, n * 8), (vout, n * 4), (sk_out, m * 4), (sc_out, m * 4)):
            clops.generate(q, b, nb // 4, clops.GEN_U32_HI, 0xDEAD + rep)
        q.finish()
        g.launch()
        q.finish()
        got = outputs()
        for a, b in zip(got, want):
            assert np.array_equal(a, b), rep
    g.close()


def test_graph_replay_follows_new_input(clops, oracle, ctx):
    """The pass plan is made on the device at each replay: a replay over keys
    whose high digits became constant skips those passes and still sorts."""
    q = clops.Queue(ctx)
    n = 200003
    kin = clops.Buffer(ctx, n * 8)
    vin = clops.Buffer.from_array(ctx, np.arange(n, dtype=np.uint32))
    kout, vout = clops.Buffer(ctx, n * 8), clops.Buffer(ctx, n * 4)
    srt = clops.Sorter(ctx, "satradix", np.uint64)
    clops.generate(q, kin, n, clops.GEN_U64, 0x71)
    srt.sort_kv_device(q, kin, vin, kout, vout, n)
    q.finish()
    g = clops.Graph(q)
    g.begin()
    try:
        srt.sort_kv_device(q, kin, vin, kout, vout, n)
    finally:
        g.end()
    for rep, (kind, param) in enumerate([(clops.GEN_U64_MASK, (1 << 20) - 1),
                                         (clops.GEN_U64, 0), (clops.GEN_U64_MASK, 0xFF)]):
        clops.generate(q, kin, n, kind, 0x80 + rep, param)
        q.finish()
        keys = kin.read(q, np.uint64, n)
        g.launch()
        q.finish()
        rk, rv = oracle.stable_sort(keys, np.arange(n, dtype=np.uint32))
        assert np.array_equal(kout.read(q, np.uint64, n), rk)
        assert np.array_equal(vout.read(q, np.uint32, n), rv)
    g.close()


def test_capture_without_warmup_fails_cleanly(clops, oracle, ctx):
    q = clops.Queue(ctx)
    n = 100000
    k = oracle.generate(oracle.GEN_U32_HI, 0x91, n)
    b = clops.Buffer.from_array(ctx, k)
    s = clops.Sorter(ctx, "satradix", np.uint32)  # no eager call: scratch unsized
    g = clops.Graph(q)
    g.begin()
    with pytest.raises(clops.ClopsError) as e:
        s.sort_device(q, b, None, n)
    assert e.value.code == clops.CLO_ERROR_ARGS
    g.end()  # closes the capture (an empty graph)
    g.close()
    s.sort_device(q, b, None, n)  # eager use still works
    q.finish()
    assert np.array_equal(b.read(q, np.uint32, n), np.sort(k))


def test_graph_api_errors(clops, ctx):
    q = clops.Queue(ctx)
    L = clops.lib()
    with pytest.raises(clops.ClopsError):
        clops._call(L.clo_graph_end_capture, q.ptr)  # not capturing
    g = clops.Graph(q)
    g.begin()
    with pytest.raises(clops.ClopsError):
        clops._call(L.clo_graph_begin_capture, q.ptr)  # already capturing
    g.end()
    g.close()
