"""Buffer placement rules at the boundary, graphs that hold library scratch,
malformed segment offsets and the device-side pairing check.

Every error case is refused by the host before anything is enqueued; the
cases that run check their results against the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _kv(clops, oracle, ctx, n, seed):
    k = oracle.generate(oracle.GEN_U64, seed, n)
    v = np.arange(n, dtype=np.uint32)
    return k, v, clops.Buffer.from_array(ctx, k), clops.Buffer.from_array(ctx, v)


@pytest.mark.parametrize("algo", ["satradix", "gselect"])
def test_kv_mixed_placement_refused(clops, oracle, ctx, algo):
    q = clops.Queue(ctx)
    n = 3000
    k, v, kin, vin = _kv(clops, oracle, ctx, n, 0x51)
    kout, vout = clops.Buffer(ctx, n * 8), clops.Buffer(ctx, n * 4)
    s = clops.Sorter(ctx, algo, np.uint64)
    big = clops.Buffer(ctx, n * 16)
    cases = [
        (kout, vin),                   # keys out of place, values in place
        (kin, vout),                   # keys in place, values out of place
        (None, vout),                  # the same through NULL
        (kout, None),
        (big.view(8, n * 8), vout),    # disjoint (a view): allowed
    ]
    for ko, vo in cases[:4]:
        with pytest.raises(clops.ClopsError) as e:
            s.sort_kv_device(q, kin, vin, ko, vo, n)
        assert e.value.code == clops.CLO_ERROR_ARGS, (ko, vo)
    # Partial overlaps: an output shifted inside its input; keys over values.
    kb = clops.Buffer(ctx, n * 8 + 64)
    kb.write(q, k)
    with pytest.raises(clops.ClopsError) as e:
        s.sort_kv_device(q, kb, vin, kb.view(8, n * 8), vout, n)
    assert e.value.code == clops.CLO_ERROR_ARGS
    with pytest.raises(clops.ClopsError) as e:
        s.sort_kv_device(q, kin, vin, kout, kout.view(0, n * 4), n)
    assert e.value.code == clops.CLO_ERROR_ARGS
    q.finish()
    # Inputs untouched by the refused calls.
    assert np.array_equal(kin.read(q, np.uint64, n), k)
    assert np.array_equal(vin.read(q, np.uint32, n), v)
    rk, rv = oracle.stable_sort(k, v)
    # Allowed: both in place through outputs equal to the inputs ...
    s.sort_kv_device(q, kin, vin, kin, vin, n)
    q.finish()
    assert np.array_equal(kin.read(q, np.uint64, n), rk)
    assert np.array_equal(vin.read(q, np.uint32, n), rv)
    # ... and both into disjoint buffers (one of them a view).
    kin.write(q, k)
    vin.write(q, v)
    s.sort_kv_device(q, kin, vin, cases[4][0], vout, n)
    q.finish()
    assert np.array_equal(cases[4][0].read(q, np.uint64, n), rk)
    assert np.array_equal(vout.read(q, np.uint32, n), rv)


def test_segmented_mixed_placement_refused(clops, oracle, ctx):
    q = clops.Queue(ctx)
    n = 5000
    k, v, kin, vin = _kv(clops, oracle, ctx, n, 0x52)
    off = clops.Buffer.from_array(ctx, np.array([0, 1000, 3000], dtype=np.uint32))
    s = clops.Sorter(ctx, "satradix", np.uint64)
    vout = clops.Buffer(ctx, n * 4)
    with pytest.raises(clops.ClopsError) as e:
        s.sort_segmented_device(q, kin, off, 3, n, vin=vin, kout=kin, vout=vout)
    assert e.value.code == clops.CLO_ERROR_ARGS


def test_sort_partial_overlap_refused(clops, oracle, ctx):
    q = clops.Queue(ctx)
    n = 1000
    b = clops.Buffer(ctx, n * 4 + 64)
    b.write(q, oracle.generate(oracle.GEN_U32_HI, 0x53, n))
    for algo in ("satradix", "abitonic", "gselect"):
        s = clops.Sorter(ctx, algo, np.uint32)
        with pytest.raises(clops.ClopsError) as e:
            s.sort_device(q, b, b.view(4, n * 4), n)
        assert e.value.code == clops.CLO_ERROR_ARGS


@pytest.mark.parametrize("n", [1000, 300007, 5000000])
def test_scan_in_place_and_overlap(clops, oracle, ctx, n):
    """In place with equal widths runs (and takes the three-kernel scan for
    sizes the single pass would serve); other overlaps are refused."""
    q = clops.Queue(ctx)
    a = oracle.generate(oracle.GEN_U32_MASK, 0x54 + n, n, 0xFF)
    b = clops.Buffer.from_array(ctx, a)
    sc = clops.Scanner(ctx, np.uint32, np.uint32)
    sc.scan_device(q, b, b, n)
    q.finish()
    assert np.array_equal(b.read(q, np.uint32, n), oracle.scan(a, np.uint32))
    wide = clops.Scanner(ctx, np.uint32, np.uint64)
    big = clops.Buffer(ctx, n * 8 + 64)
    big.write(q, a)
    with pytest.raises(clops.ClopsError) as e:
        wide.scan_device(q, big, big, n)  # u32 -> u64 over its own input
    assert e.value.code == clops.CLO_ERROR_ARGS
    with pytest.raises(clops.ClopsError) as e:
        sc.scan_device(q, big, big.view(4, n * 4), n)
    assert e.value.code == clops.CLO_ERROR_ARGS
    off = clops.Buffer.from_array(ctx, np.array([0, n // 3], dtype=np.uint32))
    with pytest.raises(clops.ClopsError) as e:
        wide.scan_segmented_device(q, big, big, off, 2, n)
    assert e.value.code == clops.CLO_ERROR_ARGS


@pytest.mark.parametrize("kv", [False, True])
def test_malformed_segment_offsets_stay_in_bounds(clops, oracle, ctx, kv):
    """Decreasing offsets and offsets past numel: the order is unspecified,
    but the call completes and leaves memory outside the arrays alone (a guard
    region after each output is checked), and well-formed segments still sort."""
    q = clops.Queue(ctx)
    n = 200000
    k = oracle.generate(oracle.GEN_U32_HI, 0x55, n)
    # [0,100) and [100,300) are well formed; then a decrease (300 -> 250),
    # overlapping long segments, an offset past numel and one near 2^32.
    off = np.array([0, 100, 300, 250, 90000, 20000, 150000, 300000, 199000, 4000000000],
                   dtype=np.uint32)
    guard = 4096
    kin = clops.Buffer.from_array(ctx, k)
    kout = clops.Buffer(ctx, (n + guard) * 4)
    clops.generate(q, kout, n + guard, clops.GEN_U32_HI, 0x56)
    before = kout.read(q, np.uint32, n + guard)[n:]
    vin = clops.Buffer.from_array(ctx, np.arange(n, dtype=np.uint32)) if kv else None
    vout = clops.Buffer(ctx, n * 4) if kv else None
    offb = clops.Buffer.from_array(ctx, off)
    s = clops.Sorter(ctx, "satradix", np.uint32)
    s.sort_segmented_device(q, kin, offb, off.size, n, vin=vin, kout=kout, vout=vout)
    q.finish()
    assert np.array_equal(kout.read(q, np.uint32, n + guard)[n:], before)
    # Segment 0 = [0, 100) is well formed.
    assert np.array_equal(kout.read(q, np.uint32, 100), np.sort(k[:100]))
    sc = clops.Scanner(ctx, np.uint32, np.uint32)
    sout = clops.Buffer(ctx, (n + guard) * 4)
    clops.generate(q, sout, n + guard, clops.GEN_U32_HI, 0x57)
    before = sout.read(q, np.uint32, n + guard)[n:]
    sc.scan_segmented_device(q, kin, sout, offb, off.size, n)
    q.finish()
    assert np.array_equal(sout.read(q, np.uint32, n + guard)[n:], before)


def test_check_pairs_detects_mismatch(clops, oracle, ctx):
    q = clops.Queue(ctx)
    n = 100003
    k, v, kin, vin = _kv(clops, oracle, ctx, n, 0x58)
    rk, rv = oracle.stable_sort(k, v)
    ko, vo = clops.Buffer.from_array(ctx, rk), clops.Buffer.from_array(ctx, rv)
    assert clops.check_pairs(q, kin, ko, vo, n, 8) == 0
    assert clops.check_sorted(q, ko, n, 8, vo) == 0
    # Two payloads exchanged between distinct keys: two mismatched keys.
    bad = rv.copy()
    bad[[10, 20000]] = bad[[20000, 10]]
    vo.write(q, bad)
    assert clops.check_pairs(q, kin, ko, vo, n, 8) == 2
    # A payload repeated (and one index missing), and one out of range.
    bad = rv.copy()
    bad[5] = bad[6]
    bad[7] = n + 5
    vo.write(q, bad)
    assert clops.check_pairs(q, kin, ko, vo, n, 8) >= 3


def test_graph_holds_scratch(clops, oracle, ctx):
    """A captured sort keeps its scratch: a larger eager call is refused while
    the graph lives, smaller ones run, and replays stay exact even after the
    sorter is destroyed."""
    q = clops.Queue(ctx)
    n = 200003
    k, v, kin, vin = _kv(clops, oracle, ctx, n, 0x59)
    kout, vout = clops.Buffer(ctx, n * 8), clops.Buffer(ctx, n * 4)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    s.sort_kv_device(q, kin, vin, kout, vout, n)
    q.finish()
    g = clops.Graph(q)
    g.begin()
    try:
        s.sort_kv_device(q, kin, vin, kout, vout, n)
    finally:
        g.end()
    m = 4 * n
    k2, v2, kin2, vin2 = _kv(clops, oracle, ctx, m, 0x5A)
    kout2, vout2 = clops.Buffer(ctx, m * 8), clops.Buffer(ctx, m * 4)
    with pytest.raises(clops.ClopsError) as e:
        s.sort_kv_device(q, kin2, vin2, kout2, vout2, m)
    assert e.value.code == clops.CLO_ERROR_ARGS
    small = 5000
    s.sort_kv_device(q, kin2, vin2, kout2, vout2, small)  # fits the scratch
    q.finish()
    rk, rv = oracle.stable_sort(k2[:small], v2[:small])
    assert np.array_equal(kout2.read(q, np.uint64, small), rk)
    rk, rv = oracle.stable_sort(k, v)
    for rep in range(3):
        clops.generate(q, kout, n * 2, clops.GEN_U32_HI, 0x5B + rep)
        if rep == 1:
            s.close()  # the graph keeps the scratch alive
        q.finish()
        g.launch()
        q.finish()
        assert np.array_equal(kout.read(q, np.uint64, n), rk), rep
        assert np.array_equal(vout.read(q, np.uint32, n), rv), rep
    g.close()
    # After the graph is gone, a fresh sorter grows its own scratch freely.
    s2 = clops.Sorter(ctx, "satradix", np.uint64)
    s2.sort_kv_device(q, kin2, vin2, kout2, vout2, m)
    q.finish()
    assert clops.check_pairs(q, kin2, kout2, vout2, m, 8) == 0
