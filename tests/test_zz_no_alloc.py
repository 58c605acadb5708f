"""clo_scan_with_device_data_noalloc through the C ABI against
clo_scan_with_device_data: bit-exact on two calls in a row per case, the
second with a smaller input, on both sides of the single-pass / three-kernel
switch; limits and aliasing rules."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(clops, sc, q, bi, bo, n):
    ok = clops._call(clops.lib().clo_scan_with_device_data_noalloc, sc.ptr, q.ptr, bi.ptr, bo.ptr,
                     n)
    assert ok == 1


CASES = [
    # (element, sum, first n, second n): 4096-element tiles, the single pass
    # up to 512 tiles
    (np.uint32, np.uint32, 5_000_003, 300_007),
    (np.uint32, np.uint64, (1 << 21) + 5, 1000),
    (np.uint64, np.uint64, 1_000_000, 4097),
    (np.int32, np.int64, 3_000_000, 2_999_999),
    (np.uint8, np.uint32, 70_000, 1),
]


@pytest.mark.parametrize("elem,sumt,n1,n2", CASES)
def test_noalloc_scan_matches_scan(clops, oracle, ctx, elem, sumt, n1, n2):
    q = clops.Queue(ctx)
    sc = clops.Scanner(ctx, elem, sumt)
    ew, sw = np.dtype(elem).itemsize, np.dtype(sumt).itemsize
    for rep, n in enumerate((n1, n2)):
        kind = clops.GEN_U64 if ew == 8 else clops.GEN_U32_HI
        raw = clops.Buffer(ctx, n * 8)
        clops.generate(q, raw, n * 2 if kind == clops.GEN_U32_HI else n, kind, 0x7A + rep)
        x = raw.read(q, np.dtype("u%d" % ew), n).view(elem)
        bi = clops.Buffer.from_array(ctx, x)
        want_b = clops.Buffer(ctx, n * sw)
        sc.scan_device(q, bi, want_b, n)
        got_b = clops.Buffer(ctx, n * sw)
        clops.generate(q, got_b, n * sw // 4, clops.GEN_U32_HI, 0xBAD)  # stale output
        _noalloc(clops, sc, q, bi, got_b, n)
        q.finish()
        got, want = got_b.read(q, sumt, n), want_b.read(q, sumt, n)
        assert np.array_equal(got, want), (rep, n)
        assert np.array_equal(got, oracle.scan(x, sumt)), (rep, n)


def test_noalloc_scan_limits(clops, ctx):
    L = clops.lib()
    q = clops.Queue(ctx)
    sc = clops.Scanner(ctx, np.uint32, np.uint64)
    maxn = L.clo_scan_noalloc_max_numel(sc.ptr)
    assert maxn >= (1 << 28)
    b = clops.Buffer(ctx, 64)
    with pytest.raises(clops.ClopsError) as e:
        # Refused on the size before the buffers are looked at.
        _noalloc(clops, sc, q, b, b, maxn + 1)
    assert e.value.code == clops.CLO_ERROR_ARGS and "limit" in e.value.message
    with pytest.raises(clops.ClopsError) as e:
        _noalloc(clops, sc, q, b, b, 8)  # u32 -> u64 over its own input
    assert e.value.code == clops.CLO_ERROR_ARGS
    _noalloc(clops, sc, q, b, clops.Buffer(ctx, 64), 0)  # empty: nothing to do


def test_noalloc_scan_profiling_queue_in_place(clops, oracle, ctx):
    """On a profiling queue it takes no records; in place with equal widths."""
    q = clops.Queue(ctx, profiling=True)
    sc = clops.Scanner(ctx, np.uint32, np.uint32)
    n = 100_000
    x = oracle.generate(oracle.GEN_U32_MASK, 0x7B, n, 0xFF)
    b = clops.Buffer.from_array(ctx, x)
    q.prof_reset()
    _noalloc(clops, sc, q, b, b, n)
    q.finish()
    assert q.prof_names() == []
    assert np.array_equal(b.read(q, np.uint32, n), oracle.scan(x, np.uint32))
