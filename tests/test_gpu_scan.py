"""blelloch (decoupled look-back) scan on the GPU vs the oracle, bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INT_TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]


def _vals(oracle, dtype, n, seed):
    u = oracle.generate(oracle.GEN_U64, seed, n)
    return u.astype(np.dtype("u%d" % np.dtype(dtype).itemsize)).view(dtype)


# 2^21 +- 1: both sides of the switch from the single-pass scan (<= 512 tiles
# of 4096) to reduce-then-scan.
@pytest.mark.parametrize("n", [0, 1, 2, 255, 4095, 4096, 4097, 8193, 100000, 1 << 20,
                               (1 << 21) - 1, 1 << 21, (1 << 21) + 1])
@pytest.mark.parametrize("elem,sumt", [(np.uint32, np.uint32), (np.uint32, np.uint64),
                                       (np.uint64, np.uint64)])
def test_scan_sizes(clops, oracle, ctx, queue, n, elem, sumt):
    sc = clops.Scanner(ctx, elem, sumt)
    a = _vals(oracle, elem, n, 0x5C + n)
    want = oracle.scan(a, sumt)
    assert np.array_equal(sc.scan_host(a), want)
    if n:  # device data: one call over the whole array (no host chunking)
        bi = clops.Buffer.from_array(ctx, a)
        bo = clops.Buffer(ctx, n * np.dtype(sumt).itemsize)
        sc.scan_device(queue, bi, bo, n)
        queue.finish()
        assert np.array_equal(bo.read(queue, sumt, n), want)


@pytest.mark.parametrize("elem", INT_TYPES)
@pytest.mark.parametrize("sumt", INT_TYPES)
def test_scan_type_pairs(clops, oracle, ctx, elem, sumt):
    n = 9000
    sc = clops.Scanner(ctx, elem, sumt)
    a = _vals(oracle, elem, n, 0x77)
    assert np.array_equal(sc.scan_host(a), oracle.scan(a, sumt))


def test_scan_large_u32_wrap(clops, oracle, ctx, queue):
    """2^26 values in [0, 255] scanned to uint32; checks against the oracle."""
    n = 1 << 26
    a = (oracle.generate(oracle.GEN_U32_HI, 0x2C, n) & np.uint32(0xFF))
    bi = clops.Buffer.from_array(ctx, a)
    bo = clops.Buffer(ctx, n * 4)
    sc = clops.Scanner(ctx, np.uint32, np.uint32)
    sc.scan_device(queue, bi, bo, n)
    queue.finish()
    out = bo.read(queue, np.uint32, n)
    assert np.array_equal(out, oracle.scan(a, np.uint32))


def test_scan_repeat_epochs(clops, oracle, ctx, queue):
    """Many launches on one scanner (status words reused across epochs)."""
    sc = clops.Scanner(ctx, np.uint32, np.uint64)
    for i, n in enumerate([5000, 300000, 17, 300000, 70000]):
        a = _vals(oracle, np.uint32, n, 0x99 + i)
        assert np.array_equal(sc.scan_host(a), oracle.scan(a, np.uint64))


@pytest.mark.parametrize("polls", [0, 1])
@pytest.mark.parametrize("elem,sumt", [(np.uint32, np.uint32), (np.uint32, np.uint64),
                                       (np.uint64, np.uint64)])
def test_scan_lookback_fallback(clops, oracle, ctx, queue, polls, elem, sumt):
    """The single-pass scan with its look-back wait cut to `polls` rounds:
    a tile that finds a predecessor unpublished sums that tile's input
    itself.  The results stay exact, also over repeated launches (epochs and
    the started count reused) and through the host path's chunk chain."""
    L = clops.lib()
    sc = clops.Scanner(ctx, elem, sumt)
    L.clo_scan_debug_lookback_polls(polls)
    try:
        for rep in range(3):
            for n in [8193, 100000, 1 << 20, (1 << 21) - 1, 1 << 21]:
                a = _vals(oracle, elem, n, 0xFB + n + rep)
                want = oracle.scan(a, sumt)
                bi = clops.Buffer.from_array(ctx, a)
                bo = clops.Buffer(ctx, n * np.dtype(sumt).itemsize)
                sc.scan_device(queue, bi, bo, n)
                queue.finish()
                assert np.array_equal(bo.read(queue, sumt, n), want), (rep, n)
                assert np.array_equal(sc.scan_host(a), want), (rep, n)
    finally:
        L.clo_scan_debug_lookback_polls(-1)
    a = _vals(oracle, elem, 1 << 20, 0x3)
    assert np.array_equal(sc.scan_host(a), oracle.scan(a, sumt))


def test_scan_options_rejected(clops, ctx):
    with pytest.raises(clops.ClopsError) as e:
        clops.Scanner(ctx, np.uint32, np.uint32, options="x=1")
    assert e.value.code == clops.CLO_ERROR_ARGS
    with pytest.raises(clops.ClopsError) as e:
        clops.Scanner(ctx, np.uint32, np.uint32, algo="hillis")
    assert e.value.code == clops.CLO_ERROR_IMPL_NOT_FOUND
