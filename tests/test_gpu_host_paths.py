"""Host-data paths (pinned staging, chunked copy/compute overlap) against
the oracle at sizes spanning several staging chunks (8 MiB each, three per
direction), with and without a separate transfer queue."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("elem,sumt,n", [
    (np.uint32, np.uint64, 7000003),   # 1 Mi elements per chunk: 7 chunks
    (np.uint8, np.uint16, 9000001),    # 4 Mi per chunk: 3 chunks
    (np.int16, np.int32, 5000000),
    (np.uint32, np.uint32, 2097152),   # exactly one chunk
    (np.int64, np.int64, 3000017),
    (np.uint32, np.uint32, 1),
    (np.uint32, np.uint32, 0),
    # One chunk of 2 MiB + 1 B in, 4 MiB + 2 B out: byte counts whose share
    # per copy thread, floor(bytes / threads), is a multiple of 64 for 2, 4
    # and 8 threads, so a split that rounds the share down drops the tail.
    (np.uint8, np.uint16, 2097153),
])
@pytest.mark.parametrize("two_queues", [False, True])
def test_scan_host_chunks(clops, oracle, ctx, elem, sumt, n, two_queues):
    q = clops.Queue(ctx)
    qc = clops.Queue(ctx) if two_queues else None
    rs = np.random.default_rng(n)
    info = np.iinfo(elem)
    x = rs.integers(max(info.min, -1000), min(info.max, 1000), n, endpoint=True).astype(elem)
    sc = clops.Scanner(ctx, elem, sumt)
    got = sc.scan_host(x, q, qc)
    assert np.array_equal(got, oracle.scan(x, sumt))
    # the same object again: cached device buffers, chain reset
    assert np.array_equal(sc.scan_host(x[: n // 2], q, qc), oracle.scan(x[: n // 2], sumt))


@pytest.mark.parametrize("algo,dtype,n", [
    ("satradix", np.uint32, 5000011),   # 3 chunks of keys
    ("satradix", np.int64, 3000017),
    ("abitonic", np.uint64, 2500001),
    ("gselect", np.uint32, 3001),       # not in place: a second device buffer
    ("sbitonic", np.int8, 10000000),
    ("satradix", np.uint8, 4194305),    # one chunk of 4 MiB + 1 B (see the scan case)
])
@pytest.mark.parametrize("two_queues", [False, True])
def test_sort_host_chunks(clops, oracle, ctx, algo, dtype, n, two_queues):
    q = clops.Queue(ctx)
    qc = clops.Queue(ctx) if two_queues else None
    x = oracle.generate(oracle.GEN_U64, n + 5, n).astype(
        np.dtype("u%d" % np.dtype(dtype).itemsize)).view(dtype)
    s = clops.Sorter(ctx, algo, dtype)
    u = "u%d" % np.dtype(dtype).itemsize

    def want(a):  # satradix orders signed keys by their unsigned bits, as the reference
        return np.sort(a.view(u)).view(dtype) if algo == "satradix" else np.sort(a)
    assert np.array_equal(s.sort_host(x, q, qc), want(x))
    assert np.array_equal(s.sort_host(x[: n // 3], q, qc), want(x[: n // 3]))


@pytest.mark.parametrize("n", [3000017, 1048577])  # 1048577: values of 4 MiB + 4 B
@pytest.mark.parametrize("two_queues", [False, True])
def test_sort_kv_host_chunks(clops, oracle, ctx, two_queues, n):
    q = clops.Queue(ctx)
    qc = clops.Queue(ctx) if two_queues else None
    k = oracle.generate(oracle.GEN_U64_MASK, 7, n, param=(1 << 12) - 1)
    v = np.arange(n, dtype=np.uint32)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    rk, rv = oracle.stable_sort(k, v)
    gk, gv = s.sort_kv_host(k, v, q, qc)
    assert np.array_equal(gk, rk) and np.array_equal(gv, rv)
    g = clops.Sorter(ctx, "gselect", np.uint64)
    gk, gv = g.sort_kv_host(k[:2000], v[:2000], q, qc)
    rk, rv = oracle.stable_sort(k[:2000], v[:2000])
    assert np.array_equal(gk, rk) and np.array_equal(gv, rv)


def test_snippet_sorter_host(clops, ctx):
    n = 3000000
    x = np.random.default_rng(3).standard_normal(n).astype(np.float32)
    s = clops.Sorter(ctx, "satradix", np.float32, key_dtype=np.uint32,
                     get_key="(as_uint(x) ^ ((uint)(as_int(x) >> 31) | 0x80000000u))")
    assert np.array_equal(s.sort_host(x), np.sort(x))
