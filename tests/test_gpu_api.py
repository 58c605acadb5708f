"""Boundary behaviour of the C API on the GPU: factory errors, introspection,
queues, events, profiling and the allocation-failure path."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_sort_factory_errors(clops, ctx):
    with pytest.raises(clops.ClopsError) as e:
        clops.Sorter(ctx, "quicksort", np.uint32)
    assert e.value.code == clops.CLO_ERROR_IMPL_NOT_FOUND
    for bad in ("radix=3", "radix=512", "nope=1", "radix"):
        with pytest.raises(clops.ClopsError) as e:
            clops.Sorter(ctx, "satradix", np.uint32, options=bad)
        assert e.value.code == clops.CLO_ERROR_ARGS
    with pytest.raises(clops.ClopsError) as e:
        clops.Sorter(ctx, "satradix", np.uint32, options="scan=hillis")
    assert e.value.code == clops.CLO_ERROR_IMPL_NOT_FOUND
    with pytest.raises(clops.ClopsError) as e:
        clops.Sorter(ctx, "sbitonic", np.uint32, compare="((a).x > (b).x)")
    assert e.value.code == clops.CLO_ERROR_ARGS


def test_introspection(clops, ctx):
    s = clops.Sorter(ctx, "satradix", np.uint64, options="radix=256,scan=blelloch,scanfoo=1")
    L = clops.lib()
    assert L.clo_sort_get_element_size(s.ptr) == 8
    assert L.clo_sort_get_key_type(s.ptr) == clops.CLO_ULONG
    assert s.kernel_names() == ["satradix_upsweep", "satradix_plan", "satradix_reduce",
                                "satradix_scan", "satradix_downsweep", "satradix_copy"]
    assert clops._call(L.clo_sort_get_localmem_usage, s.ptr, 4, 256, 1000) > 0
    sc = clops.Scanner(ctx, np.uint32, np.uint64)
    assert L.clo_scan_get_sum_size(sc.ptr) == 8
    assert [clops._call(L.clo_scan_get_kernel_name, sc.ptr, i) for i in range(4)] == [
        b"blelloch_tile_reduce", b"blelloch_tiles_scan", b"blelloch_tile_scan",
        b"blelloch_onepass"]


def test_profiling_records_kernels(clops, oracle, ctx):
    q = clops.Queue(ctx, profiling=True)
    n = 1 << 16
    k = oracle.generate(oracle.GEN_U64, 1, n)
    b = clops.Buffer.from_array(ctx, k)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    q.prof_reset()
    s.sort_device(q, b, None, n)
    cnt, ms = q.prof("satradix_downsweep")
    assert cnt == 8 and ms > 0
    assert "satradix_upsweep" in q.prof_names()
    # With a name filter only the matching kernels are recorded.
    q.prof_filter("satradix_downsweep")
    q.prof_reset()
    s.sort_device(q, b, None, n)
    assert q.prof_names() == ["satradix_downsweep"]
    cnt2, ms2 = q.prof("")
    assert cnt2 == 8 and ms2 > 0
    q.prof_filter("")
    q.prof_reset()
    s.sort_device(q, b, None, n)
    assert "satradix_upsweep" in q.prof_names()
    assert np.array_equal(b.read(q, np.uint64, n), np.sort(k))


def test_alloc_failure_reported(clops, oracle, ctx, queue):
    L = clops.lib()
    s = clops.Sorter(ctx, "satradix", np.uint32)
    n = 5000
    k = oracle.generate(oracle.GEN_U32_HI, 3, n)
    b = clops.Buffer.from_array(ctx, k)
    L.ccl_debug_fail_next_allocs(1)
    with pytest.raises(clops.ClopsError) as e:
        s.sort_device(queue, b, None, n)
    assert e.value.domain == L.ccl_error_quark()
    L.ccl_debug_fail_next_allocs(0)
    s.sort_device(queue, b, None, n)  # recovers
    queue.finish()
    assert np.array_equal(b.read(queue, np.uint32, n), np.sort(k))


def test_buffer_too_small(clops, ctx, queue):
    s = clops.Sorter(ctx, "satradix", np.uint32)
    b = clops.Buffer(ctx, 16)
    with pytest.raises(clops.ClopsError) as e:
        s.sort_device(queue, b, None, 100)
    assert e.value.code == clops.CLO_ERROR_ARGS


def test_two_queues_host_data(clops, oracle, ctx):
    qe, qc = clops.Queue(ctx), clops.Queue(ctx)
    s = clops.Sorter(ctx, "satradix", np.uint32)
    k = oracle.generate(oracle.GEN_U32_HI, 8, 77777)
    L = clops.lib()
    import ctypes
    out = np.empty_like(k)
    clops._call(L.clo_sort_with_host_data, s.ptr, qe.ptr, qc.ptr,
                k.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p), k.size, 0)
    assert np.array_equal(out, np.sort(k))


def test_device_generators_match_oracle(clops, oracle, ctx, queue):
    n = 100000
    for kind, dt, param in [(clops.GEN_U64, np.uint64, 0), (clops.GEN_U32_HI, np.uint32, 0),
                            (clops.GEN_U64_MASK, np.uint64, (1 << 20) - 1),
                            (clops.GEN_U32_MOD, np.uint32, 1000)]:
        b = clops.Buffer(ctx, n * np.dtype(dt).itemsize)
        clops.generate(queue, b, n, kind, 0xABC, param, 5)
        assert np.array_equal(b.read(queue, dt, n), oracle.generate(kind, 0xABC, n, param, 5))


def test_debug_log_lines(clops, oracle, ctx, capfd):
    """Level 1 logs one line per call with the chosen path and sizes, level 2
    also every kernel launch; level 0 is silent."""
    import ctypes
    L = clops.lib()
    L.clo_set_log_level.argtypes = [ctypes.c_int]
    q = clops.Queue(ctx)
    n = 100000
    k = oracle.generate(oracle.GEN_U64, 3, n)
    b, o = clops.Buffer.from_array(ctx, k), clops.Buffer(ctx, n * 8)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    sc = clops.Scanner(ctx, np.uint32, np.uint32)
    a = clops.Buffer.from_array(ctx, np.arange(n, dtype=np.uint32))
    capfd.readouterr()
    L.clo_set_log_level(1)
    try:
        s.sort_device(q, b, o, n)
        sc.scan_device(q, a, o, n)
        q.finish()
    finally:
        L.clo_set_log_level(0)
    err = capfd.readouterr().err
    assert "clo: satradix: numel=100000 key=8 B radix=256 (bits_in_digit=8)" in err, err
    assert "clo: blelloch: numel=100000 elem=4 B sum=4 B" in err, err
    assert "clo: launch" not in err
    L.clo_set_log_level(2)
    try:
        s.sort_device(q, b, o, n)
        q.finish()
    finally:
        L.clo_set_log_level(0)
    err = capfd.readouterr().err
    assert "clo: launch satradix_downsweep" in err, err
    s.sort_device(q, b, o, n)
    q.finish()
    assert "clo:" not in capfd.readouterr().err
    assert np.array_equal(o.read(q, np.uint64, n), np.sort(k))
