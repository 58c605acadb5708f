"""satradix on the GPU vs the CPU oracle, bit for bit (keys, stable KV order)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 3, 7, 64, 255, 256, 257, 1000, 4095, 4096, 4097, 12345, 65536, 100003]


def _keys(oracle, dtype, n, seed, dist="uniform"):
    w = np.dtype(dtype).itemsize
    u = oracle.generate(oracle.GEN_U64, seed, n)
    if dist == "uniform":
        k = u
    elif dist == "equal":
        k = np.full(n, 0x5A5A5A5A5A5A5A5A, dtype=np.uint64)
    elif dist == "sorted":
        k = np.sort(u)
    elif dist == "reversed":
        k = np.sort(u)[::-1].copy()
    elif dist == "few":
        k = u % np.uint64(16)
    elif dist == "highbit":
        k = (u & np.uint64(1)) << np.uint64(8 * w - 1)
    else:
        raise ValueError(dist)
    return k.astype(np.dtype("u%d" % w)).view(dtype)


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64, np.uint8, np.uint16, np.int32, np.int64])
@pytest.mark.parametrize("n", SIZES)
def test_keys_only_sizes(clops, oracle, ctx, dtype, n):
    s = clops.Sorter(ctx, "satradix", dtype)
    k = _keys(oracle, dtype, n, 0xA1 + n)
    out = s.sort_host(k)
    ref = oracle.stable_sort(k.view("u%d" % k.dtype.itemsize)).view(dtype)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("dist", ["uniform", "equal", "sorted", "reversed", "few", "highbit"])
@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
def test_kv_distributions(clops, oracle, ctx, dtype, dist):
    n = 50001
    s = clops.Sorter(ctx, "satradix", dtype)
    k = _keys(oracle, dtype, n, 0xB2, dist)
    v = np.arange(n, dtype=np.uint32)
    ko, vo = s.sort_kv_host(k, v)
    rk, rv = oracle.stable_sort(k, v)
    assert np.array_equal(ko, rk)
    assert np.array_equal(vo, rv)


@pytest.mark.parametrize("radix", [2, 4, 16, 32, 64, 128, 256])
def test_radix_option(clops, oracle, ctx, radix):
    n = 20000
    s = clops.Sorter(ctx, "satradix", np.uint32, options="radix=%d" % radix)
    k = _keys(oracle, np.uint32, n, 0xC3 + radix)
    v = np.arange(n, dtype=np.uint32)
    ko, vo = s.sort_kv_host(k, v)
    rk, rv = oracle.satradix(k, v, bits=int(np.log2(radix)), lws=256)
    assert np.array_equal(ko, rk) and np.array_equal(vo, rv)


@pytest.mark.parametrize("opt", ["wg=256,ipt=32", "wg=512,ipt=16", "wg=256,ipt=16",
                                 "wg=512,ipt=32"])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("n", [1, 4095, 4097, 8193, 50001, 300007])
def test_tile_shapes(clops, oracle, ctx, opt, dtype, n):
    s = clops.Sorter(ctx, "satradix", dtype, options=opt)
    k = _keys(oracle, dtype, n, 0xE7 + n, "uniform" if n % 2 else "few")
    v = np.arange(n, dtype=np.uint32)
    ko, vo = s.sort_kv_host(k, v)
    rk, rv = oracle.stable_sort(k, v)
    assert np.array_equal(ko, rk) and np.array_equal(vo, rv)
    assert np.array_equal(s.sort_host(k), rk)


def test_tile_shape_option_errors(clops, ctx):
    for bad in ("wg=128", "ipt=8", "ipt=64", "radix=16,wg=512"):
        with pytest.raises(clops.ClopsError) as e:
            clops.Sorter(ctx, "satradix", np.uint32, options=bad)
        assert e.value.code == clops.CLO_ERROR_ARGS


def test_config1_2p20_u32(clops, oracle, ctx, queue):
    """2^20 uint32 keys over the full 32-bit range (seed 0xC1), keys only."""
    n = 1 << 20
    k = oracle.generate(oracle.GEN_U32_HI, 0xC1, n)
    b = clops.Buffer.from_array(ctx, k)
    s = clops.Sorter(ctx, "satradix", np.uint32)
    s.sort_device(queue, b, None, n)
    queue.finish()
    out = b.read(queue, np.uint32, n)
    assert np.array_equal(out, np.sort(k))
    # the reference-structured oracle (4-bit digits, lws 256) gives the same array
    assert np.array_equal(out, oracle.satradix(k, bits=4, lws=256, threads=8))


def test_out_of_place_keeps_input(clops, oracle, ctx, queue):
    n = 30000
    k = _keys(oracle, np.uint64, n, 0xD4)
    bi = clops.Buffer.from_array(ctx, k)
    bo = clops.Buffer(ctx, k.nbytes)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    s.sort_device(queue, bi, bo, n)
    queue.finish()
    assert np.array_equal(bo.read(queue, np.uint64, n), np.sort(k))
    assert np.array_equal(bi.read(queue, np.uint64, n), k)


def test_constant_digit_passes_skipped(clops, oracle, ctx, queue):
    n = 100000
    k = oracle.generate(oracle.GEN_U64_MASK, 0xC3B, n, param=(1 << 20) - 1)
    v = np.arange(n, dtype=np.uint32)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    for inplace in (True, False):
        kb, vb = clops.Buffer.from_array(ctx, k), clops.Buffer.from_array(ctx, v)
        ko = None if inplace else clops.Buffer(ctx, k.nbytes)
        vo = None if inplace else clops.Buffer(ctx, v.nbytes)
        s.sort_kv_device(queue, kb, vb, ko, vo, n)
        queue.finish()
        assert s.last_passes() == 3
        rk, rv = oracle.stable_sort(k, v)
        assert np.array_equal((ko or kb).read(queue, np.uint64, n), rk)
        assert np.array_equal((vo or vb).read(queue, np.uint32, n), rv)


def test_repeat_deterministic(clops, oracle, ctx, queue):
    n = 1 << 18
    k = _keys(oracle, np.uint64, n, 0xE5)
    v = np.arange(n, dtype=np.uint32)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    outs = []
    for _ in range(3):
        outs.append(s.sort_kv_host(k, v))
    rk, rv = oracle.stable_sort(k, v)
    for ko, vo in outs:
        assert np.array_equal(ko, rk) and np.array_equal(vo, rv)


def test_zipf_kv(clops, oracle, ctx, queue):
    n = 1 << 20
    cdf = oracle.zipf_cdf(1 << 16, 1.1)
    k = oracle.generate_zipf(cdf, 0xC3C, n)
    v = np.arange(n, dtype=np.uint32)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    ko, vo = s.sort_kv_host(k, v)
    rk, rv = oracle.stable_sort(k, v)
    assert np.array_equal(ko, rk) and np.array_equal(vo, rv)


@pytest.mark.parametrize("opt", ["", "wg=512,ipt=16", "wg=256,ipt=16"])
@pytest.mark.parametrize("delta", [-1, 0, 1])
def test_counter_scan_boundary(clops, oracle, ctx, opt, delta):
    """Sorts of at most 256 tiles have each downsweep block sum the tile
    counts itself; one more tile switches to the grouped reduce + scan.  Both
    sides of the switch, for the tile sizes of each shape (8192 or 4096 keys),
    keys-only and KV."""
    tile = 4096 if opt == "wg=256,ipt=16" else 8192
    n = 256 * tile + delta
    s = clops.Sorter(ctx, "satradix", np.uint64, options=opt or None)
    k = _keys(oracle, np.uint64, n, 0xB0 + delta)
    v = np.arange(n, dtype=np.uint32)
    ko, vo = s.sort_kv_host(k, v)
    rk, rv = oracle.stable_sort(k, v)
    assert np.array_equal(ko, rk) and np.array_equal(vo, rv)
    s32 = clops.Sorter(ctx, "satradix", np.uint32, options=opt or None)
    k32 = k.astype(np.uint32)
    assert np.array_equal(s32.sort_host(k32), np.sort(k32))
