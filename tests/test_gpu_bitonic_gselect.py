"""sbitonic, abitonic and gselect on the GPU vs the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INT_TYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]


def _keys(oracle, dtype, n, seed):
    u = oracle.generate(oracle.GEN_U64, seed, n)
    return u.astype(np.dtype("u%d" % np.dtype(dtype).itemsize)).view(dtype)


@pytest.mark.parametrize("algo", ["sbitonic", "abitonic"])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 8, 100, 1000, 1024, 4097, 8192, 8193, 70000])
def test_bitonic_sizes(clops, oracle, ctx, algo, n):
    s = clops.Sorter(ctx, algo, np.uint32)
    k = _keys(oracle, np.uint32, n, 0xB1 + n)
    assert np.array_equal(s.sort_host(k), np.sort(k))


@pytest.mark.parametrize("algo", ["sbitonic", "abitonic"])
@pytest.mark.parametrize("n", [2047, 2048, 4095, 4096, 4097, 1 << 15, (1 << 17) + 3])
def test_bitonic_sizes_u64(clops, oracle, ctx, algo, n):
    """8-byte keys run 2^12-key tiles: sizes around the tile and the
    register/LDS step split, power-of-two (check-free) and padded networks."""
    s = clops.Sorter(ctx, algo, np.uint64)
    k = _keys(oracle, np.uint64, n, 0xB3 + n)
    assert np.array_equal(s.sort_host(k), np.sort(k))


@pytest.mark.parametrize("algo", ["sbitonic", "abitonic"])
@pytest.mark.parametrize("dtype", INT_TYPES)
@pytest.mark.parametrize("desc", [False, True])
def test_bitonic_types_order(clops, oracle, ctx, algo, dtype, desc):
    n = 5000
    s = clops.Sorter(ctx, algo, dtype, compare="((a) < (b))" if desc else None)
    k = _keys(oracle, dtype, n, 0xB7)
    ref = np.sort(k)[::-1] if desc else np.sort(k)
    assert np.array_equal(s.sort_host(k), ref)


@pytest.mark.parametrize("dtype", [np.uint32, np.int64])
def test_bitonic_po2_matches_reference_network(clops, oracle, ctx, dtype):
    n = 1 << 14
    k = _keys(oracle, dtype, n, 0xB9)
    for algo in ("sbitonic", "abitonic"):
        s = clops.Sorter(ctx, algo, dtype)
        assert np.array_equal(s.sort_host(k), oracle.sbitonic(k))


@pytest.mark.parametrize("opts", ["minps=1,maxps=1", "minps=2,maxps=2", "minps=3,maxps=3",
                                  "maxps=4", "maxsfs=0", "maxsfs=5", "minps=1,maxps=3,maxsfs=9"])
def test_abitonic_options(clops, oracle, ctx, opts):
    n = 1 << 17
    s = clops.Sorter(ctx, "abitonic", np.uint32, options=opts)
    k = _keys(oracle, np.uint32, n, 0xBA)
    assert np.array_equal(s.sort_host(k), np.sort(k))


@pytest.mark.parametrize("regkeys", [0, 16, 32])
@pytest.mark.parametrize("maxsfs", [5, 9, 11, 12, 13, 14, 15])
@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint32, np.int64])
def test_abitonic_tile_kernels(clops, oracle, ctx, regkeys, maxsfs, dtype):
    """Every tile kernel (plain LDS steps, 16 or 32 register keys per
    thread) at every tile size, on power-of-two and ragged sizes."""
    s = clops.Sorter(ctx, "abitonic", dtype, options="regkeys=%d,maxsfs=%d" % (regkeys, maxsfs))
    for n in (1 << 15, 40003):
        k = _keys(oracle, dtype, n, 0xBB + n)
        assert np.array_equal(s.sort_host(k), np.sort(k))


@pytest.mark.parametrize("dtype", INT_TYPES)
@pytest.mark.parametrize("desc", [False, True])
def test_abitonic_local_passes(clops, oracle, ctx, dtype, desc):
    """Sizes of several default tiles (2^12..2^14 keys), whole and ragged, so
    every stage above the tile ends in a local pass of half-cleaners, in both
    orders and for signed keys."""
    s = clops.Sorter(ctx, "abitonic", dtype, compare="((a) < (b))" if desc else None)
    for n in (1 << 17, (1 << 16) + 4099):
        k = _keys(oracle, dtype, n, 0xBC + n)
        ref = np.sort(k)[::-1] if desc else np.sort(k)
        assert np.array_equal(s.sort_host(k), ref)


def test_abitonic_bad_options(clops, ctx):
    for bad in ("minps=0", "maxps=5", "minps=3,maxps=2", "foo=1", "minps", "regkeys=8"):
        with pytest.raises(clops.ClopsError) as e:
            clops.Sorter(ctx, "abitonic", np.uint32, options=bad)
        assert e.value.code == clops.CLO_ERROR_ARGS


def test_abitonic_2p24_sweeps(clops, oracle, ctx, queue):
    n = 1 << 24
    k = oracle.generate(oracle.GEN_U32_HI, 0xC3D, n)
    b = clops.Buffer.from_array(ctx, k)
    s = clops.Sorter(ctx, "abitonic", np.uint32)
    s.sort_device(queue, b, None, n)
    queue.finish()
    assert np.array_equal(b.read(queue, np.uint32, n), np.sort(k))
    assert s.last_sweeps() > 0


@pytest.mark.parametrize("dtype", INT_TYPES)
@pytest.mark.parametrize("desc", [False, True])
def test_gselect_kv_stable(clops, oracle, ctx, dtype, desc):
    n = 3000
    k = (_keys(oracle, dtype, n, 0x95).view("u%d" % np.dtype(dtype).itemsize) % 50).astype(
        "u%d" % np.dtype(dtype).itemsize).view(dtype)
    v = np.arange(n, dtype=np.uint32)
    s = clops.Sorter(ctx, "gselect", dtype, compare="((a) < (b))" if desc else None)
    ko, vo = s.sort_kv_host(k, v)
    rk, rv = oracle.gselect(k, v, desc=desc)
    assert np.array_equal(ko, rk) and np.array_equal(vo, rv)
    assert np.array_equal(s.sort_host(k), rk)


@pytest.mark.parametrize("n", [0, 1, 2, 255, 256, 1023, 1024, 1025, 5000])
def test_gselect_sizes(clops, oracle, ctx, n):
    s = clops.Sorter(ctx, "gselect", np.uint64)
    k = _keys(oracle, np.uint64, n, 0x96 + n)
    assert np.array_equal(s.sort_host(k), np.sort(k))


def test_bitonic_rejects_kv(clops, ctx, queue):
    s = clops.Sorter(ctx, "abitonic", np.uint32)
    b = clops.Buffer(ctx, 64)
    with pytest.raises(clops.ClopsError) as e:
        s.sort_kv_device(queue, b, b, None, None, 4)
    assert e.value.code == clops.CLO_ERROR_ARGS


def test_abitonic_default_step_groups(clops, oracle, ctx):
    """Without maxps, 4-byte keys fuse up to 5 global steps per launch
    (abit_priv_5s32v); maxps=4 keeps the reference's limit.  Same result."""
    n = (1 << 20) + 12345
    k = _keys(oracle, np.uint32, n, 0xAB)
    q = clops.Queue(ctx, profiling=True)
    outs = {}
    for opt in (None, "maxps=4"):
        s = clops.Sorter(ctx, "abitonic", np.uint32, options=opt)
        kb = clops.Buffer.from_array(ctx, k)
        q.prof_reset()
        s.sort_device(q, kb, None, n)
        q.finish()
        outs[opt] = (kb.read(q, np.uint32, n), q.prof_names())
    assert np.array_equal(outs[None][0], np.sort(k))
    assert np.array_equal(outs["maxps=4"][0], np.sort(k))
    assert "abit_priv_5s32v" in outs[None][1]
    assert "abit_priv_5s32v" not in outs["maxps=4"][1]
