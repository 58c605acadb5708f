"""Sorters built from user compare / key snippets (hipRTC-compiled at
clo_sort_new) against the oracle: the reference's bitonic network bit for
bit (ties included), gselect's rank formula, and satradix's stable order by
the key bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U64 = np.uint64


def _hi32(x):
    return (x >> U64(32)).astype(np.uint32)


def _tied_u64(rs, n, distinct_hi=16):
    """u64 words whose top halves repeat (many ties under the _hi32 key) and
    whose low halves are distinct (so a tie order shows in the output)."""
    hi = rs.integers(0, distinct_hi, n).astype(U64)
    lo = rs.permutation(n).astype(U64)
    return (hi << U64(32)) | lo


def _sort(clops, ctx, q, s, x, out_of_place=True):
    b = clops.Buffer.from_array(ctx, x)
    o = clops.Buffer(ctx, max(x.nbytes, 1)) if out_of_place else None
    s.sort_device(q, b, o, x.size)
    q.finish()
    got = (o if out_of_place else b).read(q, x.dtype, x.size)
    if out_of_place:
        u = "u%d" % x.itemsize  # bitwise: NaN inputs compare unequal as floats
        assert np.array_equal(b.read(q, x.dtype, x.size).view(u), x.view(u))  # input untouched
    return got


@pytest.mark.parametrize("algo", ["sbitonic", "abitonic"])
@pytest.mark.parametrize("n", [2, 4, 64, 2048, 4096, 1 << 16, 1 << 20])
def test_bitonic_key_snippet_ties_in_reference_order(clops, oracle, ctx, queue, algo, n):
    """Key = high 32 bits of a u64: many ties, which the reference's network
    leaves in one specific order; the result matches it word for word.  The
    sizes cross the LDS tile (2048) and run global steps."""
    rs = np.random.default_rng(n)
    x = _tied_u64(rs, n)
    s = clops.Sorter(ctx, algo, np.uint64, key_dtype=np.uint32, get_key="((uint)((x) >> 32))")
    want = oracle.bitonic_network(x, key=_hi32)
    assert np.array_equal(_sort(clops, ctx, queue, s, x), want)
    assert np.array_equal(_sort(clops, ctx, queue, s, x, out_of_place=False), want)


@pytest.mark.parametrize("algo", ["sbitonic", "abitonic"])
@pytest.mark.parametrize("n", [3, 1000, 2049, 5000, 300007])
def test_bitonic_other_sizes(clops, oracle, ctx, queue, algo, n):
    """Sizes the reference cannot take: the padded network (flagged slots
    order last), word for word."""
    rs = np.random.default_rng(n + 1)
    x = _tied_u64(rs, n)
    s = clops.Sorter(ctx, algo, np.uint64, key_dtype=np.uint32, get_key="((uint)((x) >> 32))",
                     compare="((a) < (b))")
    want = oracle.bitonic_network(x, key=_hi32, compare=np.less, padded=True)
    assert np.all(np.diff(_hi32(want).astype(np.int64)) <= 0)
    assert np.array_equal(_sort(clops, ctx, queue, s, x), want)


@pytest.mark.parametrize("algo", ["sbitonic", "abitonic", "gselect"])
@pytest.mark.parametrize("dtype", [np.float16, np.float32, np.float64])
@pytest.mark.parametrize("desc", [False, True])
def test_floating_point_elements(clops, oracle, ctx, queue, algo, dtype, desc):
    """Floating-point elements with the default and the descending compare;
    +0.0 and -0.0 compare equal, so their order is the network's (or the
    rank formula's) tie order."""
    n = 4096
    rs = np.random.default_rng(7)
    x = (rs.standard_normal(n) * 100).astype(dtype)
    x[rs.integers(0, n, 200)] = dtype(0.0)
    x[rs.integers(0, n, 200)] = dtype(-0.0)
    x[rs.integers(0, n, 50)] = dtype(1.5)
    cmp = np.less if desc else np.greater
    s = clops.Sorter(ctx, algo, dtype, compare="((a) < (b))" if desc else None)
    if algo == "gselect":
        want = oracle.gselect_rank(x, compare=cmp)
    else:
        want = oracle.bitonic_network(x, compare=cmp)
    got = _sort(clops, ctx, queue, s, x)
    assert np.array_equal(got.view(f"u{x.itemsize}"), want.view(f"u{x.itemsize}"))


@pytest.mark.parametrize("algo", ["sbitonic", "abitonic"])
def test_network_with_nan(clops, oracle, ctx, queue, algo):
    """NaN compares false both ways; the network is still a fixed sequence
    of comparators, so the output is the reference's, bit for bit."""
    n = 1 << 14
    rs = np.random.default_rng(11)
    x = rs.standard_normal(n).astype(np.float32)
    x[rs.integers(0, n, 300)] = np.nan
    s = clops.Sorter(ctx, algo, np.float32)
    want = oracle.bitonic_network(x)
    got = _sort(clops, ctx, queue, s, x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("n", [1, 2, 1000, 4097, 20000])
def test_gselect_key_snippet_kv(clops, oracle, ctx, queue, n):
    """gselect with a key getter and payloads: the rank formula, i.e. the
    stable sort by the low byte, descending."""
    rs = np.random.default_rng(n)
    x = rs.integers(0, 1 << 63, n, dtype=np.int64).astype(U64)
    v = rs.integers(0, 1 << 32, n, dtype=np.int64).astype(np.uint32)
    s = clops.Sorter(ctx, "gselect", np.uint64, key_dtype=np.uint8, get_key="((uchar)(x))",
                     compare="((a) < (b))")
    lowbyte = lambda a: (a & U64(0xFF)).astype(np.uint8)  # noqa: E731
    wk, wv = oracle.gselect_rank(x, v, key=lowbyte, compare=np.less)
    kb, vb = clops.Buffer.from_array(ctx, x), clops.Buffer.from_array(ctx, v)
    ko, vo = clops.Buffer(ctx, max(8 * n, 8)), clops.Buffer(ctx, max(4 * n, 4))
    s.sort_kv_device(queue, kb, vb, ko, vo, n)
    queue.finish()
    assert np.array_equal(ko.read(queue, U64, n), wk)
    assert np.array_equal(vo.read(queue, np.uint32, n), wv)
    # in place, through the host path
    hk, hv = s.sort_kv_host(x, v, queue)
    assert np.array_equal(hk, wk) and np.array_equal(hv, wv)


def _float_key(x):
    """Order-preserving u32 image of float32 bits (the snippet below)."""
    b = x.view(np.uint32)
    m = (x.view(np.int32) >> 31).astype(np.uint32) | np.uint32(0x80000000)
    return b ^ m


FLOAT_KEY = "(as_uint(x) ^ ((uint)(as_int(x) >> 31) | 0x80000000u))"


@pytest.mark.parametrize("n", [1, 2, 1000, 65537, 1 << 22])
@pytest.mark.parametrize("radix", [256, 16])
def test_satradix_float_keys(clops, ctx, queue, n, radix):
    """satradix sorts float32 elements through an integer key snippet: the
    stable order of the key bits (numpy's stable argsort of the same map)."""
    rs = np.random.default_rng(n + radix)
    x = (rs.standard_normal(n) * 1e3).astype(np.float32)
    if n > 10:
        x[rs.integers(0, n, n // 10)] = 0.0
        x[rs.integers(0, n, n // 10)] = -0.0
        x[rs.integers(0, n, 5)] = np.nan
        x[rs.integers(0, n, 5)] = -np.inf
    s = clops.Sorter(ctx, "satradix", np.float32, options="radix=%d" % radix,
                     key_dtype=np.uint32, get_key=FLOAT_KEY)
    want = x[np.argsort(_float_key(x), kind="stable")]
    got = _sort(clops, ctx, queue, s, x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    got = _sort(clops, ctx, queue, s, x, out_of_place=False)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("n", [5, 4096, 300001])
def test_satradix_narrow_key_kv_stable(clops, ctx, queue, n):
    """u64 elements keyed by their top 16 bits (a ushort key, shift given as
    a compiler option): stable, payload carried; the compare snippet is
    ignored, as the reference's radix kernels never use it."""
    rs = np.random.default_rng(n)
    x = rs.integers(0, 1 << 63, n, dtype=np.int64).astype(U64) << U64(1)
    x[: n // 2] &= U64(0x0003FFFFFFFFFFFF)  # many equal keys
    v = np.arange(n, dtype=np.uint32)
    s = clops.Sorter(ctx, "satradix", np.uint64, key_dtype=np.uint16,
                     get_key="((ushort)((x) >> SHIFT))", compare="((a) < (b))",
                     compiler_opts="-DSHIFT=48 -cl-fast-relaxed-math")
    order = np.argsort((x >> U64(48)).astype(np.uint16), kind="stable")
    for oop in (True, False):  # keys only: index sort and gather
        assert np.array_equal(_sort(clops, ctx, queue, s, x, out_of_place=oop), x[order])
    hk, hv = s.sort_kv_host(x, v, queue)
    assert np.array_equal(hk, x[order]) and np.array_equal(hv, v[order])
    kb, vb = clops.Buffer.from_array(ctx, x), clops.Buffer.from_array(ctx, v)
    ko, vo = clops.Buffer(ctx, 8 * n), clops.Buffer(ctx, 4 * n)
    s.sort_kv_device(queue, kb, vb, ko, vo, n)
    queue.finish()
    assert np.array_equal(ko.read(queue, U64, n), x[order])
    assert np.array_equal(vo.read(queue, np.uint32, n), v[order])


@pytest.mark.parametrize("n", [2, 777, 100003])
def test_satradix_narrow_elements(clops, ctx, queue, n):
    """Elements of 1 and 2 bytes (char, half) keyed through snippets: the
    element rides as the payload of the key sort and is narrowed back."""
    rs = np.random.default_rng(n)
    h = (rs.standard_normal(n) * 50).astype(np.float16)
    s = clops.Sorter(ctx, "satradix", np.float16, key_dtype=np.uint16,
                     get_key="((ushort)(as_ushort(x) ^ ((as_short(x) >> 15) | 0x8000)))")
    m = (h.view(np.int16) >> 15).astype(np.uint16) | np.uint16(0x8000)
    want = h[np.argsort(h.view(np.uint16) ^ m, kind="stable")]
    for oop in (True, False):
        got = _sort(clops, ctx, queue, s, h, out_of_place=oop)
        assert np.array_equal(got.view(np.uint16), want.view(np.uint16))
    c = rs.integers(-128, 128, n).astype(np.int8)
    s = clops.Sorter(ctx, "satradix", np.int8, key_dtype=np.uint8,
                     get_key="((uchar)(((uchar)(x)) >> 4))")  # high nibble, unsigned
    want = c[np.argsort(c.view(np.uint8) >> 4, kind="stable")]
    for oop in (True, False):
        assert np.array_equal(_sort(clops, ctx, queue, s, c, out_of_place=oop), want)


def test_same_snippets_twice_and_source(clops, ctx, queue):
    a = clops.Sorter(ctx, "sbitonic", np.uint64, key_dtype=np.uint32,
                     get_key="((uint)((x) >> 32))")
    b = clops.Sorter(ctx, "gselect", np.uint64, key_dtype=np.uint32,
                     get_key="((uint)((x) >> 32))")
    src = a.source()
    assert "#define CLO_SORT_KEY_GET(x) ((uint)((x) >> 32))" in src
    assert "#define CLO_SORT_ELEM_TYPE ulong" in src and "clo_snip_gselect" in src
    assert b.source() == src
    assert clops.Sorter(ctx, "satradix", np.uint32).source() is None
    assert clops.Sorter(ctx, "sbitonic", np.int32, compare="((a) < (b))").source() is None


def test_snippet_errors(clops, ctx, queue):
    with pytest.raises(clops.ClopsError) as e:
        clops.Sorter(ctx, "gselect", np.uint32, compare="((a).x > (b).x)")
    assert e.value.code == clops.CLO_ERROR_ARGS and "did not compile" in str(e.value)
    with pytest.raises(clops.ClopsError) as e:  # radix digits need an integer key
        clops.Sorter(ctx, "satradix", np.float32)
    assert e.value.code == clops.CLO_ERROR_ARGS
    s = clops.Sorter(ctx, "sbitonic", np.uint64, key_dtype=np.uint32,
                     get_key="((uint)(x))")
    with pytest.raises(clops.ClopsError) as e:  # not stable: no key-value sort
        s.sort_kv_host(np.arange(4, dtype=U64), np.arange(4, dtype=np.uint32), queue)
    assert e.value.code == clops.CLO_ERROR_ARGS
    r = clops.Sorter(ctx, "satradix", np.uint64, key_dtype=np.uint32, get_key="((uint)(x))")
    x = clops.Buffer.from_array(ctx, np.arange(8, dtype=U64))
    offs = clops.Buffer.from_array(ctx, np.array([0, 4], dtype=np.uint32))
    with pytest.raises(clops.ClopsError) as e:
        r.sort_segmented_device(queue, x, offs, 2, 8)
    assert e.value.code == clops.CLO_ERROR_ARGS
