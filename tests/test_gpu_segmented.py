"""Batched segmented sort and scan on the GPU vs the oracle applied per segment."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lengths(oracle, nseg, max_len, seed):
    # log-uniform lengths in [1, max_len] (the distribution of the 32768-segment benchmark workload)
    u = oracle.generate(oracle.GEN_U64, seed, nseg)
    x = (u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.clip(np.floor(np.exp(x * np.log(max_len))), 1, max_len).astype(np.uint32)


def _offsets(lens):
    off = np.zeros(lens.size, dtype=np.uint32)
    off[1:] = np.cumsum(lens[:-1], dtype=np.uint64).astype(np.uint32)
    return off


def _ref_sort(oracle, k, v, off, n):
    rk, rv = k.copy(), None if v is None else v.copy()
    ends = np.append(off[1:], n)
    for s, e in zip(off, ends):
        if e - s > 1:
            if v is None:
                rk[s:e] = oracle.stable_sort(k[s:e])
            else:
                rk[s:e], rv[s:e] = oracle.stable_sort(k[s:e], v[s:e])
    return rk, rv


CASES = [
    # (nseg, max_len) covering chunks, medium and large segments
    (1, 100000), (7, 3), (300, 64), (500, 5000), (200, 70000), (2000, 20000),
]


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
@pytest.mark.parametrize("kv", [False, True])
@pytest.mark.parametrize("nseg,max_len", CASES)
def test_segmented_sort(clops, oracle, ctx, queue, dtype, kv, nseg, max_len):
    lens = _lengths(oracle, nseg, max_len, 0xC4A + nseg)
    off = _offsets(lens)
    n = int(lens.sum(dtype=np.uint64))
    k = oracle.generate(oracle.GEN_U64, 0xC4B, n).astype(dtype)
    v = np.arange(n, dtype=np.uint32) if kv else None
    s = clops.Sorter(ctx, "satradix", dtype)
    kb = clops.Buffer.from_array(ctx, k)
    vb = clops.Buffer.from_array(ctx, v) if kv else None
    ob = clops.Buffer.from_array(ctx, off)
    s.sort_segmented_device(queue, kb, ob, nseg, n, vin=vb)
    queue.finish()
    rk, rv = _ref_sort(oracle, k, v, off, n)
    assert np.array_equal(kb.read(queue, dtype, n), rk)
    if kv:
        assert np.array_equal(vb.read(queue, np.uint32, n), rv)


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_segmented_sort_tiles_out_of_place(clops, oracle, ctx, queue, tile):
    nseg = 3000
    lens = _lengths(oracle, nseg, 30000, 0x51)
    lens[::7] = 0  # empty segments
    off = _offsets(lens)
    n = int(lens.sum(dtype=np.uint64))
    k = (oracle.generate(oracle.GEN_U64, 0x52, n) % np.uint64(97)).astype(np.uint32)
    v = np.arange(n, dtype=np.uint32)
    s = clops.Sorter(ctx, "satradix", np.uint32, options="segtile=%d,seggrid=300" % tile)
    kb, vb = clops.Buffer.from_array(ctx, k), clops.Buffer.from_array(ctx, v)
    ko, vo = clops.Buffer(ctx, n * 4), clops.Buffer(ctx, n * 4)
    ob = clops.Buffer.from_array(ctx, off)
    s.sort_segmented_device(queue, kb, ob, nseg, n, vin=vb, kout=ko, vout=vo)
    queue.finish()
    rk, rv = _ref_sort(oracle, k, v, off, n)
    assert np.array_equal(ko.read(queue, np.uint32, n), rk)
    assert np.array_equal(vo.read(queue, np.uint32, n), rv)
    assert np.array_equal(kb.read(queue, np.uint32, n), k)  # input untouched


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("dtype,kv", [(np.uint32, False), (np.uint32, True), (np.uint64, True)])
def test_segmented_sort_class_boundaries(clops, oracle, ctx, queue, tile, dtype, kv):
    """Segment lengths on both sides of each work-class boundary: chunked
    small segments (<= W), one-tile medium segments (<= CAP) and multi-chunk
    large ones, with CAP = 256 * tile items and W = CAP / 2; many equal keys,
    so stability is checked through the payloads."""
    if dtype == np.uint64 and tile == 32:
        pytest.skip("64-bit keys take tiles of 8 or 16 items")
    cap = 256 * tile
    w = cap // 2
    edges = [0, 1, 2, w - 1, w, w + 1, cap - 1, cap, cap + 1, 2 * cap - 1, 2 * cap,
             2 * cap + 1, 3 * cap + 7]
    rs = np.random.default_rng(tile)
    lens = []
    for e in edges:  # each boundary length between short random segments
        lens += [int(rs.integers(0, 40)), e, int(rs.integers(0, 40))]
    lens = np.array(lens, dtype=np.uint32)
    off = _offsets(lens)
    nseg, n = lens.size, int(lens.sum(dtype=np.uint64))
    k = (oracle.generate(oracle.GEN_U64, 0x5E + tile, n) % np.uint64(1000)).astype(dtype)
    v = np.arange(n, dtype=np.uint32) if kv else None
    s = clops.Sorter(ctx, "satradix", dtype, options="segtile=%d" % tile)
    kb = clops.Buffer.from_array(ctx, k)
    vb = clops.Buffer.from_array(ctx, v) if kv else None
    ob = clops.Buffer.from_array(ctx, off)
    s.sort_segmented_device(queue, kb, ob, nseg, n, vin=vb)
    queue.finish()
    rk, rv = _ref_sort(oracle, k, v, off, n)
    assert np.array_equal(kb.read(queue, dtype, n), rk)
    if kv:
        assert np.array_equal(vb.read(queue, np.uint32, n), rv)


def _ref_segscan(oracle, a, off, n, sumt):
    out = np.empty(n, dtype=sumt)
    starts = np.append(off, n)
    if off.size and off[0] > 0:
        out[:off[0]] = oracle.scan(a[:off[0]], sumt)
    for s, e in zip(starts[:-1], starts[1:]):
        out[s:e] = oracle.scan(a[s:e], sumt)
    return out


@pytest.mark.parametrize("elem,sumt", [(np.uint32, np.uint32), (np.uint32, np.uint64),
                                       (np.int16, np.int64), (np.uint64, np.uint64)])
@pytest.mark.parametrize("nseg,max_len", [(1, 50000), (1000, 10), (3000, 2000), (100, 70000)])
def test_segmented_scan(clops, oracle, ctx, queue, elem, sumt, nseg, max_len):
    lens = _lengths(oracle, nseg, max_len, 0xC4C + nseg)
    lens[3::11] = 0
    off = _offsets(lens)
    n = int(lens.sum(dtype=np.uint64))
    a = oracle.generate(oracle.GEN_U64, 0x5E, n).astype(
        np.dtype(elem).str.replace("i", "u")).view(elem)
    sc = clops.Scanner(ctx, elem, sumt)
    bi = clops.Buffer.from_array(ctx, a)
    bo = clops.Buffer(ctx, max(n, 1) * np.dtype(sumt).itemsize)
    ob = clops.Buffer.from_array(ctx, off)
    sc.scan_segmented_device(queue, bi, bo, ob, nseg, n)
    queue.finish()
    assert np.array_equal(bo.read(queue, sumt, n), _ref_segscan(oracle, a, off, n, sumt))


def _mix64(x):
    z = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_config4_shape(clops, oracle, ctx, queue):
    """32768 log-uniform segments in [1, 65536]: device checks + a bit-exact sample."""
    nseg = 32768
    lens = _lengths(oracle, nseg, 65536, 0xC4A)
    off = _offsets(lens)
    n = int(lens.sum(dtype=np.uint64))
    kb = clops.Buffer(ctx, n * 4)
    clops.generate(queue, kb, n, clops.GEN_U32_HI, 0xC4B)
    ob = clops.Buffer.from_array(ctx, off)
    h0 = clops.multiset_hash(queue, kb, n, 4)
    s = clops.Sorter(ctx, "satradix", np.uint32)
    s.sort_segmented_device(queue, kb, ob, nseg, n)
    queue.finish()
    assert clops.multiset_hash(queue, kb, n, 4) == h0
    got = kb.read(queue, np.uint32, n)
    k = oracle.generate(oracle.GEN_U32_HI, 0xC4B, n)
    ends = np.append(off[1:], n)
    # Every segment: non-decreasing inside, and the same multiset as its input
    # (per-segment sums of the splitmix64 finalizer of each key, mod 2^64).
    inner = np.ones(max(n - 1, 0), dtype=bool)
    inner[off[1:][off[1:] > 0] - 1] = False
    assert not np.any((got[1:] < got[:-1]) & inner)
    nz = lens > 0
    with np.errstate(over="ignore"):
        hs_in = np.add.reduceat(_mix64(k), off[nz])
        hs_out = np.add.reduceat(_mix64(got), off[nz])
    assert np.array_equal(hs_in, hs_out)
    for j in list(range(0, nseg, 97)) + [int(np.argmax(lens))]:
        s0, e0 = int(off[j]), int(ends[j])
        assert np.array_equal(got[s0:e0], np.sort(k[s0:e0])), j
    sc = clops.Scanner(ctx, np.uint32, np.uint32)
    so = clops.Buffer(ctx, n * 4)
    sc.scan_segmented_device(queue, kb, so, ob, nseg, n)
    queue.finish()
    res = so.read(queue, np.uint32, n)
    for j in range(0, nseg, 89):
        s0, e0 = int(off[j]), int(ends[j])
        assert np.array_equal(res[s0:e0], oracle.scan(got[s0:e0], np.uint32)), j


def _keys_for(oracle, dist, n, dtype, seed):
    """Keys-only distributions for the unstable MSD bucketing paths: spread,
    constant, two-valued, clustered with outliers (large buckets that need
    further levels or the digit-pass fallback), many duplicates, heavy head,
    ordered and reversed."""
    bits = np.dtype(dtype).itemsize * 8
    u = oracle.generate(oracle.GEN_U64, seed, n)
    top = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    if dist == "uniform":
        k = u
    elif dist == "constant":
        k = np.full(n, 0x1234567, dtype=np.uint64)
    elif dist == "two":
        k = np.where(u & np.uint64(1), top, np.uint64(0))
    elif dist == "clustered":
        # 99% of keys in [2^20, 2^20 + 4096), the rest anywhere
        k = np.where(u % np.uint64(100) == 0, u, np.uint64(1 << 20) + (u >> np.uint64(40)) % np.uint64(4096))
    elif dist == "dups":
        k = u % np.uint64(97)
    elif dist == "head":
        # geometric-like: key = number of trailing zero bits of u (0..63)
        k = np.zeros(n, dtype=np.uint64)
        x = u.copy()
        for _ in range(63):
            m = (x & np.uint64(1)) == 0
            k += m.astype(np.uint64)
            x = np.where(m, x >> np.uint64(1), np.uint64(1))
    elif dist == "sorted":
        k = np.arange(n, dtype=np.uint64) * np.uint64(3)
    elif dist == "reversed":
        k = (np.uint64(n) - np.arange(n, dtype=np.uint64)) << np.uint64(bits - 24 if bits > 24 else 0)
    else:
        raise ValueError(dist)
    return (k & top).astype(dtype)


@pytest.mark.parametrize("tile", [8, 16, 32])
@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
@pytest.mark.parametrize("dist", ["uniform", "constant", "two", "clustered", "dups", "head",
                                  "sorted", "reversed"])
@pytest.mark.parametrize("out_of_place", [False, True])
@pytest.mark.parametrize("msd", [0, 1])
def test_segmented_keys_only_distributions(clops, oracle, ctx, queue, tile, dtype, dist,
                                           out_of_place, msd):
    """Keys alone, with the MSD bucketing kernel for medium and large segments
    (segmsd=1; large segments whose keys crowd into part of their range go on
    to the digit passes) or the digit passes only (segmsd=0): every
    distribution against the oracle's per-segment sort."""
    if dtype == np.uint64 and tile == 32:
        pytest.skip("64-bit keys take tiles of 8 or 16 items")
    cap = 256 * tile
    lens = np.array([0, 1, 5, cap // 2 + 3, cap - 1, cap, cap + 1, 3 * cap + 5, 70000, 2,
                     cap * 2, 17, 40000], dtype=np.uint32)
    off = _offsets(lens)
    nseg, n = lens.size, int(lens.sum(dtype=np.uint64))
    k = _keys_for(oracle, dist, n, dtype, 0xD15 + tile)
    s = clops.Sorter(ctx, "satradix", dtype, options="segtile=%d,segmsd=%d" % (tile, msd))
    kb = clops.Buffer.from_array(ctx, k)
    ob = clops.Buffer.from_array(ctx, off)
    ko = clops.Buffer(ctx, n * np.dtype(dtype).itemsize) if out_of_place else None
    s.sort_segmented_device(queue, kb, ob, nseg, n, kout=ko)
    queue.finish()
    rk, _ = _ref_sort(oracle, k, None, off, n)
    got = (ko if out_of_place else kb).read(queue, dtype, n)
    assert np.array_equal(got, rk)
    if out_of_place:
        assert np.array_equal(kb.read(queue, dtype, n), k)  # input untouched


@pytest.mark.parametrize("tile", [8, 16])
@pytest.mark.parametrize("dist", ["uniform", "constant", "two", "clustered", "dups", "head",
                                  "sorted", "reversed"])
@pytest.mark.parametrize("msd", [0, 1])
def test_segmented_kv_out_of_place(clops, oracle, ctx, queue, tile, dist, msd):
    """Key-value, 32-bit keys, out of place: with segmsd=1 the bucketing kernel
    sorts 64-bit composites (key, position in the segment), and segments it
    cannot bucket go to the key-value digit passes in the output; with
    segmsd=0 the digit passes take everything.  Stable order (equal keys keep
    their input order) and the values that travel with the keys, against the
    oracle's per-segment stable sort; inputs untouched."""
    cap = 256 * tile
    lens = np.array([0, 1, 5, cap // 2 + 3, cap - 1, cap, cap + 1, 3 * cap + 5, 70000, 2,
                     cap * 2, 17, 40000, 300000], dtype=np.uint32)
    off = _offsets(lens)
    nseg, n = lens.size, int(lens.sum(dtype=np.uint64))
    k = _keys_for(oracle, dist, n, np.uint32, 0xE15 + tile)
    v = (oracle.generate(oracle.GEN_U64, 0xE16, n) >> np.uint64(32)).astype(np.uint32)
    s = clops.Sorter(ctx, "satradix", np.uint32, options="segtile=%d,segmsd=%d" % (tile, msd))
    kb = clops.Buffer.from_array(ctx, k)
    vb = clops.Buffer.from_array(ctx, v)
    ob = clops.Buffer.from_array(ctx, off)
    ko = clops.Buffer(ctx, n * 4)
    vo = clops.Buffer(ctx, n * 4)
    s.sort_segmented_device(queue, kb, ob, nseg, n, vin=vb, kout=ko, vout=vo)
    queue.finish()
    rk, rv = _ref_sort(oracle, k, v, off, n)
    assert np.array_equal(ko.read(queue, np.uint32, n), rk)
    assert np.array_equal(vo.read(queue, np.uint32, n), rv)
    assert np.array_equal(kb.read(queue, np.uint32, n), k)
    assert np.array_equal(vb.read(queue, np.uint32, n), v)
