"""Pins the CPU oracle (CPU only).

The reference's own tests hold no value checks (src/tests/test_rng.c asserts
only "no error"), so every oracle function is pinned here by something it
does not share code with: published known answers, numpy's independent
implementations, exhaustive enumeration on tiny inputs, and algebraic
properties of the generators (periods, skip-ahead identities).
"""
import itertools
import math

import numpy as np
import pytest

M64 = (1 << 64) - 1


# --------------------------------------------------------------------------
# Seeded input generator (splitmix64)


def test_splitmix64_known_outputs(oracle):
    # splitmix64 from state 0 (Vigna's reference C code): the first outputs.
    u = oracle.generate(oracle.GEN_U64, 0, 3)
    assert [int(x) for x in u] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_generator_kinds(oracle):
    u = oracle.generate(oracle.GEN_U64, 77, 1000, start=5)
    assert np.array_equal(oracle.generate(oracle.GEN_U32_HI, 77, 1000, start=5),
                          (u >> np.uint64(32)).astype(np.uint32))
    assert np.array_equal(oracle.generate(oracle.GEN_U64_MASK, 77, 1000, param=0xFFFFF, start=5),
                          u & np.uint64(0xFFFFF))
    assert np.array_equal(oracle.generate(oracle.GEN_U32_MOD, 77, 1000, param=1000, start=5),
                          (u >> np.uint64(32)).astype(np.uint32) % np.uint32(1000))
    assert np.array_equal(oracle.generate(oracle.GEN_IOTA_U32, 0, 10, start=3),
                          np.arange(3, 13, dtype=np.uint32))
    # Presorted with 1% random swaps: over whole groups of 1024 blocks of 100
    # positions the source positions form a permutation made of disjoint
    # transpositions (an involution), one moved position per block, each
    # pairing block b with block b ^ m for one m in [1, 1023].
    n = 2 * 102400
    ns = oracle.generate(oracle.GEN_NEARLY_SORTED_U64, 77, n, param=1000)
    assert np.all(ns % np.uint64(1000) == 0)
    s = (ns // np.uint64(1000)).astype(np.int64)
    idx = np.arange(n, dtype=np.int64)
    assert np.array_equal(np.sort(s), idx)
    assert np.array_equal(s[s], idx)
    moved = np.nonzero(s != idx)[0]
    assert moved.size == n // 100
    assert np.array_equal(np.bincount(moved // 100, minlength=n // 100), np.ones(n // 100))
    m = np.unique((moved // 100) ^ (s[moved] // 100))
    assert m.size == 1 and 1 <= int(m[0]) <= 1023
    assert np.all(np.diff(np.delete(s, moved)) > 0)
    ns32 = oracle.generate(oracle.GEN_NEARLY_SORTED_U32, 77, n, param=1000)
    assert np.array_equal(ns32, (ns & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    # A prefix and a chunk from `start` agree with the whole sequence.
    assert np.array_equal(oracle.generate(oracle.GEN_NEARLY_SORTED_U64, 77, 777, 1000, start=5000),
                          ns[5000:5777])
    idx = np.arange(n, dtype=np.uint64)
    rv = oracle.generate(oracle.GEN_REVERSED_U32, 0, n, param=7)
    assert np.array_equal(rv, (0xFFFFFFFF - 7 * idx).astype(np.uint32))
    assert np.all(np.diff(rv.astype(np.int64)) < 0)
    rv64 = oracle.generate(oracle.GEN_REVERSED_U64, 0, n, param=3)
    assert rv64[0] == np.uint64(2 ** 64 - 1) and np.all(rv64[:-1] > rv64[1:])


def test_zipf_distribution(oracle):
    # H(2^16, 1.1) by direct summation; the top rank takes 1/H of the mass.
    ranks, s = 1 << 16, 1.1
    H = math.fsum((r + 1) ** -s for r in range(ranks))
    cdf = oracle.zipf_cdf(ranks, s)
    assert abs(cdf[0] - 1.0 / H) < 1e-12 and cdf[-1] == 1.0 and np.all(np.diff(cdf) > 0)
    k = oracle.generate_zipf(cdf, 0xC3C, 200000)
    top = oracle.mix64(0)
    frac = float(np.mean(k == np.uint64(top)))
    assert abs(frac - 1.0 / H) < 0.01


# --------------------------------------------------------------------------
# Sorts


def _stable_ref(k, v):
    idx = np.argsort(k, kind="stable")
    return k[idx], v[idx]


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("n", [0, 1, 2, 17, 256, 257, 5000])
def test_stable_sort_matches_numpy(oracle, dtype, n):
    k = oracle.generate(oracle.GEN_U64, n + 1, n).astype(dtype)
    k[: n // 3] = k[0] if n else k[: n // 3]
    v = np.arange(n, dtype=np.uint32)
    rk, rv = oracle.stable_sort(k, v)
    ek, ev = _stable_ref(k, v)
    assert np.array_equal(rk, ek) and np.array_equal(rv, ev)


@pytest.mark.parametrize("bits", [1, 2, 3, 4, 5, 7, 8])
@pytest.mark.parametrize("lws", [16, 256, 1000])
@pytest.mark.parametrize("dtype", [np.uint16, np.uint32, np.uint64])
def test_satradix_restatement(oracle, bits, lws, dtype):
    n = 3001
    k = (oracle.generate(oracle.GEN_U64, bits * 7 + lws, n) % np.uint64(5000)).astype(dtype)
    v = np.arange(n, dtype=np.uint32)
    rk, rv = oracle.satradix(k, v, bits=bits, lws=lws, threads=3)
    ek, ev = _stable_ref(k, v)
    assert np.array_equal(rk, ek) and np.array_equal(rv, ev)


def test_sbitonic_exhaustive_tiny(oracle):
    # every permutation (with repeats) of small power-of-two arrays
    for n in (2, 4, 8):
        vals = list(range(n // 2)) * 2
        for perm in set(itertools.permutations(vals)):
            a = np.array(perm, dtype=np.int32)
            assert np.array_equal(oracle.sbitonic(a), np.sort(a))
            assert np.array_equal(oracle.sbitonic(a, desc=True), np.sort(a)[::-1])


@pytest.mark.parametrize("dtype", [np.int8, np.uint16, np.int32, np.uint64, np.int64])
def test_sbitonic_po2(oracle, dtype):
    n = 1 << 12
    k = oracle.generate(oracle.GEN_U64, 3, n).astype(np.dtype(dtype).str.replace("i", "u")).view(
        dtype)
    assert np.array_equal(oracle.sbitonic(k), np.sort(k))
    assert np.array_equal(oracle.sbitonic(k, desc=True), np.sort(k)[::-1])


@pytest.mark.parametrize("dtype", [np.int8, np.uint32, np.int64])
@pytest.mark.parametrize("desc", [False, True])
def test_gselect_restatement(oracle, dtype, desc):
    n = 700
    k = (oracle.generate(oracle.GEN_U64, 9, n) % np.uint64(40)).astype(np.int64) - 20
    k = k.astype(dtype)
    v = np.arange(n, dtype=np.uint32)
    rk, rv = oracle.gselect(k, v, desc=desc)
    idx = np.argsort(-k.astype(np.float64) if desc else k, kind="stable")
    assert np.array_equal(rk, k[idx]) and np.array_equal(rv, v[idx])


# --------------------------------------------------------------------------
# Scan


@pytest.mark.parametrize("elem", [np.int8, np.uint8, np.int16, np.uint32, np.int64, np.uint64])
@pytest.mark.parametrize("sumt", [np.uint8, np.int16, np.uint32, np.int32, np.uint64, np.int64])
def test_scan_matches_numpy(oracle, elem, sumt):
    n = 2000
    a = oracle.generate(oracle.GEN_U64, 4, n).astype(np.dtype(elem).str.replace("i", "u")).view(
        elem)
    want = np.zeros(n, dtype=sumt)
    # numpy converts as C does (sign extension), and cumsum wraps
    want[1:] = np.cumsum(a[:-1].astype(sumt), dtype=sumt)
    assert np.array_equal(oracle.scan(a, sumt), want)


@pytest.mark.parametrize("n", [1, 511, 512, 4096, 100000])
@pytest.mark.parametrize("lws", [4, 64, 256])
def test_blelloch_restatement(oracle, n, lws):
    a = oracle.generate(oracle.GEN_U32_HI, 11, n)
    for st in (np.uint32, np.uint64):
        assert np.array_equal(oracle.scan_blelloch(a, st, lws=lws, threads=4), oracle.scan(a, st))


# --------------------------------------------------------------------------
# RNG: known answers


def test_parkmiller_known_answer(oracle):
    # Park & Miller (CACM 31(10), 1988): from seed 1, the 10000th state is 1043618065.
    st = oracle.rng_from_seed("parkmiller", 1)
    for _ in range(10000):
        out = oracle.rng_next("parkmiller", st)
    assert int(st.view(np.int32)[0]) == 1043618065
    assert out == (1043618065 << 1) & 0xFFFFFFFF


def test_parkmiller_negative_states_truncate(oracle):
    # HOST_MT seeds make negative int states; OpenCL's % truncates toward zero.
    for s0 in (-1, -2, -123456789, -(1 << 31), 2147483646, 0, 2147483647):
        st = oracle.rng_from_seed("parkmiller", s0 & 0xFFFFFFFF)
        s = s0
        for _ in range(50):
            r = oracle.rng_next("parkmiller", st)
            s = int(math.fmod(s * 16807, 2147483647))  # exact: |s*16807| < 2^53
            assert int(st.view(np.int32)[0]) == s
            assert r == (s & 0xFFFFFFFF) << 1 & 0xFFFFFFFF


def test_lcg_java_random(oracle):
    # java.util.Random(0): scrambled seed 0 ^ 0x5DEECE66D; nextInt() = -1155484576.
    st = oracle.rng_from_seed("lcg", 0x5DEECE66D)
    assert oracle.rng_next("lcg", st) == (-1155484576) & 0xFFFFFFFF
    # java.util.Random(42).nextInt() = -1170105035
    st = oracle.rng_from_seed("lcg", 42 ^ 0x5DEECE66D)
    assert oracle.rng_next("lcg", st) == (-1170105035) & 0xFFFFFFFF


def test_lcg_skip_ahead(oracle):
    a, c, m = 0x5DEECE66D, 0xB, 1 << 48
    st = oracle.rng_from_seed("lcg", 12345)
    for _ in range(1000):
        oracle.rng_next("lcg", st)
    # x_n = a^n x_0 + c (a^n - 1)/(a - 1)  (mod m), computed with big ints
    an = pow(a, 1000, m * (a - 1))
    xn = (an * 12345 + c * ((an - 1) // (a - 1))) % m
    assert int(st.view(np.uint64)[0]) == xn


def test_xorshift128_marsaglia(oracle):
    # Marsaglia (2003) xor128 with x,y,z,w = 123456789, 362436069, 521288629, 88675123
    st = np.array([123456789, 362436069, 521288629, 88675123], dtype=np.uint32).view(np.uint8)
    got = [oracle.rng_next("xorshift128", st) for _ in range(3)]
    assert got == [3701687786, 458299110, 2500872618]


def test_mwc64x_skip_ahead(oracle):
    # MWC: S' = A*x + c with S = c*2^32 + x, hence S_n = A^n S_0 mod (A*2^32 - 1).
    A = 4294883355
    M = A * (1 << 32) - 1
    seed = 0x0123456789ABCDEF % M
    st = oracle.rng_from_seed("mwc64x", seed)
    outs = []
    for _ in range(777):
        outs.append(oracle.rng_next("mwc64x", st))
    x, c = (int(w) for w in st.view(np.uint32))
    assert c * (1 << 32) + x == pow(A, 777, M) * seed % M
    s1 = A * seed % M
    assert outs[1] == (s1 & 0xFFFFFFFF) ^ (s1 >> 32)


def _gf2_matrix(step, bits):
    cols = []
    for i in range(bits):
        cols.append(step(1 << i))
    return cols  # column i = image of basis vector i


def _gf2_apply(cols, v):
    r = 0
    i = 0
    while v:
        if v & 1:
            r ^= cols[i]
        v >>= 1
        i += 1
    return r


def _gf2_mul(A, B):
    return [_gf2_apply(A, b) for b in B]


def _gf2_pow(A, e, bits):
    R = [1 << i for i in range(bits)]
    while e:
        if e & 1:
            R = _gf2_mul(A, R)
        A = _gf2_mul(A, A)
        e >>= 1
    return R


def _full_period(cols, bits, factors):
    ident = [1 << i for i in range(bits)]
    order = (1 << bits) - 1
    if _gf2_pow(cols, order, bits) != ident:
        return False
    return all(_gf2_pow(cols, order // p, bits) != ident for p in factors)


def test_xorshift64_full_period(oracle):
    # The (21, 35, 4) triple has period 2^64 - 1 (Marsaglia 2003); a wrong
    # shift or direction would not.
    def step(v):
        st = np.array([v], dtype=np.uint64).view(np.uint8)
        oracle.rng_next("xorshift64", st)
        return int(st.view(np.uint64)[0])

    cols = _gf2_matrix(step, 64)
    f64 = [3, 5, 17, 257, 641, 65537, 6700417]
    assert math.prod(f64) == (1 << 64) - 1
    assert _full_period(cols, 64, f64)
    # linearity of the oracle step
    a, b = 0x123456789ABCDEF0, 0x0FEDCBA987654321
    assert step(a ^ b) == step(a) ^ step(b)


def test_xorshift128_full_period(oracle):
    def step(v):
        st = np.array([v & 0xFFFFFFFF, (v >> 32) & 0xFFFFFFFF, (v >> 64) & 0xFFFFFFFF,
                       (v >> 96) & 0xFFFFFFFF], dtype=np.uint32).view(np.uint8)
        oracle.rng_next("xorshift128", st)
        w = [int(x) for x in st.view(np.uint32)]
        return w[0] | w[1] << 32 | w[2] << 64 | w[3] << 96

    cols = _gf2_matrix(step, 128)
    f128 = [3, 5, 17, 257, 641, 65537, 274177, 6700417, 67280421310721]
    assert math.prod(f128) == (1 << 128) - 1
    assert _full_period(cols, 128, f128)


def test_mt19937_known_answer(oracle):
    # C++ [rand.predef]: the 10000th output of mt19937 with default seed 5489.
    assert int(oracle.mt19937(5489, 10000)[-1]) == 4123659995


def _xs64_unstep(s):
    """Inverse of the xorshift64 state update: y = x ^ (x << a) is undone by
    x = y ^ (y << a) ^ (y << 2a) ^ ..., likewise for right shifts; the three
    shifts are undone in reverse order."""
    def unshl(y, a):
        x, t = y, y
        while True:
            t = (t << a) & M64
            if not t:
                return x
            x ^= t

    def unshr(y, a):
        x, t = y, y
        while True:
            t >>= a
            if not t:
                return x
            x ^= t
    return unshl(unshr(unshl(s, 4), 35), 21)


def test_xorshift64_output_word(oracle):
    """The draw is convert_uint of the new state, i.e. its least significant
    32 bits (OpenCL 1.2 6.2.3.3, clo_rng_xorshift64.cl:52).  States are chosen
    through the inverse update so that the new state is a given value: new
    states equal in the low half give equal draws whatever the high half, and
    the draw is that low half.  Taking the high half, or any other window,
    fails."""
    def draw(state):
        st = np.array([state], dtype=np.uint64).view(np.uint8)
        out = oracle.rng_next("xorshift64", st)
        return out, int(st.view(np.uint64)[0])

    rng = np.random.default_rng(0x64)
    for s in [int(x) for x in rng.integers(0, 1 << 63, 16, dtype=np.int64)] + [1, M64]:
        assert draw(_xs64_unstep(s))[1] == s  # the inverse is right
    for lo in (0, 1, 0x80000000, 0xFFFFFFFF, 0x12345678):
        outs = set()
        for hi in (0, 1, 0xFFFFFFFF, 0xDEADBEEF):
            target = (hi << 32) | lo
            out, new = draw(_xs64_unstep(target))
            assert new == target
            outs.add(out)
        assert outs == {lo}, (hex(lo), outs)


def test_seed_conversion_single_bits(oracle):
    """Seeds with one bit set, expected words derived by hand from the
    conversions (clo_rng_xorshift128.cl:30: words = seed >> 0, 16, 32, 46
    masked to 32 bits; clo_rng_tauslcg.cl:35: as_uint4 of (seed, seed), low
    half first).  Any other shift moves a bit out of its expected place."""
    xs128 = {
        0: [1, 0, 0, 0],
        15: [1 << 15, 0, 0, 0],
        16: [1 << 16, 1, 0, 0],
        31: [1 << 31, 1 << 15, 0, 0],
        32: [0, 1 << 16, 1, 0],
        45: [0, 1 << 29, 1 << 13, 0],
        46: [0, 1 << 30, 1 << 14, 1],
        47: [0, 1 << 31, 1 << 15, 2],
        48: [0, 0, 1 << 16, 4],
        63: [0, 0, 1 << 31, 1 << 17],
    }
    for k, want in xs128.items():
        got = [int(x) for x in oracle.rng_from_seed("xorshift128", 1 << k).view(np.uint32)]
        assert got == want, k
    taus = {
        0x0000000100000002: [2, 1, 2, 1],
        1 << 63: [0, 1 << 31, 0, 1 << 31],
        0xFFFFFFFF: [0xFFFFFFFF, 0, 0xFFFFFFFF, 0],
    }
    for s, want in taus.items():
        got = [int(x) for x in oracle.rng_from_seed("tauslcg", s).view(np.uint32)]
        assert got == want, hex(s)


def test_xs1_hash_inverts(oracle):
    """XS1 (clo_rng_init.cl:30-33) on 64-bit values is a bijection made of
    x ^ (x >> 16) and a multiply by 0x45d9f3b mod 2^64.  Undoing those steps,
    x ^ (x >> 16) by x ^ (x >> 16) ^ (x >> 32) ^ (x >> 48) and the multiply by
    the constant's inverse mod 2^64, must give back gid + seed for every state
    the oracle wrote; a different shift or constant does not invert."""
    cinv = pow(0x45D9F3B, -1, 1 << 64)

    def unxs(y):
        return y ^ (y >> 16) ^ (y >> 32) ^ (y >> 48)

    def inverse(h):
        h = unxs(h)
        h = unxs((h * cinv) & M64)
        return unxs((h * cinv) & M64)

    for seed in (0, 1, 0xFFFF, 1 << 40, M64 - 100):
        st = oracle.rng_init_dev_gid("lcg", 64, seed, 2).view(np.uint64)
        for gid in range(64):
            assert inverse(int(st[gid])) == (gid + seed) & M64, (seed, gid)
    # XS1(0) = 0.
    assert int(oracle.rng_init_dev_gid("lcg", 1, 0, 2).view(np.uint64)[0]) == 0


def test_seed_conversions(oracle):
    s = 0x0123456789ABCDEF
    xs = oracle.rng_from_seed("xorshift128", s).view(np.uint32)
    assert [int(x) for x in xs] == [s & M64 & 0xFFFFFFFF, (s >> 16) & 0xFFFFFFFF,
                                    (s >> 32) & 0xFFFFFFFF, (s >> 46) & 0xFFFFFFFF]
    mw = oracle.rng_from_seed("mwc64x", s).view(np.uint32)
    assert [int(x) for x in mw] == [s & 0xFFFFFFFF, s >> 32]
    pm = oracle.rng_from_seed("parkmiller", 0xFFFFFFFF80000001)
    assert int(pm.view(np.int32)[0]) == -(1 << 31) + 1


def test_dev_gid_hashes(oracle):
    count, seed = 64, 1234
    for h, fn in ((0, lambda x: x),
                  (1, lambda x: (x * 2654435761) % (1 << 32)),
                  (2, None)):
        st = oracle.rng_init_dev_gid("lcg", count, seed, h).view(np.uint64)
        for gid in range(count):
            x = gid + seed
            if h == 2:
                x = (((x >> 16) ^ x) * 0x45D9F3B) & M64
                x = (((x >> 16) ^ x) * 0x45D9F3B) & M64
                x = (x >> 16) ^ x
            else:
                x = fn(x) & M64
            assert int(st[gid]) == x


def test_host_mt_layout(oracle):
    words = oracle.mt19937(1234, 8)
    st = oracle.rng_init_host_mt("xorshift128", 2, 1234)
    assert np.array_equal(st.view(np.uint32), words)
    st = oracle.rng_init_host_mt("lcg", 4, (1 << 40) + 1234)  # seed truncated to 32 bits
    assert np.array_equal(st.view(np.uint32), words)


def test_fill_layout(oracle):
    # out[r*G + gid] is the r-th draw of work-item gid
    G, draws = 37, 5
    for gen in oracle.RNG_NAMES:
        st = oracle.rng_init_dev_gid(gen, G, 9, 0)
        per = st.copy().reshape(G, -1)
        want = np.zeros((draws, G), dtype=np.uint32)
        for gid in range(G):
            s = per[gid].copy()
            for r in range(draws):
                want[r, gid] = oracle.rng_next(gen, s) >> 7
        got = oracle.rng_fill(gen, st, draws, bits=25, threads=3)
        assert np.array_equal(got, want.reshape(-1))


def _glib_fixture():
    import json
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "glib_grand.json")
    with open(p) as f:
        return json.load(f)


def test_mt19937_matches_glib_grand_fixture(oracle):
    # GLib's GRand words (fixture written by tests/golden/make_glib_grand.py
    # from the system libglib-2.0) are what the reference's HOST_MT seeding
    # copies into the seed buffer (src/cl_ops/rng/clo_rng.c:196-203).
    fx = _glib_fixture()
    for seed, rec in fx["seeds"].items():
        words = oracle.mt19937(int(seed), 1000)
        assert [int(x) for x in words[:32]] == rec["first32"], seed
        assert int(words[999]) == rec["word1000"], seed


def test_host_mt_seeds_match_glib_grand_fixture(oracle):
    fx = _glib_fixture()
    for seed, rec in fx["seeds"].items():
        st = oracle.rng_init_host_mt("xorshift128", 8, int(seed))  # 8 states x 4 words
        assert [int(x) for x in st.view(np.uint32)] == rec["first32"], seed
        st = oracle.rng_init_host_mt("mwc64x", 16, int(seed))  # 16 states x 2 words
        assert [int(x) for x in st.view(np.uint32)] == rec["first32"], seed


# --------------------------------------------------------------------------
# Cost-model MLP (FP64 oracle): finite differences and closed forms.

def _mlp_case(oracle, n_in=7, rows=33, seed=5):
    rng = np.random.default_rng(seed)
    o = oracle.mlp_layout(n_in)
    p = rng.normal(0, 0.4, o["count"])
    X = rng.uniform(-1, 1, (rows, n_in))
    t = rng.normal(0, 1, rows)
    return p, X, t


def test_mlp_gradient_finite_differences(oracle):
    p, X, t = _mlp_case(oracle)
    loss, g = oracle.mlp_grad(p, X, t)
    assert loss == pytest.approx(oracle.mlp_loss(p, X, t), rel=1e-14)
    h = 1e-6
    for q in range(p.size):
        e = np.zeros_like(p)
        e[q] = h
        fd = (oracle.mlp_loss(p + e, X, t) - oracle.mlp_loss(p - e, X, t)) / (2 * h)
        assert g[q] == pytest.approx(fd, rel=1e-5, abs=1e-9), q


def test_mlp_closed_forms(oracle):
    n_in = 3
    o = oracle.mlp_layout(n_in)
    assert o["count"] == 32 * 3 + 32 + 32 * 32 + 32 + 32 + 1
    p = np.zeros(o["count"])
    p[o["b3"]] = 2.5  # all weights zero: y = b3 everywhere
    X = np.ones((4, n_in))
    assert np.allclose(oracle.mlp_forward(p, X), 2.5)
    loss, g = oracle.mlp_grad(p, X, np.zeros(4))
    assert loss == pytest.approx(0.5 * 2.5 ** 2)
    assert g[o["b3"]] == pytest.approx(2.5) and np.count_nonzero(g) == 1
    # one path: x0 -> unit 0 -> unit 0 -> y, all positive: y = w * x0 + b
    p[:] = 0
    p[o["w1"]] = 2.0
    p[o["w2"]] = 3.0
    p[o["w3"]] = 0.5
    p[o["b3"]] = 1.0
    X = np.array([[1.0, 9, 9], [2.0, 9, 9]])
    assert np.allclose(oracle.mlp_forward(p, X), [4.0, 7.0])


def test_adam_closed_form(oracle):
    # first step moves every weight by lr * sign(g) (bias-corrected moments)
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -7.0, 1e-3])
    p1, m, v = oracle.adam_step(p, np.zeros(3), np.zeros(3), g, 1, 0.01, eps=0.0)
    assert np.allclose(p1, p - 0.01 * np.sign(g))


# --------------------------------------------------------------------------
# tauslcg (src/cl_ops/rng/clo_rng_tauslcg.cl): Tausworthe components and the
# LCG, each pinned by what the mathematics fixes.

def _taus_fn(oracle, s1, s2, s3, m):
    return lambda z: int(oracle.lib().orc_taus(z & 0xFFFFFFFF, s1, s2, s3, m))


def test_tauslcg_component1_period(oracle):
    # (13, 19, 12, ~1) is L'Ecuyer's taus88 first component: a 31-bit LFSR of
    # period 2^31 - 1 (a prime), so T^(2^31-1) fixes every reachable state and
    # T itself fixes none.
    T = _taus_fn(oracle, 13, 19, 12, 0xFFFFFFFE)
    cols = _gf2_matrix(T, 32)
    P = _gf2_pow(cols, (1 << 31) - 1, 32)
    for z0 in (0x12345678, 0xDEADBEEF, 0x80000001):
        z1 = T(z0)
        assert _gf2_apply(P, z1) == z1
        assert T(z1) != z1


def test_tauslcg_component2_period(oracle):
    # (2, 25, 4, ~7) is taus88's second component: period 2^29 - 1
    # = 233 * 1103 * 2089.
    T = _taus_fn(oracle, 2, 25, 4, 0xFFFFFFF8)
    cols = _gf2_matrix(T, 32)
    e = (1 << 29) - 1
    assert 233 * 1103 * 2089 == e
    P = _gf2_pow(cols, e, 32)
    for z0 in (0x12345678, 0xCAFEBABE):
        z = T(T(T(z0)))
        assert _gf2_apply(P, z) == z
        for p in (233, 1103, 2089):
            assert _gf2_apply(_gf2_pow(cols, e // p, 32), z) != z


def test_tauslcg_lcg_word(oracle):
    # The fourth word is Numerical Recipes' quick generator (ranqd1)
    # 1664525 * x + 1013904223; its published sequence from 0:
    seq = [0, 1013904223, 1196435762, 3519870697, 2868466484, 1649599747, 2670642822,
           1476291629, 2748932008, 2180890343, 2498801434]
    for a, b in zip(seq, seq[1:]):
        st = np.array([a, 0, 0, 0], dtype=np.uint32).view(np.uint8)
        oracle.rng_next("tauslcg", st)
        assert int(st.view(np.uint32)[3]) == b


def test_tauslcg_seed_and_rotation(oracle):
    s = 0x0123456789ABCDEF
    st = oracle.rng_from_seed("tauslcg", s).view(np.uint32)
    assert [int(w) for w in st] == [0x89ABCDEF, 0x01234567, 0x89ABCDEF, 0x01234567]
    # one step moves each word one slot towards x through its own map
    T1 = _taus_fn(oracle, 13, 19, 12, 0xFFFFFFFE)
    T2 = _taus_fn(oracle, 2, 25, 4, 0xFFFFFFF8)
    T3 = _taus_fn(oracle, 3, 11, 17, 0xFFFFFFFE)
    x, y, z, w = (int(v) for v in st)
    b = st.view(np.uint8).copy()
    out = oracle.rng_next("tauslcg", b)
    nx, ny, nz, nw = (int(v) for v in b.view(np.uint32))
    assert (nx, ny, nz) == (T1(y), T2(z), T3(w)) and out == nx
    assert nw == (1664525 * x + 1013904223) & 0xFFFFFFFF
