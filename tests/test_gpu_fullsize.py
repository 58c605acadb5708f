"""The benchmark configurations at their stated sizes, checked on the GPU.

C3: 2^28 u64-key/u32-value pairs under the three key distributions, each
proven to be exactly the stable sort by four device checks (keys
non-decreasing, values increasing inside runs of equal keys, values a
permutation of 0..n-1, each output key the input key its value names), plus
the stable-sort position of a seeded sample of input pairs computed on the
host from its definition.  sbitonic/abitonic on 2^24 u32 keys against numpy.
C2: xorshift128 and mwc64x fills of 2^26 work-items x 16 draws in both seed
modes against the oracle's streams, then the scans of the masked draws.
C5: a bounded prefix of the mixed request stream, every feasible
configuration of every request against the oracle.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N28 = 1 << 28


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def kv28(clops, ctx):
    """Device buffers for 2^28 pairs, values 0..n-1."""
    q = clops.Queue(ctx)
    b = {name: clops.Buffer(ctx, N28 * w) for name, w in
         (("kin", 8), ("vin", 4), ("kout", 8), ("vout", 4))}
    clops.generate(q, b["vin"], N28, clops.GEN_IOTA_U32, 0)
    q.finish()
    yield q, b
    for x in b.values():
        x.close()


def _zipf_cdf(clops, ranks, s):
    cdf = np.empty(ranks, dtype=np.float64)
    clops.lib().clo_util_zipf_cdf(cdf.ctypes.data_as(ctypes.c_void_p), cdf.size, s)
    return cdf


def _stable_positions(keys, idx):
    """Position of pair i in the stable sort: #{j: k_j < k_i} + #{j < i: k_j = k_i}."""
    out = []
    for i in idx:
        k = keys[i]
        out.append(int(np.count_nonzero(keys < k)) + int(np.count_nonzero(keys[:i] == k)))
    return out


@pytest.mark.parametrize("dist", ["uniform", "below2^20", "zipf1.1"])
def test_c3_kv_2p28(clops, oracle, ctx, kv28, dist):
    q, b = kv28
    n = N28
    _say("C3 KV 2^28 u64/u32, distribution", dist)
    if dist == "uniform":
        clops.generate(q, b["kin"], n, clops.GEN_U64, 0xC3A)
    elif dist == "below2^20":
        clops.generate(q, b["kin"], n, clops.GEN_U64_MASK, 0xC3B, (1 << 20) - 1)
    else:
        cdf = clops.Buffer.from_array(ctx, _zipf_cdf(clops, 1 << 24, 1.1))
        clops._call(clops.lib().clo_util_generate_zipf, q.ptr, b["kin"].ptr, n, cdf.ptr,
                    1 << 24, 0xC3C)
    s = clops.Sorter(ctx, "satradix", np.uint64)
    s.sort_kv_device(q, b["kin"], b["vin"], b["kout"], b["vout"], n)
    q.finish()
    assert clops.check_sorted(q, b["kout"], n, 8, b["vout"]) == 0
    assert clops.check_pairs(q, b["kin"], b["kout"], b["vout"], n, 8) == 0
    assert clops.multiset_hash(q, b["kin"], n, 8) == clops.multiset_hash(q, b["kout"], n, 8)
    if dist == "below2^20":
        assert s.last_passes() == 3  # digits 3..7 are zero in every key
    # The stable-sort position of a seeded sample of input pairs, from the
    # definition, on the host.
    keys = b["kin"].read(q, np.uint64, n)
    rng = np.random.default_rng(0xC3)
    idx = np.sort(rng.choice(n, 12, replace=False))
    if dist == "zipf1.1":
        idx[0] = int(np.argmax(keys == keys[idx[1]]))  # an early copy of a repeated key
    for i, p in zip(idx, _stable_positions(keys, idx)):
        assert b["kout"].read(q, np.uint64, 1, offset=8 * p)[0] == keys[i]
        assert b["vout"].read(q, np.uint32, 1, offset=4 * p)[0] == i


def test_zipf_generator_matches_oracle(clops, oracle, ctx):
    ranks, n = 1 << 24, 1 << 20
    cdf_h = _zipf_cdf(clops, ranks, 1.1)
    assert np.array_equal(cdf_h, oracle.zipf_cdf(ranks, 1.1))
    q = clops.Queue(ctx)
    cdf = clops.Buffer.from_array(ctx, cdf_h)
    out = clops.Buffer(ctx, n * 8)
    clops._call(clops.lib().clo_util_generate_zipf, q.ptr, out.ptr, n, cdf.ptr, ranks, 0xC3C)
    got = out.read(q, np.uint64, n)
    assert np.array_equal(got, oracle.generate_zipf(cdf_h, 0xC3C, n))
    # The top rank (mix64(0)) takes 1/H(2^24, 1.1) of the draws, about 11.5%.
    top = np.count_nonzero(got == np.uint64(oracle.mix64(0))) / n
    assert 0.11 < top < 0.12


@pytest.mark.parametrize("algo", ["sbitonic", "abitonic"])
def test_c3_bitonic_2p24(clops, ctx, algo):
    n = 1 << 24
    q = clops.Queue(ctx)
    kin = clops.Buffer(ctx, n * 4)
    kout = clops.Buffer(ctx, n * 4)
    clops.generate(q, kin, n, clops.GEN_U32_HI, 0xC3D)
    s = clops.Sorter(ctx, algo, np.uint32)
    s.sort_device(q, kin, kout, n)
    q.finish()
    assert np.array_equal(kout.read(q, np.uint32, n), np.sort(kin.read(q, np.uint32, n)))


def _compare_fill(oracle, q, out, gen, states, G, draws, chunk=1 << 23):
    """Device output out[r*G + gid] against the oracle, chunks of gids."""
    ss = states.size // G
    for a in range(0, G, chunk):
        m = min(chunk, G - a)
        st = states[a * ss:(a + m) * ss].copy()
        ref = oracle.rng_fill(gen, st, draws, threads=16).reshape(draws, m)
        for r in range(draws):
            got = out.read(q, np.uint32, m, offset=4 * (r * G + a))
            assert np.array_equal(got, ref[r]), (gen, a, r)


@pytest.mark.parametrize("gen", ["xorshift128", "mwc64x"])
@pytest.mark.parametrize("mode", ["host_mt", "dev_gid"])
def test_c2_fill_and_scan_pipeline(clops, oracle, ctx, gen, mode):
    G, draws = 1 << 26, 16
    q = clops.Queue(ctx)
    _say("C2", gen, mode)
    if mode == "host_mt":
        r = clops.Rng(ctx, gen, clops.SEED_HOST_MT, G, main_seed=0)
        states = oracle.rng_init_host_mt(gen, G, 0)
    else:
        r = clops.Rng(ctx, gen, clops.SEED_DEV_GID, G, main_seed=0, q=q)
        states = oracle.rng_init_dev_gid(gen, G, 0, 0)
    q.finish()
    assert np.array_equal(r.read_states(q), states)
    out = clops.Buffer(ctx, G * draws * 4)
    r.fill(q, out, draws)
    q.finish()
    _compare_fill(oracle, q, out, gen, states, G, draws)
    if mode != "host_mt":
        return
    # The scans of the configuration, on the draws themselves.
    if gen == "xorshift128":
        a = out.read(q, np.uint32, N28) & np.uint32(0xFF)
        bi = clops.Buffer.from_array(ctx, a)
        for sumt in (np.uint32, np.uint64):
            sc = clops.Scanner(ctx, np.uint32, sumt)
            bo = clops.Buffer(ctx, N28 * np.dtype(sumt).itemsize)
            sc.scan_device(q, bi, bo, N28)
            q.finish()
            assert np.array_equal(bo.read(q, sumt, N28), oracle.scan(a, sumt)), sumt
            bo.close()
    else:
        n = 1 << 27
        a = out.read(q, np.uint32, n).astype(np.uint64)
        bi = clops.Buffer.from_array(ctx, a)
        sc = clops.Scanner(ctx, np.uint64, np.uint64)
        bo = clops.Buffer(ctx, n * 8)
        sc.scan_device(q, bi, bo, n)
        q.finish()
        assert np.array_equal(bo.read(q, np.uint64, n), oracle.scan(a, np.uint64))


@pytest.fixture(scope="module")
def c5_requests(clops, ctx):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    q = clops.Queue(ctx)
    data = bench.RequestData(ctx, q, 1 << 24)
    return bench, q, data, bench.request_stream(0xC5, 200, 24)


@pytest.mark.parametrize("part", [0, 1])
def test_c5_stream_prefix_every_config(clops, oracle, ctx, c5_requests, part):
    """Requests 100*part .. 100*part+99 of the seed-0xC5 stream (sizes capped
    at 2^24): every configuration the tuner, table or model could choose is
    run and compared word for word with the oracle's result."""
    bench, q, data, reqs = c5_requests
    d = clops.Dispatch(ctx)
    ref = clops.Buffer(ctx, (1 << 24) * 8)
    checked = 0
    for j in range(100 * part, 100 * part + 100):
        op, n, width, dist, seed = reqs[j]
        data.fill(n, width, dist, seed)
        dt = np.uint64 if width == 8 else np.uint32
        x = data.inb.read(q, dt, n)
        want = np.sort(x) if op == clops.OP_SORT else oracle.scan(x, dt)
        ref.write(q, want)
        for c in range(clops.N_CONFIGS):
            if not d.feasible(c, op, n, width):
                continue
            d.run(q, c, op, data.inb, data.outb, n, width)
            diff = clops.count_diff(q, data.outb, ref, n * width // 4)
            assert diff == 0, (j, op, n, width, dist, d.config_name(c))
            checked += 1
    _say("C5 part", part, "outputs checked", checked)
