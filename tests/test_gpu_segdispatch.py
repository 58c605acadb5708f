"""Segmented sort under the dispatcher: every tile/grid configuration and
every per-length work-class table is bit-exact against the oracle, the
segmented features match their closed form, the tuners return what they
promise, and the learned per-segment class choice is bit-exact and no slower
than the length thresholds alone."""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lengths(oracle, nseg, max_len, seed):
    u = oracle.generate(oracle.GEN_U64, seed, nseg)
    x = (u >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.clip(np.floor(np.exp(x * np.log(max_len))), 1, max_len).astype(np.uint32)


def _offsets(lens):
    off = np.zeros(lens.size, dtype=np.uint32)
    off[1:] = np.cumsum(lens[:-1], dtype=np.uint64).astype(np.uint32)
    return off


def _ref(oracle, k, off, n):
    out = k.copy()
    for s, e in zip(off, np.append(off[1:], n)):
        if e - s > 1:
            out[s:e] = np.sort(k[s:e])
    return out


def _case(clops, oracle, ctx, dtype, nseg, max_len, seed):
    lens = _lengths(oracle, nseg, max_len, seed)
    lens[::17] = 0  # empty segments among the others
    off = _offsets(lens)
    n = int(lens.sum(dtype=np.uint64))
    k = oracle.generate(oracle.GEN_U64, seed + 1, n).astype(dtype)
    return k, off, n, clops.Buffer.from_array(ctx, k), clops.Buffer.from_array(ctx, off)


@pytest.mark.parametrize("width", [4, 8])
def test_every_segmented_config_bit_exact(clops, oracle, ctx, width):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    dtype = np.uint64 if width == 8 else np.uint32
    for nseg, max_len in ((3000, 3000), (300, 40000)):
        k, off, n, kb, ob = _case(clops, oracle, ctx, dtype, nseg, max_len, 0x5D1 + nseg)
        want = _ref(oracle, k, off, n)
        out = clops.Buffer(ctx, n * width)
        ran = 0
        for c in range(clops.N_CONFIGS):
            if not d.feasible(c, clops.OP_SEGSORT, n, width):
                continue
            d.run_segmented(q, c, kb, out, ob, nseg, n, width)
            q.finish()
            assert np.array_equal(out.read(q, dtype, n), want), d.config_name(c)
            ran += 1
        assert ran == (12 if width == 4 else 8)
        assert np.array_equal(kb.read(q, dtype, n), k)  # input untouched


def _tables(rng):
    yield "zero", np.zeros(64, dtype=np.uint8)
    yield "medium", np.ones(64, dtype=np.uint8)
    yield "large", np.full(64, 2, dtype=np.uint8)
    for i in range(3):
        yield "random%d" % i, rng.integers(0, 3, 64).astype(np.uint8)


@pytest.mark.parametrize("cfg", [12, 18, 22])
def test_class_tables_bit_exact(clops, oracle, ctx, cfg):
    """Lengths over every bin up to 2^15 (so every class boundary of the
    tile), keys only through the dispatcher, for every table."""
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    rng = np.random.default_rng(cfg)
    k, off, n, kb, ob = _case(clops, oracle, ctx, np.uint32, 4000, 1 << 15, 0x5D2 + cfg)
    nseg = off.size
    want = _ref(oracle, k, off, n)
    out = clops.Buffer(ctx, n * 4)
    for name, t in _tables(rng):
        d.run_segmented(q, cfg, kb, out, ob, nseg, n, 4, classes=t)
        q.finish()
        assert np.array_equal(out.read(q, np.uint32, n), want), name


@pytest.mark.parametrize("dtype", [np.uint32, np.uint64])
@pytest.mark.parametrize("inplace", [False, True])
def test_segclass_option_kv(clops, oracle, ctx, queue, dtype, inplace):
    """The sorter option, key-value data, in place and out of place: stable
    order identical for every table."""
    rng = np.random.default_rng(7)
    lens = _lengths(oracle, 2500, 6000, 0x5D3)
    off = _offsets(lens)
    n = int(lens.sum(dtype=np.uint64))
    k = (oracle.generate(oracle.GEN_U64, 0x5D4, n) % np.uint64(50)).astype(dtype)  # many ties
    v = np.arange(n, dtype=np.uint32)
    rk, rv = k.copy(), v.copy()
    for s, e in zip(off, np.append(off[1:], n)):
        if e - s > 1:
            rk[s:e], rv[s:e] = oracle.stable_sort(k[s:e], v[s:e])
    ob = clops.Buffer.from_array(ctx, off)
    for name, t in _tables(rng):
        s = clops.Sorter(ctx, "satradix", dtype, options="segclass=" + "".join(map(str, t)))
        kb, vb = clops.Buffer.from_array(ctx, k), clops.Buffer.from_array(ctx, v)
        if inplace:
            s.sort_segmented_device(queue, kb, ob, off.size, n, vin=vb)
            ko, vo = kb, vb
        else:
            ko = clops.Buffer(ctx, n * np.dtype(dtype).itemsize)
            vo = clops.Buffer(ctx, n * 4)
            s.sort_segmented_device(queue, kb, ob, off.size, n, vin=vb, kout=ko, vout=vo)
        queue.finish()
        assert np.array_equal(ko.read(queue, dtype, n), rk), name
        assert np.array_equal(vo.read(queue, np.uint32, n), rv), name


def test_segclass_option_rejects_bad_digits(clops, ctx):
    for bad in ("3", "01a", "0" * 65):
        with pytest.raises(clops.ClopsError) as e:
            clops.Sorter(ctx, "satradix", np.uint32, options="segclass=" + bad)
        assert e.value.code == clops.CLO_ERROR_ARGS


def test_segmented_features_closed_form(clops, oracle, ctx):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    for nseg, max_len in ((5000, 9000), (100, 100), (10000, 65536)):
        k, off, n, kb, ob = _case(clops, oracle, ctx, np.uint32, nseg, max_len, 0x5D5 + nseg)
        f = d.features_segmented(q, kb, ob, nseg, n, 4)
        fs = d.features(q, clops.OP_SORT, kb, n, 4)
        assert np.allclose(f[:7], fs[:7]) and f[7] == 2
        S = min(nseg, 4096)
        j = (np.arange(S, dtype=np.uint64) * np.uint64(nseg)) // np.uint64(S)
        lens = np.append(off[1:], n).astype(np.int64)[j] - off[j].astype(np.int64)
        assert f[8] == pytest.approx(np.mean(np.log2(lens + 1.0)) / 16, rel=1e-4)
        big = lens[lens > 4096].sum() / lens.sum()
        assert f[9] == pytest.approx(big, abs=1e-4)


def test_segmented_tuner_and_table(clops, oracle, ctx):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    k, off, n, kb, ob = _case(clops, oracle, ctx, np.uint32, 2000, 20000, 0x5D6)
    nseg = off.size
    out = clops.Buffer(ctx, n * 4)
    t = d.tune_segmented(q, kb, out, ob, nseg, n, 4, reps=2)
    for c in range(clops.N_CONFIGS):
        assert (t[c] > 0) == d.feasible(c, clops.OP_SEGSORT, n, 4), c
    assert np.array_equal(out.read(q, np.uint32, n), _ref(oracle, k, off, n))
    f = d.features_segmented(q, kb, ob, nseg, n, 4)
    d.table_add(clops.OP_SEGSORT, n, 4, f, t)
    best = int(np.argmin(np.where(t > 0, t, np.inf)))
    ch = d.choose(None, clops.POLICY_TABLE, [clops.OP_SEGSORT], [n], [4], f[None, :])
    assert ch[0] == best


def test_segment_class_timings(clops, ctx):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    # tile 16: CAP 4096, W 2048
    for length, feasible in ((1, [1, 1, 1]), (2048, [1, 1, 1]), (2049, [0, 1, 1]),
                             (4096, [0, 1, 1]), (4097, [0, 0, 1]), (65536, [0, 0, 1])):
        t = d.tune_segment_class(q, 18, 4, length, reps=1)
        assert [int(x > 0) for x in t] == feasible, (length, t)
        assert all(x == -1 for x, f in zip(t, feasible) if not f)


def _time_best_pair(q, fa, fb, reps):
    """Least times of fa and fb over reps interleaved runs each (a run
    measured first would otherwise carry the warm-up of both)."""
    fa()
    fb()
    q.finish()
    best = [math.inf, math.inf]
    for _ in range(reps):
        for i, fn in enumerate((fa, fb)):
            t0 = time.perf_counter()
            fn()
            q.finish()
            best[i] = min(best[i], time.perf_counter() - t0)
    return best


def test_learned_segment_classes(clops, oracle, ctx):
    """Train the per-segment cost model on work-class timings from the tuner
    (every length bin up to CAP, each feasible class), take its class table,
    and run a log-uniform segmented workload with it: bit-exact with the
    length-threshold rule, and no slower than it (5% timing tolerance, best
    of 7 runs each)."""
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    cfg = 18
    X, y, meas = [], [], {}
    for b in range(clops.SEG_BINS):
        L = d.seg_bin_length(b)
        if L == 0 or L > 4096:
            continue
        t = d.tune_segment_class(q, cfg, 4, L, reps=2)
        meas[b] = t
        for c in range(3):
            if t[c] > 0:
                X.append(d.segment_row(L, 4, cfg, c))
                y.append(math.log2(t[c]))
    m = clops.Mlp(ctx, clops.SEG_INPUTS, seed=0x5E)
    loss = m.train(q, np.array(X), np.array(y), 3000, 3e-3)
    assert loss < 0.05
    d.set_segment_model(m)
    classes = d.segment_classes(q, cfg, 4, 0.05)
    assert classes.size == 64 and classes.max() <= 2
    for b, t in meas.items():
        if classes[b]:
            # A raised class is one the tuner timed as feasible for the bin.
            assert t[classes[b]] > 0
    # Without a model the table is the threshold rule.
    d2 = clops.Dispatch(ctx)
    assert not d2.segment_classes(None, cfg, 4).any()
    nseg = 16384
    lens = _lengths(oracle, nseg, 8192, 0x5D7)
    off = _offsets(lens)
    n = int(lens.sum(dtype=np.uint64))
    kb = clops.Buffer(ctx, n * 4)
    clops.generate(q, kb, n, clops.GEN_U32_HI, 0x5D8)
    ob = clops.Buffer.from_array(ctx, off)
    o1, o2 = clops.Buffer(ctx, n * 4), clops.Buffer(ctx, n * 4)
    t_fixed, t_model = _time_best_pair(
        q, lambda: d.run_segmented(q, cfg, kb, o1, ob, nseg, n, 4),
        lambda: d.run_segmented(q, cfg, kb, o2, ob, nseg, n, 4, classes=classes), 9)
    assert clops.count_diff(q, o1, o2, n) == 0
    k = kb.read(q, np.uint32, n)
    got = o2.read(q, np.uint32, n)
    assert np.array_equal(got, _ref(oracle, k, off, n))
    print("learned classes", "".join(map(str, classes)), "fixed %.3f ms, model %.3f ms"
          % (t_fixed * 1e3, t_model * 1e3), flush=True)
    assert t_model <= 1.05 * t_fixed
