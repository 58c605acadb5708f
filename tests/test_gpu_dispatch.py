"""Dispatcher on the GPU: every configuration is bit-exact, the feature kernel
matches its closed form, the tuner and its table behave, and the learned
choice beats the fixed configuration on a cost surface with a known optimum."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [1, 2, 100, 4095, 4096, 5000, 65536, 70001, 300007]


def _data(clops, oracle, ctx, q, kind_name, width, n, seed):
    dt = np.uint64 if width == 8 else np.uint32
    kinds = {
        "uniform": (oracle.GEN_U64 if width == 8 else oracle.GEN_U32_HI, 0),
        "few": (oracle.GEN_U64_MOD if width == 8 else oracle.GEN_U32_MOD, 16),
        "nearly": (oracle.GEN_NEARLY_SORTED_U64 if width == 8 else oracle.GEN_NEARLY_SORTED_U32,
                   1 << 20),
        "reversed": (oracle.GEN_REVERSED_U64 if width == 8 else oracle.GEN_REVERSED_U32, 3),
    }
    kind, param = kinds[kind_name]
    x = oracle.generate(kind, seed, n, param=param).astype(dt)
    return x, clops.Buffer.from_array(ctx, x)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", SIZES)
def test_every_sort_config_bit_exact(clops, oracle, ctx, width, n):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    for dist in ("uniform", "few", "nearly", "reversed"):
        x, b = _data(clops, oracle, ctx, q, dist, width, n, n + width)
        want = np.sort(x)
        out = clops.Buffer(ctx, max(n * width, 4))
        for c in range(clops.N_CONFIGS):
            if not d.feasible(c, clops.OP_SORT, n, width):
                continue
            d.run(q, c, clops.OP_SORT, b, out, n, width)
            q.finish()
            got = out.read(q, x.dtype, n)
            assert np.array_equal(got, want), (d.config_name(c), dist)
            assert np.array_equal(b.read(q, x.dtype, n), x)  # input left unchanged


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", SIZES + [1 << 20])
def test_every_scan_config_bit_exact(clops, oracle, ctx, width, n):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    x, b = _data(clops, oracle, ctx, q, "uniform", width, n, 7 * n)
    want = oracle.scan(x, x.dtype)
    out = clops.Buffer(ctx, n * width)
    for c in range(clops.N_CONFIGS):
        if not d.feasible(c, clops.OP_SCAN, n, width):
            continue
        d.run(q, c, clops.OP_SCAN, b, out, n, width)
        q.finish()
        assert np.array_equal(out.read(q, x.dtype, n), want), d.config_name(c)


def _features_ref(x, op):
    n = x.size
    S = min(n, 4096)
    pos = (np.arange(S, dtype=np.uint64) * np.uint64(n)) // np.uint64(S)
    ok = pos + 1 < n
    a, b = x[pos[ok]], x[pos[ok] + 1]
    pairs = max(int(ok.sum()), 0)
    w = x.dtype.itemsize
    ents = []
    for byte in range(w):
        v = ((x[pos] >> np.array(8 * byte, dtype=x.dtype)) & np.array(255, dtype=x.dtype))
        c = np.bincount(v.astype(np.int64), minlength=256)
        p = c[c > 0] / S
        ents.append(float(-(p * np.log2(p)).sum()))
    ents = np.array(ents)
    # One segment of n: mean log2(len + 1) / 16 and the share of keys in
    # segments longer than 4096.
    return np.array([math.log2(n) / 32, w / 8, ents.mean() / 8, ents.max() / 8,
                     np.mean(ents < 1e-6), (a < b).sum() / pairs if pairs else 0,
                     (a > b).sum() / pairs if pairs else 0, op,
                     math.log2(n + 1) / 16, 1.0 if n > 4096 else 0.0], dtype=np.float64)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("dist", ["uniform", "few", "nearly", "reversed"])
@pytest.mark.parametrize("n", [1, 3, 1000, 4096, 123457])
def test_features_closed_form(clops, oracle, ctx, width, dist, n):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    x, b = _data(clops, oracle, ctx, q, dist, width, n, 11)
    f = d.features(q, clops.OP_SCAN, b, n, width)
    assert np.allclose(f, _features_ref(x, 1), atol=2e-5)


def test_feature_signatures(clops, oracle, ctx):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    n = 200000
    _, b = _data(clops, oracle, ctx, q, "nearly", 8, n, 1)
    f = d.features(q, clops.OP_SORT, b, n, 8)
    assert f[5] > 0.95 and f[7] == 0
    _, b = _data(clops, oracle, ctx, q, "reversed", 4, n, 1)
    assert d.features(q, clops.OP_SORT, b, n, 4)[6] == 1.0
    _, b = _data(clops, oracle, ctx, q, "few", 8, n, 1)
    f = d.features(q, clops.OP_SORT, b, n, 8)
    assert f[3] <= 0.5 + 1e-6 and f[4] == pytest.approx(7 / 8)  # 16 values: 4 bits, 7 zero bytes


def test_tuner_and_table(clops, oracle, ctx, tmp_path):
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    n = 3000
    x, b = _data(clops, oracle, ctx, q, "uniform", 4, n, 5)
    out = clops.Buffer(ctx, n * 4)
    t = d.tune(q, clops.OP_SORT, b, out, n, 4, reps=2)
    for c in range(clops.N_CONFIGS):
        assert (t[c] > 0) == d.feasible(c, clops.OP_SORT, n, 4)
    assert np.array_equal(out.read(q, np.uint32, n), np.sort(x))
    f = d.features(q, clops.OP_SORT, b, n, 4)
    assert d.table_choose(clops.OP_SORT, n, 4, f) == -1
    d.table_add(clops.OP_SORT, n, 4, f, t)
    best = int(np.argmin(np.where(t > 0, t, np.inf)))
    assert d.table_choose(clops.OP_SORT, n, 4, f) == best
    path = str(tmp_path / "tune.txt")
    d.table_save(path)
    d2 = clops.Dispatch(ctx)
    d2.table_load(path)
    assert d2.table_choose(clops.OP_SORT, n, 4, f) == best
    # fixed / table policies
    ch = d2.choose(None, clops.POLICY_TABLE, [0, 0], [n, 1 << 24], [4, 8], np.stack([f, f]))
    assert ch[0] == best and ch[1] == clops.Dispatch.fixed(clops.OP_SORT, 1 << 24, 8)


def _synthetic_cost(ops, ns, widths, feats, c):
    """A cost surface (microseconds) with different winners in different
    regions: radix-like linear costs, a bitonic-like n log^2 n, a
    one-block-like cost that only fits small n, and scan tiles."""
    lg = np.log2(ns)
    w = widths / 4.0
    sortedness = feats[:, 5]
    if c == 0:
        return 20 + 0.004 * ns * w
    if c == 1:
        return 35 + 0.0032 * ns * w * (1 + 0.5 * sortedness)
    if c == 2:
        return 12 + 0.006 * ns * w
    if c == 3:
        return 8 + 0.00002 * ns * lg * lg * w
    if c == 4:
        return 4 + 0.0009 * ns * w
    if c == 5:
        return 3 + 0.0000008 * ns * ns
    if c == 6:
        return 6 + 0.0011 * ns * w
    if c == 7:
        return 9 + 0.0008 * ns * w
    if c == 9:
        return 30 + 0.008 * ns * w
    if c == 10:
        return 25 + 0.006 * ns * w
    if c == 11:
        return 22 + 0.005 * ns * w
    return 15 + 0.0007 * ns * w


def clops_n_features():
    import clops
    return clops.N_FEATURES


def _requests(rng, count):
    ops = rng.integers(0, 2, count).astype(np.int32)
    lg = rng.uniform(10, 24, count)
    ns = np.floor(2 ** lg).astype(np.uint64)
    widths = rng.choice([4, 8], count).astype(np.uint64)
    feats = np.zeros((count, clops_n_features()), dtype=np.float32)
    feats[:, 0] = np.log2(ns) / 32
    feats[:, 1] = widths / 8
    feats[:, 2] = rng.uniform(0.3, 1, count)
    feats[:, 3] = np.maximum(feats[:, 2], rng.uniform(0.3, 1, count))
    feats[:, 5] = rng.uniform(0, 1, count)
    feats[:, 6] = (1 - feats[:, 5]) * rng.uniform(0, 1, count)
    feats[:, 7] = ops
    feats[:, 8] = np.log2(ns + 1.0) / 16
    feats[:, 9] = ns > 4096
    return ops, ns, widths, feats


def test_learned_choice_beats_fixed_on_known_costs(clops, ctx):
    """Train the cost model on samples of a known cost surface and check the
    choice on new requests: regret <= 5% against the true optimum, and
    faster than the fixed configuration."""
    D = clops.Dispatch
    rng = np.random.default_rng(0xC5)
    ops, ns, widths, feats = _requests(rng, 3000)
    X, y = [], []
    for r in range(ops.size):
        for c in range(clops.N_CONFIGS):
            if D.feasible(c, int(ops[r]), int(ns[r]), int(widths[r])):
                X.append(D.model_row(feats[r], c))
                y.append(math.log2(_synthetic_cost(ops[r:r + 1], ns[r:r + 1].astype(np.float64),
                                                   widths[r:r + 1].astype(np.float64),
                                                   feats[r:r + 1], c)[0]))
    X, y = np.array(X, dtype=np.float32), np.array(y, dtype=np.float32)
    q = clops.Queue(ctx)
    m = clops.Mlp(ctx, clops.MODEL_INPUTS, seed=1)
    loss = m.train(q, X, y, 4000, 3e-3)
    assert loss < 0.05
    d = clops.Dispatch(ctx)
    d.set_model(m)
    ops, ns, widths, feats = _requests(rng, 1000)
    choice = d.choose(q, clops.POLICY_MODEL, ops, ns, widths, feats)
    fixed = d.choose(q, clops.POLICY_FIXED, ops, ns, widths, feats)

    def total(cfgs):
        s = 0.0
        for r in range(ops.size):
            s += _synthetic_cost(ops[r:r + 1], ns[r:r + 1].astype(np.float64),
                                 widths[r:r + 1].astype(np.float64), feats[r:r + 1],
                                 int(cfgs[r]))[0]
        return s

    best = []
    for r in range(ops.size):
        cand = [c for c in range(clops.N_CONFIGS)
                if D.feasible(c, int(ops[r]), int(ns[r]), int(widths[r]))]
        best.append(min(cand, key=lambda c: _synthetic_cost(
            ops[r:r + 1], ns[r:r + 1].astype(np.float64), widths[r:r + 1].astype(np.float64),
            feats[r:r + 1], c)[0]))
        assert D.feasible(int(choice[r]), int(ops[r]), int(ns[r]), int(widths[r]))
    t_model, t_best, t_fixed = total(choice), total(best), total(fixed)
    assert t_model / t_best - 1 <= 0.05
    assert t_model < t_fixed


def test_features_batch_matches_single(clops, oracle, ctx):
    """One launch over requests of both widths, both ops and every
    distribution (views into one pool): each row equals the single-request
    features."""
    q = clops.Queue(ctx)
    d = clops.Dispatch(ctx)
    pool = clops.Buffer(ctx, 1 << 24)
    views, ops, ns, ws, singles = [], [], [], [], []
    off = 0
    for i, (dist, width, n) in enumerate([("uniform", 4, 1), ("few", 8, 3), ("nearly", 4, 4096),
                                          ("reversed", 8, 123457), ("uniform", 8, 70001),
                                          ("few", 4, 1000)]):
        x, _ = _data(clops, oracle, ctx, q, dist, width, n, 40 + i)
        v = pool.view(off, n * width)
        v.write(q, x)
        off += (n * width + 255) // 256 * 256
        op = clops.OP_SORT if i % 2 else clops.OP_SCAN
        views.append(v)
        ops.append(op)
        ns.append(n)
        ws.append(width)
        singles.append(d.features(q, op, v, n, width))
    f = d.features_batch(q, ops, views, ns, ws)
    assert np.allclose(f, np.array(singles), rtol=0, atol=1e-6)
    with pytest.raises(clops.ClopsError):
        d.features_batch(q, [clops.OP_SEGSORT], views[:1], ns[:1], ws[:1])


def test_barrier_orders_queues(clops, oracle, ctx):
    """Work on one queue waits for an event of another: a fill on the side
    queue, then a sort on the exec queue that waited for it."""
    side, ex = clops.Queue(ctx), clops.Queue(ctx)
    n = 1 << 22
    b = clops.Buffer(ctx, n * 4)
    o = clops.Buffer(ctx, n * 4)
    s = clops.Sorter(ctx, "satradix", np.uint32)
    for rep in range(3):
        clops.generate(side, b, n, clops.GEN_U32_HI, 0xBA + rep)
        ex.wait_for(side.marker())
        s.sort_device(ex, b, o, n)
        ex.finish()
        want = np.sort(oracle.generate(oracle.GEN_U32_HI, 0xBA + rep, n))
        assert np.array_equal(o.read(ex, np.uint32, n), want), rep
    side.gc()
    ex.gc()
