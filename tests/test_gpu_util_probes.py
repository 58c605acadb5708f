"""The bandwidth probes move exactly the bytes they claim: the stream copy is
a copy, and the scatter probe writes run d of tile t to
d * (n * run / 8192) + t * run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_stream_copy(clops, oracle, ctx):
    q = clops.Queue(ctx)
    nb = 16384 * 37
    x = oracle.generate(oracle.GEN_U64, 0x57, nb // 8)
    a, b = clops.Buffer.from_array(ctx, x), clops.Buffer(ctx, nb)
    clops._call(clops.lib().clo_util_stream_copy, q.ptr, b.ptr, a.ptr, nb)
    assert np.array_equal(b.read(q, np.uint64, nb // 8), x)
    with pytest.raises(clops.ClopsError):
        clops._call(clops.lib().clo_util_stream_copy, q.ptr, b.ptr, a.ptr, nb - 16)


@pytest.mark.parametrize("run", [4, 8, 32, 128])
def test_scatter_runs_layout(clops, oracle, ctx, run):
    q = clops.Queue(ctx)
    tiles = 24
    n = 8192 * tiles
    k = oracle.generate(oracle.GEN_U64, 0x58, n)
    v = np.arange(n, dtype=np.uint32)
    ks, vs = clops.Buffer.from_array(ctx, k), clops.Buffer.from_array(ctx, v)
    kd, vd = clops.Buffer(ctx, n * 8), clops.Buffer(ctx, n * 4)
    clops._call(clops.lib().clo_util_scatter_runs, q.ptr, kd.ptr, vd.ptr, ks.ptr, vs.ptr, n, run)
    i = np.arange(n, dtype=np.int64)
    t, w = i // 8192, i % 8192
    dst = (w // run) * (n * run // 8192) + t * run + (w % run)
    want_k = np.empty_like(k)
    want_k[dst] = k
    want_v = np.empty_like(v)
    want_v[dst] = v
    assert np.array_equal(kd.read(q, np.uint64, n), want_k)
    assert np.array_equal(vd.read(q, np.uint32, n), want_v)
