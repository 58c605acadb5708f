"""numpy front end of the CPU oracle (liboracle.so, built from oracle.c).

TEST INFRASTRUCTURE ONLY: imported by the tests, the smoke check and the
benchmark's CPU-baseline leg as the checker / timed CPU baseline.  The product
path (libcl_ops.so) never loads it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GEN_U64, GEN_U32_HI, GEN_U64_MASK, GEN_U32_MASK, GEN_IOTA_U32 = 0, 1, 2, 3, 4
GEN_U32_MOD, GEN_U64_MOD = 6, 7
GEN_NEARLY_SORTED_U32, GEN_NEARLY_SORTED_U64, GEN_REVERSED_U32, GEN_REVERSED_U64 = 8, 9, 10, 11
RNG_NAMES = ["lcg", "xorshift64", "xorshift128", "mwc64x", "parkmiller", "tauslcg"]
STATE_SIZE = {"lcg": 8, "xorshift64": 8, "xorshift128": 16, "mwc64x": 8, "parkmiller": 4,
              "tauslcg": 16}

_U = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}
_S = {1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(path)
        vp, sz, u64, u32, i32 = (ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64,
                                 ctypes.c_uint32, ctypes.c_int)
        sig = {
            "orc_mix64": (u64, [u64]),
            "orc_generate": (None, [i32, u64, u64, u64, sz, vp]),
            "orc_zipf_cdf": (None, [vp, sz, ctypes.c_double]),
            "orc_generate_zipf": (None, [vp, sz, u64, sz, vp]),
            "orc_satradix": (None, [vp, vp, sz, i32, i32, sz, i32]),
            "orc_stable_sort": (None, [vp, vp, sz, i32]),
            "orc_sbitonic": (None, [vp, sz, i32, i32, i32]),
            "orc_gselect": (None, [vp, vp, vp, vp, sz, i32, i32, i32]),
            "orc_scan": (None, [vp, i32, i32, vp, i32, sz]),
            "orc_scan_blelloch": (None, [vp, vp, i32, sz, sz, i32]),
            "orc_rng_state_size": (sz, [i32]),
            "orc_rng_from_seed": (None, [i32, u64, vp]),
            "orc_rng_next": (u32, [i32, vp]),
            "orc_rng_init_dev_gid": (None, [i32, sz, u64, i32, vp]),
            "orc_rng_init_host_mt": (None, [i32, sz, u64, vp]),
            "orc_mt19937": (None, [u32, sz, vp]),
            "orc_taus": (u32, [u32, i32, i32, i32, u32]),
            "orc_rng_fill": (None, [i32, vp, sz, u32, u32, u32, vp, i32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def mix64(z):
    return lib().orc_mix64(z & 0xFFFFFFFFFFFFFFFF)


def generate(kind, seed, n, param=0, start=0):
    dt = np.uint64 if kind in (GEN_U64, GEN_U64_MASK, GEN_U64_MOD, GEN_NEARLY_SORTED_U64,
                               GEN_REVERSED_U64) else np.uint32
    out = np.empty(n, dtype=dt)
    lib().orc_generate(kind, seed, param, start, n, _p(out))
    return out


def zipf_cdf(ranks, s):
    cdf = np.empty(ranks, dtype=np.float64)
    lib().orc_zipf_cdf(_p(cdf), ranks, s)
    return cdf


def generate_zipf(cdf, seed, n):
    out = np.empty(n, dtype=np.uint64)
    lib().orc_generate_zipf(_p(cdf), len(cdf), seed, n, _p(out))
    return out


def satradix(keys, vals=None, bits=4, lws=256, threads=1):
    """Reference-structured satradix; returns sorted copies."""
    k = np.array(keys, copy=True)
    v = None if vals is None else np.array(vals, dtype=np.uint32, copy=True)
    lib().orc_satradix(_p(k), _p(v), k.size, k.dtype.itemsize, bits, lws, threads)
    return k if v is None else (k, v)


def stable_sort(keys, vals=None):
    k = np.array(keys, copy=True)
    v = None if vals is None else np.array(vals, dtype=np.uint32, copy=True)
    lib().orc_stable_sort(_p(k), _p(v), k.size, k.dtype.itemsize)
    return k if v is None else (k, v)


def sbitonic(keys, desc=False):
    k = np.array(keys, copy=True)
    if k.size & (k.size - 1):
        raise ValueError("the reference network needs a power-of-two size")
    lib().orc_sbitonic(_p(k), k.size, k.dtype.itemsize, int(k.dtype.kind == "i"), int(desc))
    return k


def gselect(keys, vals=None, desc=False):
    k = np.ascontiguousarray(keys)
    ko = np.empty_like(k)
    v = None if vals is None else np.ascontiguousarray(vals, dtype=np.uint32)
    vo = None if v is None else np.empty_like(v)
    lib().orc_gselect(_p(k), _p(v), _p(ko), _p(vo), k.size, k.dtype.itemsize,
                      int(k.dtype.kind == "i"), int(desc))
    return ko if v is None else (ko, vo)


# ---- sorts with user key / compare functions ------------------------------
# The Python callables stand for the CLO_SORT_KEY_GET(x) and
# CLO_SORT_COMPARE(a, b) snippets; they take and return numpy arrays.

def _swap_step(d, pad, i1, i2, desc, key, compare):
    a, b = d[i1], d[i2]
    swap = np.asarray(compare(key(a), key(b)), dtype=bool) ^ desc
    if pad is not None:
        pa, pb = pad[i1], pad[i2]
        either = pa | pb
        swap = np.where(either, np.where(desc, pb & ~pa, pa & ~pb), swap)
        pad[i1[swap]], pad[i2[swap]] = pb[swap], pa[swap]
    d[i1[swap]], d[i2[swap]] = b[swap], a[swap]


def bitonic_network(data, key=lambda x: x, compare=np.greater, padded=False):
    """CL_Ops' bitonic network with a key getter and a compare: the kernel of
    sbitonic.cl:38-69 (pair i1 = gid + (gid / 2^(step-1)) * 2^(step-1),
    i2 = i1 + 2^(step-1); desc = bit (stage-1) of gid; swap iff
    COMPARE(KEY_GET(d1), KEY_GET(d2)) ^ desc) driven by the stage/step loops
    of sbitonic.c:102-118.  abitonic.cl's ABIT_CMPXCH applies the same
    comparators in the same dependency order, so this is its result too.

    The reference needs a power-of-two size.  padded=True instead runs the
    network on the next power of two, with the extra slots flagged as ordering
    after every element, and returns the first len(data) slots (the library's
    rule for other sizes)."""
    d = np.array(data, copy=True)
    n = d.size
    if n < 2:
        return d
    npow2 = 1 << (n - 1).bit_length()
    if npow2 != n and not padded:
        raise ValueError("the reference network needs a power-of-two size")
    pad = None
    if npow2 != n:
        d = np.concatenate([d, np.zeros(npow2 - n, dtype=d.dtype)])
        pad = np.arange(npow2) >= n
    gid = np.arange(npow2 // 2, dtype=np.int64)
    for stage in range(1, npow2.bit_length()):
        desc = ((gid >> (stage - 1)) & 1).astype(bool)
        for step in range(stage, 0, -1):
            ps = 1 << (step - 1)
            i1 = gid + (gid // ps) * ps
            _swap_step(d, pad, i1, i1 + ps, desc, key, compare)
    return d[:n]


def gselect_rank(data, vals=None, key=lambda x: x, compare=np.greater):
    """gselect.cl:38-57 with a key getter and a compare:
    pos(g) = #{i : COMPARE(key_g, key_i) || (key_i == key_g && i < g)},
    out[pos(g)] = data[g].  Ranks that collide or fall outside 0..n-1 (a
    compare that is not a strict order) raise instead of racing."""
    d = np.asarray(data)
    n = d.size
    k = key(d)
    pos = np.empty(n, dtype=np.int64)
    idx = np.arange(n)
    for g0 in range(0, n, 512):
        g1 = min(n, g0 + 512)
        kg = k[g0:g1, None]
        c = np.asarray(compare(kg, k[None, :]), dtype=bool) | (
            (k[None, :] == kg) & (idx[None, :] < idx[g0:g1, None]))
        pos[g0:g1] = c.sum(axis=1)
    if n and (pos.max() >= n or np.unique(pos).size != n):
        raise ValueError("ranks are not a permutation")
    out = np.empty_like(d)
    out[pos] = d
    if vals is None:
        return out
    vo = np.empty_like(np.asarray(vals))
    vo[pos] = vals
    return out, vo


def scan(values, sum_dtype):
    a = np.ascontiguousarray(values)
    out = np.empty(a.size, dtype=sum_dtype)
    lib().orc_scan(_p(a), a.dtype.itemsize, int(a.dtype.kind == "i"), _p(out),
                   out.dtype.itemsize, a.size)
    return out


def scan_blelloch(values_u32, sum_dtype=np.uint32, lws=256, threads=1):
    a = np.ascontiguousarray(values_u32, dtype=np.uint32)
    out = np.empty(a.size, dtype=sum_dtype)
    lib().orc_scan_blelloch(_p(a), _p(out), out.dtype.itemsize, a.size, lws, threads)
    return out


def _gen_id(name):
    return RNG_NAMES.index(name)


def rng_init_dev_gid(name, count, main_seed=0, hash_id=0):
    st = np.zeros(count * STATE_SIZE[name], dtype=np.uint8)
    lib().orc_rng_init_dev_gid(_gen_id(name), count, main_seed, hash_id, _p(st))
    return st


def rng_init_host_mt(name, count, main_seed=0):
    st = np.zeros(count * STATE_SIZE[name], dtype=np.uint8)
    lib().orc_rng_init_host_mt(_gen_id(name), count, main_seed, _p(st))
    return st


def rng_from_seed(name, seed):
    st = np.zeros(STATE_SIZE[name], dtype=np.uint8)
    lib().orc_rng_from_seed(_gen_id(name), seed & 0xFFFFFFFFFFFFFFFF, _p(st))
    return st


def rng_next(name, state):
    """One draw; `state` (uint8 array of one state) is advanced in place."""
    return lib().orc_rng_next(_gen_id(name), _p(state))


def mt19937(seed, n):
    out = np.empty(n, dtype=np.uint32)
    lib().orc_mt19937(seed & 0xFFFFFFFF, n, _p(out))
    return out


def rng_fill(name, states, draws, bits=32, maxint=0, threads=1):
    """Runs the bench kernel `draws` times; states (uint8 bytes) advance in place."""
    G = states.size // STATE_SIZE[name]
    out = np.empty(draws * G, dtype=np.uint32)
    lib().orc_rng_fill(_gen_id(name), _p(states), G, draws, bits, maxint, _p(out), threads)
    return out


# --------------------------------------------------------------------------
# FP64 restatement of the dispatcher's cost model (n_in -> 32 -> 32 -> 1,
# ReLU), its mean-squared-error gradient and Adam, for checking the GPU
# kernels.  No counterpart in the reference (the model is an addition); pinned
# in tests/test_oracle.py by central finite differences and closed forms.

MLP_H = 32


def mlp_layout(n_in):
    """Offsets of W1[32][n_in], b1, W2[32][32], b2, w3, b3 in the flat array."""
    h = MLP_H
    o = {"w1": 0, "b1": h * n_in}
    o["w2"] = o["b1"] + h
    o["b2"] = o["w2"] + h * h
    o["w3"] = o["b2"] + h
    o["b3"] = o["w3"] + h
    o["count"] = o["b3"] + 1
    return o


def _mlp_unpack(p, n_in):
    o, h = mlp_layout(n_in), MLP_H
    p = np.asarray(p, dtype=np.float64)
    return (p[o["w1"]:o["b1"]].reshape(h, n_in), p[o["b1"]:o["w2"]],
            p[o["w2"]:o["b2"]].reshape(h, h), p[o["b2"]:o["w3"]], p[o["w3"]:o["b3"]],
            p[o["b3"]])


def mlp_forward(p, X):
    X = np.asarray(X, dtype=np.float64)
    W1, b1, W2, b2, w3, b3 = _mlp_unpack(p, X.shape[1])
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    return h2 @ w3 + b3


def mlp_loss(p, X, t):
    e = mlp_forward(p, X) - np.asarray(t, dtype=np.float64)
    return 0.5 * float(np.mean(e * e))


def mlp_grad(p, X, t):
    """(loss, dloss/dp) of the mean 0.5 (y - t)^2, backpropagated in FP64."""
    X = np.asarray(X, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    n_in, rows = X.shape[1], X.shape[0]
    W1, b1, W2, b2, w3, b3 = _mlp_unpack(p, n_in)
    z1 = X @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    y = h2 @ w3 + b3
    e = y - t
    dy = e / rows
    d2 = np.outer(dy, w3) * (z2 > 0)
    d1 = (d2 @ W2) * (z1 > 0)
    o = mlp_layout(n_in)
    g = np.zeros(o["count"])
    g[o["w1"]:o["b1"]] = (d1.T @ X).ravel()
    g[o["b1"]:o["w2"]] = d1.sum(0)
    g[o["w2"]:o["b2"]] = (d2.T @ h1).ravel()
    g[o["b2"]:o["w3"]] = d2.sum(0)
    g[o["w3"]:o["b3"]] = dy @ h2
    g[o["b3"]] = dy.sum()
    return 0.5 * float(np.mean(e * e)), g


def adam_step(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
