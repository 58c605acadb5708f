"""RNG seeds and fills on the GPU vs the oracle, every stream word."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GENS = ["lcg", "xorshift64", "xorshift128", "mwc64x", "parkmiller", "tauslcg"]


@pytest.mark.parametrize("gen", GENS)
@pytest.mark.parametrize("hash_name,hash_id", [(None, 0), ("KNUTH(x)", 1), ("XS1(x)", 2)])
def test_dev_gid_seeds(clops, oracle, ctx, queue, gen, hash_name, hash_id):
    count = 10000
    r = clops.Rng(ctx, gen, clops.SEED_DEV_GID, count, main_seed=1234, hash=hash_name, q=queue)
    queue.finish()
    st = r.read_states(queue)
    assert np.array_equal(st, oracle.rng_init_dev_gid(gen, count, 1234, hash_id))


@pytest.mark.parametrize("gen", GENS)
@pytest.mark.parametrize("seed", [0, 1234, (1 << 40) + 7])
def test_host_mt_seeds(clops, oracle, ctx, queue, gen, seed):
    count = 3000
    r = clops.Rng(ctx, gen, clops.SEED_HOST_MT, count, main_seed=seed)
    assert np.array_equal(r.read_states(queue), oracle.rng_init_host_mt(gen, count, seed))


@pytest.mark.parametrize("gen", GENS)
@pytest.mark.parametrize("bits,maxint", [(32, 0), (17, 0), (1, 0), (32, 1000), (32, 7)])
def test_fill_matches_oracle(clops, oracle, ctx, queue, gen, bits, maxint):
    G, draws = 4099, 16
    r = clops.Rng(ctx, gen, clops.SEED_HOST_MT, G, main_seed=99)
    st = r.read_states(queue)
    out = clops.Buffer(ctx, G * draws * 4)
    r.fill(queue, out, draws, bits, maxint)
    got = out.read(queue, np.uint32, G * draws)
    ref = oracle.rng_fill(gen, st, draws, bits, maxint)
    assert np.array_equal(got, ref)
    # the device states advanced exactly `draws` steps
    assert np.array_equal(r.read_states(queue), st)


@pytest.mark.parametrize("gen", GENS)
def test_fill_resume_ext_host(clops, oracle, ctx, queue, gen):
    """States read back and re-seeded with EXT_HOST continue the stream."""
    G = 2048
    r = clops.Rng(ctx, gen, clops.SEED_DEV_GID, G, main_seed=5, q=queue)
    out = clops.Buffer(ctx, G * 8 * 4)
    r.fill(queue, out, 4)
    st = r.read_states(queue)
    r2 = clops.Rng(ctx, gen, clops.SEED_EXT_HOST, G, seeds=st)
    r.fill(queue, out, 4)
    a = out.read(queue, np.uint32, G * 4)
    r2.fill(queue, out, 4)
    b = out.read(queue, np.uint32, G * 4)
    assert np.array_equal(a, b)


def test_ext_dev_shares_buffer(clops, oracle, ctx, queue):
    G = 1000
    st = oracle.rng_init_dev_gid("xorshift128", G, 42, 0)
    buf = clops.Buffer.from_array(ctx, st)
    r = clops.Rng(ctx, "xorshift128", clops.SEED_EXT_DEV, G, seeds=buf)
    out = clops.Buffer(ctx, G * 4 * 3)
    r.fill(queue, out, 3)
    assert np.array_equal(out.read(queue, np.uint32, 3 * G),
                          oracle.rng_fill("xorshift128", st.copy(), 3))


def test_config2_fills_prefix(clops, oracle, ctx, queue):
    """Configuration 2 layout (run-major) at 2^20 work-items x 16 draws."""
    G, draws = 1 << 20, 16
    for gen in ("xorshift128", "mwc64x"):
        r = clops.Rng(ctx, gen, clops.SEED_HOST_MT, G, main_seed=0)
        st = r.read_states(queue)
        out = clops.Buffer(ctx, G * draws * 4)
        r.fill(queue, out, draws)
        got = out.read(queue, np.uint32, G * draws)
        assert np.array_equal(got, oracle.rng_fill(gen, st, draws, threads=8))


M64 = (1 << 64) - 1


def _xs1(s):
    """clo_rng_init.cl's XS1 on a 64-bit seed."""
    s = (((s >> 16) ^ s) * 0x45d9f3b) & M64
    s = (((s >> 16) ^ s) * 0x45d9f3b) & M64
    return (s >> 16) ^ s


# Seed hash snippets beyond the predefined ones, compiled at run time, with
# their meaning on a Python int.  The reference applies the snippet as the
# statement CLO_RNG_HASH(seed); (clo_rng_init.cl:54-55), so an expression
# without an assignment leaves the seed unchanged.
CUSTOM_HASHES = [
    ("x ^= x >> 33; x *= 0xff51afd7ed558ccdUL",
     lambda s: ((s ^ (s >> 33)) * 0xff51afd7ed558ccd) & M64),
    ("KNUTH(x); x += 7", lambda s: ((s * 2654435761) % (1 << 32)) + 7),
    ("XS1(x) x = ~x", lambda s: ~_xs1(s) & M64),
    ("(x * 3 + 1)", lambda s: s),
]


@pytest.mark.parametrize("gen", GENS)
@pytest.mark.parametrize("which", range(len(CUSTOM_HASHES)))
def test_dev_gid_custom_hash(clops, oracle, ctx, queue, gen, which):
    """DEV_GID seeds with a hash snippet other than KNUTH/XS1/none: the
    snippet is compiled with the generator's source and applied to
    gid + main_seed, then converted as the predefined ones are."""
    text, fn = CUSTOM_HASHES[which]
    count, main_seed = 2000, (1 << 40) + 99
    r = clops.Rng(ctx, gen, clops.SEED_DEV_GID, count, main_seed=main_seed, hash=text, q=queue)
    queue.finish()
    want = np.concatenate([oracle.rng_from_seed(gen, fn(g + main_seed)) for g in range(count)])
    assert np.array_equal(r.read_states(queue), want)


def test_custom_hash_equal_to_predefined(clops, oracle, ctx, queue):
    """A snippet spelling out KNUTH gives the predefined KNUTH seeds."""
    r = clops.Rng(ctx, "xorshift128", clops.SEED_DEV_GID, 5000, main_seed=3,
                  hash="x = ((x * 2654435761UL) % 0x100000000UL)", q=queue)
    queue.finish()
    assert np.array_equal(r.read_states(queue),
                          oracle.rng_init_dev_gid("xorshift128", 5000, 3, 1))


def test_errors(clops, ctx):
    with pytest.raises(clops.ClopsError) as e:
        clops.Rng(ctx, "mersenne", clops.SEED_DEV_GID, 10)
    assert e.value.code == clops.CLO_ERROR_IMPL_NOT_FOUND
    with pytest.raises(clops.ClopsError) as e:
        clops.Rng(ctx, "lcg", clops.SEED_DEV_GID, 10, hash="x = y")
    assert e.value.code == clops.CLO_ERROR_ARGS and "did not compile" in str(e.value)


def test_source_text(clops, ctx):
    r = clops.Rng(ctx, "mwc64x", clops.SEED_HOST_MT, 4)
    src = r.source()
    assert "namespace clo_rng_mwc64x" in src and "4294883355U" in src
    assert "clo_rng_next_int" in src


def _hiprtc_kernel(src, name):
    """Compiles `src` with hipRTC for the current GPU and returns
    (module, function) handles (ctypes voidp)."""
    import ctypes
    rtc = ctypes.CDLL("libhiprtc.so")
    hip = ctypes.CDLL("libamdhip64.so")
    prog = ctypes.c_void_p()
    assert rtc.hiprtcCreateProgram(ctypes.byref(prog), src.encode(), b"user.hip", 0, None,
                                   None) == 0
    opts = (ctypes.c_char_p * 1)(b"-O3")  # target: the current device
    rc = rtc.hiprtcCompileProgram(prog, 1, opts)
    if rc != 0:
        sz = ctypes.c_size_t()
        rtc.hiprtcGetProgramLogSize(prog, ctypes.byref(sz))
        log = ctypes.create_string_buffer(sz.value + 1)
        rtc.hiprtcGetProgramLog(prog, log)
        raise AssertionError("hipRTC failed: " + log.value.decode(errors="replace")[-2000:])
    sz = ctypes.c_size_t()
    assert rtc.hiprtcGetCodeSize(prog, ctypes.byref(sz)) == 0
    code = ctypes.create_string_buffer(sz.value)
    assert rtc.hiprtcGetCode(prog, code) == 0
    mod, fn = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipModuleLoadData(ctypes.byref(mod), code) == 0
    assert hip.hipModuleGetFunction(ctypes.byref(fn), mod, name.encode()) == 0
    return hip, mod, fn, code


@pytest.mark.parametrize("gen", ["xorshift128", "tauslcg", "mwc64x"])
def test_source_serves_user_kernels(clops, oracle, ctx, queue, gen):
    """clo_rng_get_source() is what a user kernel compiles against (the
    reference hands it to OpenCL programs): a hipRTC kernel drawing with
    clo_rng_next_int4 gives the oracle's words."""
    import ctypes
    G, nmod = 1024, 1000003
    r = clops.Rng(ctx, gen, clops.SEED_DEV_GID, 4 * G, main_seed=77)
    st = r.read_states(queue)
    src = r.source() + """
extern "C" __global__ void user_draw(clo_statetype* st, unsigned int* out, unsigned int n) {
	const unsigned int g = CLO_RNG_GID1();
	const uint4 v = clo_rng_next_int4(st, n);
	out[4 * g] = v.x;
	out[4 * g + 1] = v.y;
	out[4 * g + 2] = v.z;
	out[4 * g + 3] = v.w;
}
"""
    hip, mod, fn, _code = _hiprtc_kernel(src, "user_draw")
    out = clops.Buffer(ctx, 4 * G * 4)
    L = clops.lib()
    p_st = ctypes.c_void_p(L.ccl_buffer_get_device_ptr(L.clo_rng_get_device_seeds(r.ptr)))
    p_out = ctypes.c_void_p(L.ccl_buffer_get_device_ptr(out.ptr))
    n = ctypes.c_uint(nmod)
    args = (ctypes.c_void_p * 3)(ctypes.cast(ctypes.byref(p_st), ctypes.c_void_p),
                                 ctypes.cast(ctypes.byref(p_out), ctypes.c_void_p),
                                 ctypes.cast(ctypes.byref(n), ctypes.c_void_p))
    queue.finish()
    assert hip.hipModuleLaunchKernel(fn, G // 256, 1, 1, 256, 1, 1, 0, None, args, None) == 0
    assert hip.hipDeviceSynchronize() == 0
    got = out.read(queue, np.uint32, 4 * G).reshape(G, 4)
    ss = oracle.STATE_SIZE[gen]
    per = st.reshape(4 * G, ss)
    want = np.zeros((G, 4), dtype=np.uint32)
    for g in range(G):
        for k in range(4):
            s = per[k * G + g].copy()
            want[g, k] = oracle.rng_next(gen, s) % nmod
    assert np.array_equal(got, want)
    hip.hipModuleUnload(mod)


@pytest.mark.parametrize("gen", GENS)
@pytest.mark.parametrize("hash_name,hash_id", [(None, 0), ("XS1(x)", 2)])
def test_dev_gid_first_fill_computes_seeds(clops, oracle, ctx, queue, gen, hash_name, hash_id):
    """DEV_GID seeds are computed by the first fill (never read from memory
    before it): the draws and the states after it equal the oracle's from
    the seeded states, and a second fill continues the stream."""
    G, draws = 5003, 7
    r = clops.Rng(ctx, gen, clops.SEED_DEV_GID, G, main_seed=77, hash=hash_name, q=queue)
    st = oracle.rng_init_dev_gid(gen, G, 77, hash_id)
    out = clops.Buffer(ctx, G * draws * 4)
    for rep in range(2):
        r.fill(queue, out, draws)
        ref = oracle.rng_fill(gen, st, draws)  # advances st in place
        assert np.array_equal(out.read(queue, np.uint32, G * draws), ref), rep
    assert np.array_equal(r.read_states(queue), st)


def test_dev_gid_seeds_initialised_before_capture(clops, oracle, ctx):
    """A graph captured over the first fill of DEV_GID seeds does not
    re-seed on replay: the seeds are initialised when the capture begins,
    and every replay continues the stream."""
    q = clops.Queue(ctx)
    G, draws = 4096, 4
    r = clops.Rng(ctx, "xorshift128", clops.SEED_DEV_GID, G, main_seed=3, q=q)
    st = oracle.rng_init_dev_gid("xorshift128", G, 3, 0)
    out = clops.Buffer(ctx, G * draws * 4)
    g = clops.Graph(q)
    g.begin()
    try:
        r.fill(q, out, draws)
    finally:
        g.end()
    q.finish()
    for rep in range(3):
        g.launch()
        q.finish()
        ref = oracle.rng_fill("xorshift128", st, draws)
        assert np.array_equal(out.read(q, np.uint32, G * draws), ref), rep
    g.close()
    assert np.array_equal(r.read_states(q), st)
