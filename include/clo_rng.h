/*
 * clo_rng.h - device-side pseudo-random number generators.
 *
 * Same API as CL_Ops' src/cl_ops/rng/clo_rng.in.h:31-111: a CloRng owns a
 * device buffer of per-work-item generator states ("seeds"), initialised in
 * one of four ways.  clo_rng_get_source() returns HIP device source of the
 * generator (the state type, clo_ulong2statetype and clo_rng_next) for
 * callers that compile their own kernels with hipRTC; the same functions are
 * shipped as a header in clo_rng_device.h.
 *
 * Generators: lcg, xorshift64, xorshift128, mwc64x, parkmiller, tauslcg (the
 * reference's clo_rng_infos order, clo_rng.c:60-68).
 */
#ifndef CLO_RNG_H_
#define CLO_RNG_H_

#include "clo_common.h"
#include <string.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_RNG_IMPLS "lcg, xorshift64, xorshift128, mwc64x, parkmiller, tauslcg"

/* clo_rng.in.h:50-65 */
struct clo_rng_info {
	const char* name;
	const char* src;
	const size_t seed_size;
};
extern const struct clo_rng_info clo_rng_infos[];

/* clo_rng.in.h:78-92 */
typedef enum clo_rng_seed_type {
	CLO_RNG_SEED_DEV_GID = 0,
	CLO_RNG_SEED_HOST_MT = 1,
	CLO_RNG_SEED_EXT_DEV = 2,
	CLO_RNG_SEED_EXT_HOST = 3
} CloRngSeedType;

/* clo_rng.c:262-405.
 *  DEV_GID : seed = gid + main_seed; CLO_RNG_HASH(seed);
 *            state[gid] = clo_ulong2statetype(seed) (clo_rng_init.cl:47-60).
 *            `hash` is the body of the CLO_RNG_HASH(x) macro, a statement
 *            on the ulong x: NULL/""/"x" (none), "KNUTH(x)" and "XS1(x)"
 *            run precompiled kernels; any other snippet (which may use KNUTH
 *            and XS1) is compiled with hipRTC, after the generator's device
 *            source, when the generator is created.  A snippet that does not
 *            compile fails with CLO_ERROR_ARGS and the compiler log.
 *  HOST_MT : word i of the state buffer = i-th output of MT19937 seeded with
 *            (uint32) main_seed (GLib g_rand_new_with_seed/g_rand_int).
 *  EXT_DEV : `seeds` is a CCLBuffer* of seeds_count states (referenced).
 *  EXT_HOST: `seeds` is a host array of seeds_count states (copied).
 * HOST_MT and EXT_HOST block; DEV_GID is asynchronous on `cq`.
 * DEV_GID with a predefined hash defers the initialisation: the first
 * clo_rng_fill computes the states from gid instead of reading them (the
 * same draws and final states; the state array is written once instead of
 * written, read and written again).  clo_rng_get_device_seeds before the
 * first fill initialises them first (blocking), and so does
 * clo_graph_begin_capture for every such RNG on the device (on the queue
 * being captured, before the capture), so a graph never re-seeds. */
CloRng* clo_rng_new(const char* type, CloRngSeedType seed_type,
	void* seeds, size_t seeds_count, cl_ulong main_seed,
	const char* hash, CCLContext* ctx, CCLQueue* cq, GError** err);
void clo_rng_destroy(CloRng* rng);
const char* clo_rng_get_source(CloRng* rng);
CCLBuffer* clo_rng_get_device_seeds(CloRng* rng);
size_t clo_rng_get_size(CloRng* rng);

/* ---- addition (no counterpart in CL_Ops' library) ------------------------
 * The fill the reference's benchmark kernel performs one launch per draw
 * (src/benchmarks/clo_rng_bench.cl:23-37), done for `draws` draws in one
 * launch with the state held in registers:
 *   out[r * G + gid] = (maxint == 0) ? next >> (32 - bits) : next % maxint
 * for r in [0, draws), gid in [0, G), G = seeds_count; bits in 1..32.
 * The states advance exactly as `draws` calls of clo_rng_next would.
 * `out` must hold draws * G uint32 words.  Asynchronous. */
CCLEvent* clo_rng_fill(CloRng* rng, CCLQueue* cq, CCLBuffer* out,
	cl_uint draws, cl_uint bits, cl_uint maxint, GError** err);

#ifdef __cplusplus
}
#endif
#endif
