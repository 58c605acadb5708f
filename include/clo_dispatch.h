/*
 * clo_dispatch.h - per-request choice of kernel configuration for sorts and
 * scans: an exhaustive autotuner, its persisted per-bucket table, and a
 * learned cost model (clo_mlp.h).  No counterpart in CL_Ops (an addition);
 * every configuration computes the same bit-exact result, so the choice only
 * changes the time taken.
 *
 * Requests: sort (keys only, ascending, unsigned 32- or 64-bit keys, out of
 * place in -> out), scan (exclusive, element = sum type = the same unsigned
 * width, in -> out) or segmented sort (keys only, nseg uint32 offsets as in
 * clo_sort_segmented_with_device_data, out of place).
 *
 * Configurations (index: name, op, feasibility):
 *   0 satradix 256x32        sort     any n
 *   1 satradix 512x16        sort     any n
 *   2 satradix 256x16        sort     any n
 *   3 abitonic               sort     any n
 *   4 segment-local sort     sort     n <= 65536 (one segment, LDS-resident)
 *   5 gselect                sort     n <= 4096 (rank sort)
 *   6 blelloch 8 items       scan     any n
 *   7 blelloch 16 items      scan     any n (the plain clo_scan path)
 *   8 blelloch 32 items      scan     any n
 *   9 satradix 4-bit digits  sort     any n (radix=16, default tile shape)
 *  10 satradix 6-bit digits  sort     any n (radix=64)
 *  11 satradix 7-bit digits  sort     any n (radix=128); 8-bit digits are 0-2
 *  12 + 4 t + (g - 2)      segsort  segmented tile 256 x {8, 16, 32}[t]
 *                                     items, g = 2..5 blocks per CU in the
 *                                     persistent grid; tile 32 for 32-bit
 *                                     keys only
 * The fixed configuration is what the plain API runs: satradix's default
 * tile shape (2 up to 2^20 keys, else 0 for 32-bit and 1 for 64-bit keys),
 * 7 for scans, and the segmented default (the tuner's choice on 32768
 * log-uniform segments: tile 16, 4 blocks per CU, 18).
 *
 * Features (CLO_DISPATCH_NUM_FEATURES floats, from a GPU kernel over a
 * sample of up to 4096 evenly spaced elements and segments):
 *   0 log2(n) / 32                 5 fraction of sampled neighbour pairs
 *   1 key bytes / 8                  (x[i], x[i+1]) in ascending order
 *   2 mean key-byte entropy / 8    6 fraction in descending order
 *   3 max key-byte entropy / 8     7 op (0 sort, 1 scan, 2 segmented sort)
 *   4 fraction of constant         8 mean log2(segment length + 1) / 16
 *     key bytes                    9 share of keys in segments longer
 *                                    than 4096 (sort and scan: one
 *                                    segment of n)
 * Model input row (CLO_DISPATCH_MODEL_INPUTS floats) = the features followed
 * by a one-hot of the configuration; the model predicts a log2 time (absolute,
 * or relative to the fixed configuration of the same request: only the
 * order of one request's predictions matters), and the choice is the feasible
 * configuration with the least prediction.
 */
#ifndef CLO_DISPATCH_H_
#define CLO_DISPATCH_H_

#include "clo_ccl.h"
#include "clo_mlp.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_DISPATCH_NUM_CONFIGS 24
#define CLO_DISPATCH_NUM_FEATURES 10
#define CLO_DISPATCH_MODEL_INPUTS (CLO_DISPATCH_NUM_FEATURES + CLO_DISPATCH_NUM_CONFIGS)
/* Per-segment work classes: length bins and the segment model's inputs. */
#define CLO_DISPATCH_SEG_BINS 64
#define CLO_DISPATCH_SEG_INPUTS 7

typedef enum { CLO_OP_SORT = 0, CLO_OP_SCAN = 1, CLO_OP_SEGSORT = 2 } CloOp;

typedef enum {
	CLO_POLICY_FIXED = 0,  /* the plain path's configuration */
	CLO_POLICY_TABLE = 1,  /* the tuner table's best for the request's bucket, else fixed */
	CLO_POLICY_MODEL = 2   /* the learned model's choice, else the table, else fixed */
} CloPolicy;

typedef struct clo_dispatch CloDispatch;

CloDispatch* clo_dispatch_new(CCLContext* ctx, GError** err);
void clo_dispatch_destroy(CloDispatch* d);

const char* clo_dispatch_config_name(cl_uint cfg);
cl_bool clo_dispatch_config_feasible(cl_uint cfg, CloOp op, size_t n, size_t width);
cl_uint clo_dispatch_fixed_config(CloOp op, size_t n, size_t width);

/* Features of a sort or scan request (blocking: the sample statistics come
 * back to the host).  width = 4 or 8 bytes. */
cl_bool clo_dispatch_features(CloDispatch* d, CCLQueue* cq, CloOp op, CCLBuffer* data,
	size_t n, size_t width, float* features, GError** err);
/* Features of `count` sort or scan requests in one launch (one block per
 * request) and one copy back: request i reads ns[i] elements of widths[i]
 * bytes from data[i]; features receives count x CLO_DISPATCH_NUM_FEATURES
 * floats, each row equal to clo_dispatch_features of that request.
 * Blocking. */
cl_bool clo_dispatch_features_batch(CloDispatch* d, CCLQueue* cq, size_t count, const int* ops,
	CCLBuffer* const* data, const size_t* ns, const size_t* widths, float* features,
	GError** err);
/* Features of a segmented sort request (op feature 2). */
cl_bool clo_dispatch_features_segmented(CloDispatch* d, CCLQueue* cq, CCLBuffer* data,
	CCLBuffer* offsets, size_t nseg, size_t n, size_t width, float* features, GError** err);
/* Model input row for one configuration. */
void clo_dispatch_model_row(const float* features, cl_uint cfg, float* row);

/* Runs configuration cfg on the request (asynchronous). */
CCLEvent* clo_dispatch_run(CloDispatch* d, CCLQueue* cq, cl_uint cfg, CloOp op, CCLBuffer* in,
	CCLBuffer* out, size_t n, size_t width, GError** err);
/* Runs segmented configuration cfg; classes (CLO_DISPATCH_SEG_BINS digits,
 * or NULL for all zero) gives the least work class of each segment length
 * bin (see clo_dispatch_segment_classes). */
CCLEvent* clo_dispatch_run_segmented(CloDispatch* d, CCLQueue* cq, cl_uint cfg, CCLBuffer* in,
	CCLBuffer* out, CCLBuffer* offsets, size_t nseg, size_t n, size_t width,
	const unsigned char* classes, GError** err);

/* Exhaustive tuner: every feasible configuration of the op runs once
 * untimed, then `reps` times timed with events on cq; times_ms[cfg] = the
 * least time (ms), -1 for configurations that are not feasible or belong to
 * the other op.  Blocking. */
cl_bool clo_dispatch_tune(CloDispatch* d, CCLQueue* cq, CloOp op, CCLBuffer* in,
	CCLBuffer* out, size_t n, size_t width, cl_uint reps, float* times_ms, GError** err);
/* The same over the segmented configurations (tile and grid). */
cl_bool clo_dispatch_tune_segmented(CloDispatch* d, CCLQueue* cq, CCLBuffer* in,
	CCLBuffer* out, CCLBuffer* offsets, size_t nseg, size_t n, size_t width, cl_uint reps,
	float* times_ms, GError** err);

/* Per-segment work class.  The segmented sort puts each segment in a work
 * class: 0 chunk (small segments of one window of CAP/2 elements sorted
 * together in LDS), 1 medium (one block, all passes in LDS), 2 large (one
 * block, passes through global memory), CAP = 256 x tile items.  By default
 * a segment takes the least class that holds it; a table of one digit per
 * length bin (four bins per octave: bin = 4 floor(log2 len) + the two bits
 * after the leading one, capped at 63) can raise it.
 * clo_dispatch_seg_bin: the bin of a length; clo_dispatch_seg_bin_length:
 * the least length in a bin (0 for bins no length falls in). */
cl_uint clo_dispatch_seg_bin(size_t len);
size_t clo_dispatch_seg_bin_length(cl_uint bin);
/* Times each feasible work class (ns per key, -1 if infeasible, 3 floats) on
 * about 2^21 device-generated keys in equal segments of len <= 65536 keys,
 * every segment forced into the class, under segmented configuration cfg.
 * Blocking. */
cl_bool clo_dispatch_tune_segment_class(CloDispatch* d, CCLQueue* cq, cl_uint cfg,
	size_t width, size_t len, cl_uint reps, float* ns_per_key, GError** err);
/* Segment model input row (CLO_DISPATCH_SEG_INPUTS floats):
 * log2(len + 1) / 16, key bytes / 8, tile items / 32, blocks per CU / 4,
 * one-hot of the class; the model predicts log2 ns per key. */
void clo_dispatch_segment_row(size_t len, size_t width, cl_uint cfg, cl_uint cls, float* row);
/* Segment model (borrowed; NULL removes it). */
cl_bool clo_dispatch_set_segment_model(CloDispatch* d, CloMlp* mlp, GError** err);
/* Class table for segmented configuration cfg from the segment model: for
 * every bin up to CAP, the class predicted cheapest at the bin's least
 * length when it beats the least feasible class by more than log2_margin,
 * else 0.  All zero without a model.  classes: CLO_DISPATCH_SEG_BINS bytes. */
cl_bool clo_dispatch_segment_classes(CloDispatch* d, CCLQueue* cq, cl_uint cfg, size_t width,
	float log2_margin, unsigned char* classes, GError** err);

/* Tuner table: buckets of (op, width, floor(log2 n), order class) keep the
 * mean log2 time of every configuration timed in them. */
void clo_dispatch_table_add(CloDispatch* d, CloOp op, size_t n, size_t width,
	const float* features, const float* times_ms);
/* Best configuration of the request's bucket, -1 if the bucket is empty. */
int clo_dispatch_table_choose(CloDispatch* d, CloOp op, size_t n, size_t width,
	const float* features);
/* Text file "CLOTUNE01" followed by one bucket entry per line. */
cl_bool clo_dispatch_table_save(CloDispatch* d, const char* path, GError** err);
cl_bool clo_dispatch_table_load(CloDispatch* d, const char* path, GError** err);

/* Learned model (borrowed; NULL removes it).  Its input count must be
 * CLO_DISPATCH_MODEL_INPUTS. */
cl_bool clo_dispatch_set_model(CloDispatch* d, CloMlp* mlp, GError** err);
/* Model policy: keep the fixed configuration unless the predicted log2 time
 * of the best one is lower by more than log2_margin (default 0). */
void clo_dispatch_set_margin(CloDispatch* d, float log2_margin);
/* Batched choice for `count` requests with one model launch over every
 * (request, feasible configuration) row.  ops/ns/widths/features
 * (count x NUM_FEATURES) describe the requests; cfg_out receives the
 * choices.  Policy as in CloPolicy. */
cl_bool clo_dispatch_choose(CloDispatch* d, CCLQueue* cq, CloPolicy policy, size_t count,
	const int* ops, const size_t* ns, const size_t* widths, const float* features, int* cfg_out,
	GError** err);

#ifdef __cplusplus
}
#endif
#endif
