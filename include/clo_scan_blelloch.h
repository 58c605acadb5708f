/*
 * clo_scan_blelloch.h - exclusive scan registered as "blelloch"
 * (src/cl_ops/scan/clo_scan_blelloch.in.h:30-46).  On MI355X, plain scans of
 * up to 2^21 unsigned 32/64-bit elements on 16-byte aligned buffers run one
 * kernel with decoupled look-back (blelloch_onepass: one launch instead of
 * three).  Larger scans, other element types and segmented scans run three
 * kernels like the reference's (per-tile sums, a scan of the tile sums, the
 * tile scans with their carry), ordered reduce-then-scan: at HBM sizes reading
 * the input twice is faster than waiting on device-coherent look-back.  Any option fails with
 * CLO_ERROR_ARGS, as in the reference (clo_scan_blelloch.c:43-45).
 */
#ifndef CLO_SCAN_BLELLOCH_H_
#define CLO_SCAN_BLELLOCH_H_

#include "clo_scan_abstract.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_SCAN_BLELLOCH_NUM_KERNELS 4
#define CLO_SCAN_BLELLOCH_KNAME_REDUCE "blelloch_tile_reduce"
#define CLO_SCAN_BLELLOCH_KNAME_TILES "blelloch_tiles_scan"
#define CLO_SCAN_BLELLOCH_KNAME_SCAN "blelloch_tile_scan"
#define CLO_SCAN_BLELLOCH_KNAME_ONEPASS "blelloch_onepass"
#define CLO_SCAN_BLELLOCH_KERNELNAMES { CLO_SCAN_BLELLOCH_KNAME_REDUCE, \
	CLO_SCAN_BLELLOCH_KNAME_TILES, CLO_SCAN_BLELLOCH_KNAME_SCAN, \
	CLO_SCAN_BLELLOCH_KNAME_ONEPASS }

extern const CloScanImplDef clo_scan_blelloch_def;

/* Test hook (no counterpart in the reference): the single-pass kernel's
 * look-back waits `polls` rounds for an unpublished predecessor tile before
 * summing that tile's input itself; 0 sums every unpublished predecessor at
 * once, a negative value restores the default (4096).  The result is the same
 * for any value.  Applies to launches made after the call. */
void clo_scan_debug_lookback_polls(int polls);

#ifdef __cplusplus
}
#endif
#endif
