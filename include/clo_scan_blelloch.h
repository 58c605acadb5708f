/*
 * clo_scan_blelloch.h - exclusive scan registered as "blelloch"
 * (src/cl_ops/scan/clo_scan_blelloch.in.h:30-46).  On MI355X it is a
 * single-pass decoupled look-back scan (one read and one write of the data)
 * instead of the reference's three-kernel up/down-sweep.  Any option fails
 * with CLO_ERROR_ARGS, as in the reference (clo_scan_blelloch.c:43-45).
 */
#ifndef CLO_SCAN_BLELLOCH_H_
#define CLO_SCAN_BLELLOCH_H_

#include "clo_scan_abstract.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_SCAN_BLELLOCH_NUM_KERNELS 2
#define CLO_SCAN_BLELLOCH_KNAME_LOOKBACK "blelloch_lookback_scan"
#define CLO_SCAN_BLELLOCH_KNAME_SEGMENTED "blelloch_segmented_scan"

extern const CloScanImplDef clo_scan_blelloch_def;

#ifdef __cplusplus
}
#endif
#endif
