/*
 * clo_sort_satradix.h - stable LSD radix sort ("satradix").
 *
 * Names follow CL_Ops' src/cl_ops/sort/clo_sort_satradix.in.h:31-55.  The
 * MI355X kernels differ from the reference's three (localsort, histogram,
 * scatter + a Blelloch scan per pass): one histogram kernel for every digit
 * position, one digit-offset kernel, and one fused rank+scatter kernel per
 * digit pass with decoupled look-back for the digit offsets.
 *
 * Options ("key=value,..."): radix=<power of two, 2..256> (default 256),
 * scan=<impl> and scan<opt>=<value> (accepted for compatibility; the digit
 * offsets are computed inside the scatter pass).
 */
#ifndef CLO_SORT_SATRADIX_H_
#define CLO_SORT_SATRADIX_H_

#include "clo_sort_abstract.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_SORT_SATRADIX_NUM_KERNELS 4
#define CLO_SORT_SATRADIX_KIDX_HISTOGRAM 0
#define CLO_SORT_SATRADIX_KIDX_OFFSETS 1
#define CLO_SORT_SATRADIX_KIDX_SCATTER 2
#define CLO_SORT_SATRADIX_KIDX_COPY 3
#define CLO_SORT_SATRADIX_KNAME_HISTOGRAM "satradix_histogram"
#define CLO_SORT_SATRADIX_KNAME_OFFSETS "satradix_offsets"
#define CLO_SORT_SATRADIX_KNAME_SCATTER "satradix_scatter"
#define CLO_SORT_SATRADIX_KNAME_COPY "satradix_copy"
#define CLO_SORT_SATRADIX_KERNELNAMES { \
	CLO_SORT_SATRADIX_KNAME_HISTOGRAM, CLO_SORT_SATRADIX_KNAME_OFFSETS, \
	CLO_SORT_SATRADIX_KNAME_SCATTER, CLO_SORT_SATRADIX_KNAME_COPY }

/* clo_sort_satradix.c:678-687 */
extern const CloSortImplDef clo_sort_satradix_def;

#ifdef __cplusplus
}
#endif
#endif
