/*
 * clo_sort_satradix.h - stable LSD radix sort ("satradix").
 *
 * Names follow CL_Ops' src/cl_ops/sort/clo_sort_satradix.in.h:31-55.  The
 * MI355X kernels differ from the reference's three (localsort, histogram,
 * scatter + a Blelloch scan per pass): per digit pass an upsweep (tile digit
 * counts), a reduce and a two-kernel counter scan, and a downsweep that ranks
 * each tile with wave64 ballots and scatters it through LDS.
 *
 * Options ("key=value,..."): radix=<power of two, 2..256> (default 256),
 * scan=<impl> and scan<opt>=<value> (accepted for compatibility: "blelloch"
 * is the only scan, and the counter scan is built into the sort).
 */
#ifndef CLO_SORT_SATRADIX_H_
#define CLO_SORT_SATRADIX_H_

#include "clo_sort_abstract.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_SORT_SATRADIX_NUM_KERNELS 6
#define CLO_SORT_SATRADIX_KIDX_UPSWEEP 0
#define CLO_SORT_SATRADIX_KIDX_PLAN 1
#define CLO_SORT_SATRADIX_KIDX_REDUCE 2
#define CLO_SORT_SATRADIX_KIDX_SCAN 3
#define CLO_SORT_SATRADIX_KIDX_DOWNSWEEP 4
#define CLO_SORT_SATRADIX_KIDX_COPY 5
#define CLO_SORT_SATRADIX_KNAME_UPSWEEP "satradix_upsweep"
#define CLO_SORT_SATRADIX_KNAME_PLAN "satradix_plan"
#define CLO_SORT_SATRADIX_KNAME_REDUCE "satradix_reduce"
#define CLO_SORT_SATRADIX_KNAME_SCAN "satradix_scan"
#define CLO_SORT_SATRADIX_KNAME_DOWNSWEEP "satradix_downsweep"
#define CLO_SORT_SATRADIX_KNAME_COPY "satradix_copy"
#define CLO_SORT_SATRADIX_KERNELNAMES { \
	CLO_SORT_SATRADIX_KNAME_UPSWEEP, CLO_SORT_SATRADIX_KNAME_PLAN, \
	CLO_SORT_SATRADIX_KNAME_REDUCE, CLO_SORT_SATRADIX_KNAME_SCAN, \
	CLO_SORT_SATRADIX_KNAME_DOWNSWEEP, CLO_SORT_SATRADIX_KNAME_COPY }

/* clo_sort_satradix.c:678-687 */
extern const CloSortImplDef clo_sort_satradix_def;

#ifdef __cplusplus
}
#endif
#endif
