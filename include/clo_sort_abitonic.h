/*
 * clo_sort_abitonic.h - advanced bitonic sort: the bitonic network with its
 * steps fused into register-resident and LDS-resident kernels, as CL_Ops'
 * src/cl_ops/sort/clo_sort_abitonic.{in.h:33-116,c:58-567}.
 *
 * Options: minps=<1..4>, maxps=<1..4> (register steps fused per global
 * sweep; minps <= maxps), maxsfs=<n> (largest stage-finish fused in LDS, as
 * log2 of the span: at most 15, 14 for 8-byte keys; default 12 for 8-byte
 * keys, 14 for 4-byte keys, 15 otherwise), regkeys=0|16|32 (keys per thread
 * of the LDS tile kernels; default 16 for 8-byte keys, 32 otherwise).  They
 * change the launch strategy, never the result.
 */
#ifndef CLO_SORT_ABITONIC_H_
#define CLO_SORT_ABITONIC_H_

#include "clo_sort_abstract.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_SORT_ABITONIC_NUM_KERNELS 7
#define CLO_SORT_ABITONIC_KNAME_ANY "abit_any"
#define CLO_SORT_ABITONIC_KNAME_PRIV_2S4V "abit_priv_2s4v"
#define CLO_SORT_ABITONIC_KNAME_PRIV_3S8V "abit_priv_3s8v"
#define CLO_SORT_ABITONIC_KNAME_PRIV_4S16V "abit_priv_4s16v"
#define CLO_SORT_ABITONIC_KNAME_LOCAL "abit_local"
#define CLO_SORT_ABITONIC_KNAME_TILESORT "abit_tilesort"
#define CLO_SORT_ABITONIC_KNAME_PRIV_5S32V "abit_priv_5s32v"
#define CLO_SORT_ABITONIC_KERNELNAMES { CLO_SORT_ABITONIC_KNAME_ANY, \
	CLO_SORT_ABITONIC_KNAME_PRIV_2S4V, CLO_SORT_ABITONIC_KNAME_PRIV_3S8V, \
	CLO_SORT_ABITONIC_KNAME_PRIV_4S16V, CLO_SORT_ABITONIC_KNAME_LOCAL, \
	CLO_SORT_ABITONIC_KNAME_TILESORT, CLO_SORT_ABITONIC_KNAME_PRIV_5S32V }

extern const CloSortImplDef clo_sort_abitonic_def;

/* Number of global sweeps (kernel launches that read and write the whole
 * array) of the last abitonic call. */
cl_uint clo_sort_abitonic_get_last_sweeps(CloSort* sorter);

#ifdef __cplusplus
}
#endif
#endif
