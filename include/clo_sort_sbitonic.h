/*
 * clo_sort_sbitonic.h - simple bitonic sort: one kernel launch per network
 * step, as CL_Ops' src/cl_ops/sort/clo_sort_sbitonic.{in.h:30-36,c:73-118}.
 * Not stable; keys-only.  Options are ignored.
 */
#ifndef CLO_SORT_SBITONIC_H_
#define CLO_SORT_SBITONIC_H_

#include "clo_sort_abstract.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_SORT_SBITONIC_KNAME "sbitonic"

extern const CloSortImplDef clo_sort_sbitonic_def;

#ifdef __cplusplus
}
#endif
#endif
