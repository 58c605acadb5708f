/*
 * clo_sort_gselect.h - stable O(n^2) rank ("global selection") sort, as
 * CL_Ops' src/cl_ops/sort/clo_sort_gselect.{in.h:30-36,c:31-243,cl:38-57}.
 * Out of place; with data_out == NULL the result is copied back to data_in.
 */
#ifndef CLO_SORT_GSELECT_H_
#define CLO_SORT_GSELECT_H_

#include "clo_sort_abstract.h"

#ifdef __cplusplus
extern "C" {
#endif

#define CLO_SORT_GSELECT_KNAME "gselect"

extern const CloSortImplDef clo_sort_gselect_def;

#ifdef __cplusplus
}
#endif
#endif
