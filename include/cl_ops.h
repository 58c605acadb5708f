/*
 * cl_ops.h - umbrella header (CL_Ops' src/cl_ops/cl_ops.h:33-48).
 */
#ifndef CL_OPS_H_
#define CL_OPS_H_

#include "clo_common.h"
#include "clo_rng.h"
#include "clo_sort_abstract.h"
#include "clo_sort_abitonic.h"
#include "clo_sort_sbitonic.h"
#include "clo_sort_gselect.h"
#include "clo_sort_satradix.h"
#include "clo_scan_abstract.h"
#include "clo_scan_blelloch.h"
#include "clo_util.h"
#include "clo_graph.h"
#include "clo_mlp.h"
#include "clo_dispatch.h"

#endif
