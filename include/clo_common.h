/*
 * clo_common.h - common types, error codes and bit utilities.
 *
 * Same names, values and signatures as CL_Ops'
 * src/cl_ops/common/clo_common.in.h:32-165.
 */
#ifndef CLO_COMMON_H_
#define CLO_COMMON_H_

#include "clo_ccl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Default RNG seed (clo_common.in.h:32). */
#define CLO_DEFAULT_SEED 0

/* Error domain "clo-error-quark" (clo_common.c:221-223). */
#define CLO_ERROR clo_error_quark()

/* Helper macros (clo_common.in.h:53-70). */
#define CLO_DIV_CEIL(a, b) ((a + b - 1) / b)
#define CLO_GWS_MULT(gws, lws) (lws * CLO_DIV_CEIL(gws, lws))
#define CLO_IS_PO2(x) (((x) & ((x) - 1)) == 0)

/* Error codes (clo_common.in.h:80-95). */
enum clo_error_codes {
	CLO_SUCCESS = 0,
	CLO_ERROR_OPENFILE = 1,
	CLO_ERROR_ARGS = 2,
	CLO_ERROR_STREAM_WRITE = 3,
	CLO_ERROR_IMPL_NOT_FOUND = 5,
	CLO_ERROR_UNKNOWN_TYPE = 6,
	CLO_ERROR_LIBRARY = 7
};

/* OpenCL scalar types (clo_common.in.h:108-120). */
typedef enum {
	CLO_CHAR = 0,
	CLO_UCHAR = 1,
	CLO_SHORT = 2,
	CLO_USHORT = 3,
	CLO_INT = 4,
	CLO_UINT = 5,
	CLO_LONG = 6,
	CLO_ULONG = 7,
	CLO_HALF = 8,
	CLO_FLOAT = 9,
	CLO_DOUBLE = 10
} CloType;

typedef struct clo_type_info CloTypeInfo;
typedef struct clo_sort CloSort;
typedef struct clo_scan CloScan;
typedef struct clo_rng CloRng;

/* clo_common.c:76-124 */
const char* clo_type_get_name(CloType type);
size_t clo_type_sizeof(CloType type);
CloType clo_type_by_name(const char* name, GError** err);

/* clo_common.c:141-198 */
unsigned int clo_nlpo2(unsigned int x);
unsigned int clo_ones32(unsigned int x);
unsigned int clo_tzc(int x);
unsigned int clo_sum(unsigned int x);

/* Debug log of the library's decisions on stderr, one "clo: ..." line per
 * call (the chosen kernel path and its sizes), as CL_Ops logs them with
 * g_debug (clo_sort_satradix.c:171,192): 0 off, 1 per call, 2 also every
 * kernel launch; -1 restores the default, the CLO_LOG environment variable
 * (0 when unset).  No counterpart function in CL_Ops (an addition). */
void clo_set_log_level(int level);
int clo_get_log_level(void);

/* clo_common.c:206-208: a print handler that discards its input. */
void clo_print_to_null(const gchar* string);

/* clo_common.c:221-223 */
GQuark clo_error_quark(void);

#ifdef __cplusplus
}
#endif

#endif /* CLO_COMMON_H_ */
