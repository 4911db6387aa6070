/*
 * trid_common.h - status codes, dimension limit and linkage macro shared by
 * every tridsolver_amd entry point.
 *
 * Layout/value compatibility with the reference:
 *   - tridStatus_t keeps the reference's ten values 0..9
 *     (reference: scalar/include/trid_common.h:63-75).
 *   - MAXDIM = 3 (reference: scalar/include/trid_common.h:54).
 *   - EXTERN_C expands to extern "C" under C++ (reference:
 *     scalar/include/trid_common.h:57-61).
 *
 * Unlike the reference, which always returns TRID_STATUS_SUCCESS and calls
 * exit(-1) on device errors, this library reports:
 *   TRID_STATUS_INVALID_VALUE     bad pointer / ndim / solvedim / dims / pads
 *   TRID_STATUS_ALLOC_FAILED      a workspace allocation failed
 *   TRID_STATUS_ARCH_MISMATCH     the current device is not gfx950
 *   TRID_STATUS_EXECUTION_FAILED  a HIP launch or synchronisation failed
 *   TRID_STATUS_NOT_INITIALIZED   no HIP device is usable
 *   TRID_STATUS_INTERNAL_ERROR    a workspace is needed inside stream capture
 *                                 but was not reserved beforehand
 */
#ifndef TRID_COMMON_H__
#define TRID_COMMON_H__

#define MAXDIM 3

#ifdef __cplusplus
#  define EXTERN_C extern "C"
#else
#  define EXTERN_C
#endif

typedef enum {
  TRID_STATUS_SUCCESS                   = 0,
  TRID_STATUS_NOT_INITIALIZED           = 1,
  TRID_STATUS_ALLOC_FAILED              = 2,
  TRID_STATUS_INVALID_VALUE             = 3,
  TRID_STATUS_ARCH_MISMATCH             = 4,
  TRID_STATUS_MAPPING_ERROR             = 5,
  TRID_STATUS_EXECUTION_FAILED          = 6,
  TRID_STATUS_INTERNAL_ERROR            = 7,
  TRID_STATUS_MATRIX_TYPE_NOT_SUPPORTED = 8,
  TRID_STATUS_ZERO_PIVOT                = 9
} tridStatus_t;

#endif
