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
 *   - With FPPREC defined, FP and F name the caller's precision as in the
 *     reference (scalar/include/trid_common.h:39-47): FPPREC 0 gives
 *     FP = float and the literal suffix F = f, FPPREC 1 gives FP = double
 *     and an empty F. The reference's callers (the ADI app:
 *     apps/adi/src/adi_cuda.cu:47,88-170, preproc.hpp, compare.c,
 *     print_array.c) declare their arrays with FP. Unlike the reference,
 *     which stops with #error when FPPREC is missing, the header then simply
 *     leaves FP undefined, so callers that do not use it (and the library
 *     itself) need no FPPREC.
 *   - ALIGN, ALIGN_FLOAT, ALIGN_DOUBLE, CUDA_ALIGN_BYTE and WARP_SIZE keep the
 *     reference's values (scalar/include/trid_common.h:49-55) for callers that
 *     size allocations or tiles with them. The library does not read them: its
 *     kernels run 64-lane wavefronts and take any alignment, using 16-byte
 *     vector loads where the pointers and pitches allow.
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

#ifdef FPPREC
#  if FPPREC == 0
#    define FP float
#    define F  f
#  elif FPPREC == 1
#    define FP double
#    define F
#  else
#    error "FPPREC must be 0 (float) or 1 (double)"
#  endif
#endif

#ifndef WARP_SIZE
#  define WARP_SIZE 32
#endif
#define ALIGN           32
#define ALIGN_FLOAT     (ALIGN / 4)
#define ALIGN_DOUBLE    (ALIGN / 8)
#define CUDA_ALIGN_BYTE 32

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
