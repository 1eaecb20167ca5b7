/*
 * kblas_defs.h -- character enums and status codes of the KBLAS C API.
 *
 * Same names and values as the reference (include/kblas_defs.h:24-49) so
 * callers that test `status == KBLAS_Success` (which is 1, not 0) keep working.
 */
#ifndef KBLAS_DEFS_H_
#define KBLAS_DEFS_H_

#define KBLAS_Lower 'L'
#define KBLAS_Upper 'U'
#define KBLAS_Left 'L'
#define KBLAS_Right 'R'
#define KBLAS_Trans 'T'
#define KBLAS_NoTrans 'N'
#define KBLAS_Unit 'U'
#define KBLAS_NonUnit 'N'
#define KBLAS_Symm 'S'
#define KBLAS_NonSymm 'N'

#define KBLAS_Success 1
#define KBLAS_UnknownError 0
#define KBLAS_NotSupported -1
#define KBLAS_NotImplemented -2
#define KBLAS_cuBLAS_Error -3
#define KBLAS_WrongConfig -4
#define KBLAS_CUDA_Error -5
#define KBLAS_InsufficientWorkspace -6
#define KBLAS_Error_Allocation -7
#define KBLAS_Error_Deallocation -8
#define KBLAS_Error_NotInitialized -9
#define KBLAS_Error_WrongInput -10
#define KBLAS_MAGMA_Error -11
#define KBLAS_SVD_NoConvergence -12

/* The HIP runtime is the device API here; the reference's CUDA-named code
 * keeps its value and meaning (a device-runtime call failed). */
#define KBLAS_HIP_Error KBLAS_CUDA_Error

#endif /* KBLAS_DEFS_H_ */
