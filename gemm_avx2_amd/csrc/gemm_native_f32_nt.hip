/*
 * gemm_native_f32_nt.hip -- native FP32 tile kernels, transa N, transb T
 * (every tile configuration, one pass and split-K); see gemm_native.hip.
 */
#include "gemm_native_impl.h"

GEMMX_NATIVE_INSTANCE(float, false, true, gemmx_native_f32_nt)
