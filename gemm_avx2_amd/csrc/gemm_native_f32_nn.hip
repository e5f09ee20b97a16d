/*
 * gemm_native_f32_nn.hip -- native FP32 tile kernels, transa N, transb N
 * (every tile configuration, one pass and split-K); see gemm_native.hip.
 */
#include "gemm_native_impl.h"

GEMMX_NATIVE_INSTANCE(float, false, false, gemmx_native_f32_nn)
