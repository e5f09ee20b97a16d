/*
 * gemm_native_f64_nn.hip -- native FP64 tile kernels, transa N, transb N
 * (every tile configuration, one pass and split-K); see gemm_native.hip.
 */
#include "gemm_native_impl.h"

GEMMX_NATIVE_INSTANCE(double, false, false, gemmx_native_f64_nn)
