/*
 * gemm_native_f64_tt.hip -- native FP64 tile kernels, transa T, transb T
 * (every tile configuration, one pass and split-K); see gemm_native.hip.
 */
#include "gemm_native_impl.h"

GEMMX_NATIVE_INSTANCE(double, true, true, gemmx_native_f64_tt)
