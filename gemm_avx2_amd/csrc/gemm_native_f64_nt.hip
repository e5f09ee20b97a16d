/*
 * gemm_native_f64_nt.hip -- native FP64 tile kernels, transa N, transb T
 * (every tile configuration, one pass and split-K); see gemm_native.hip.
 */
#include "gemm_native_impl.h"

GEMMX_NATIVE_INSTANCE(double, false, true, gemmx_native_f64_nt)
