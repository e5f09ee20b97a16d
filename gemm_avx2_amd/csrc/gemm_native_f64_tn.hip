/*
 * gemm_native_f64_tn.hip -- native FP64 tile kernels, transa T, transb N
 * (every tile configuration, one pass and split-K); see gemm_native.hip.
 */
#include "gemm_native_impl.h"

GEMMX_NATIVE_INSTANCE(double, true, false, gemmx_native_f64_tn)
