/*
 * gemmx_gen.hip -- device-side input generator for benchmarks and tests
 * (libgemmx_gen.so; not part of DGEMM.so / SGEMM.so).
 *
 * Counter-based SplitMix64: the stream key is the SplitMix64 finaliser of
 * (seed, id) and element idx is the idx-th SplitMix64 output of that key, so
 * every entry is drawn independently and identically on host and device.
 * A rows x cols column-major matrix with leading dimension ld takes counter
 * index i + j*rows for element (i, j); rows..ld-1 of each column hold NaN.
 *   dist 0: U[-1,1)       FP64 (x>>11)*2^-53*2-1, FP32 (x>>40)*2^-24*2-1
 *   dist 1: (u-0.5)*exp(param*g), u ~ U[0,1), g ~ N(0,1) by Box-Muller
 *   dist 2: exp(param*g)  (log-normal)
 * dist 0 is bit-exact with any host implementation of the same rule; dist 1
 * and 2 use the device exp/log/cos and may differ from a host in the last
 * bits, so checks that need exact inputs read them back.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

__device__ __forceinline__ uint64_t sm64_fin(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t counter(uint64_t key, uint64_t idx) {
  return sm64_fin(key + (idx + 1) * 0x9E3779B97F4A7C15ull);
}

__device__ __forceinline__ double u01(uint64_t x) { return (double)(x >> 11) * 0x1p-53; }

__device__ __forceinline__ double gauss(uint64_t key, uint64_t i1, uint64_t i2) {
  const double a = 1.0 - u01(counter(key, i1));
  const double b = u01(counter(key, i2));
  return sqrt(-2.0 * log(a)) * cos(6.283185307179586 * b);
}

template <typename T>
__device__ __forceinline__ T draw(uint64_t key, uint64_t idx, int dist, double param);

template <>
__device__ __forceinline__ double draw<double>(uint64_t key, uint64_t idx, int dist, double param) {
  if (dist == 1) return (u01(counter(key, 3 * idx)) - 0.5) * exp(param * gauss(key, 3 * idx + 1, 3 * idx + 2));
  if (dist == 2) return (double)(float)exp(param * gauss(key, 2 * idx, 2 * idx + 1));
  return u01(counter(key, idx)) * 2.0 - 1.0;
}

template <>
__device__ __forceinline__ float draw<float>(uint64_t key, uint64_t idx, int dist, double param) {
  if (dist == 1) return (float)((u01(counter(key, 3 * idx)) - 0.5) * exp(param * gauss(key, 3 * idx + 1, 3 * idx + 2)));
  if (dist == 2) return (float)exp(param * gauss(key, 2 * idx, 2 * idx + 1));
  const float u = (float)(counter(key, idx) >> 40) * 0x1p-24f;
  return u * 2.0f - 1.0f;
}

template <typename T>
__global__ void __launch_bounds__(256)
fill_kernel(T *__restrict__ p, int rows, int cols, int64_t ld, uint64_t key,
            int dist, double param) {
  const int64_t total = ld * (int64_t)cols;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / ld, i = t - j * ld;
    p[t] = (i < rows) ? draw<T>(key, (uint64_t)i + (uint64_t)j * (uint64_t)rows, dist, param)
                      : (T)__builtin_nan("");
  }
}

uint64_t host_fin(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
int fill(T *p, int rows, int cols, int ld, uint64_t seed, uint64_t id, int dist,
         double param, hipStream_t s) {
  if (rows < 0 || cols < 0 || ld < rows || ld < 1) return 1;
  const uint64_t key = host_fin(seed * 0x9E3779B97F4A7C15ull + (id + 1) * 0xD1B54A32D192ED03ull);
  const int64_t total = (int64_t)ld * cols;
  if (total == 0) return 0;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL((fill_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, s,
                     p, rows, cols, (int64_t)ld, key, dist, param);
  return hipGetLastError() == hipSuccess ? 0 : 4;
}

} // namespace

extern "C" __attribute__((visibility("default"))) int
gemmx_fill_f64(double *p, int rows, int cols, int ld, uint64_t seed, uint64_t id,
               int dist, double param, void *stream) {
  return fill<double>(p, rows, cols, ld, seed, id, dist, param, (hipStream_t)stream);
}

extern "C" __attribute__((visibility("default"))) int
gemmx_fill_f32(float *p, int rows, int cols, int ld, uint64_t seed, uint64_t id,
               int dist, double param, void *stream) {
  return fill<float>(p, rows, cols, ld, seed, id, dist, param, (hipStream_t)stream);
}
