// rtmc_probe.hip -- rtmc_probe_atomics: throughput of no-return global
// atomics on this GPU, the ceiling of the histogram deposits when the
// histogram lives in global memory.  Measurement aid; not part of the path.
#include "rtmc_common.hpp"

namespace rtmc {
namespace {

__device__ __forceinline__ uint64_t xorshift(uint64_t x) {
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return x;
}

// kind 0: uint32 at random words; 1: uint64 at random aligned pairs;
// 2 / 3: a random word then the next 1 / 7 words from the same lane (runs of 2 / 8);
// 4: uint64 at random aligned pairs returning the old value (the values are
// summed and the sum decides a store at the end, so every return is waited for)
__global__ __launch_bounds__(256) void k_probe(uint32_t* __restrict__ E, uint64_t words, int kind, int64_t iters) {
  uint64_t x = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 0x9E3779B97F4A7C15ull + 7u;
  const int run = kind == 2 ? 2 : 8;
  uint64_t base = 0;
  unsigned long long acc = 0;
  for (int64_t i = 0; i < iters; ++i) {
    if (kind == 4) {
      x = xorshift(x);
      acc += __hip_atomic_fetch_add((unsigned long long*)E + (x % (words / 2)), 0x100000001ull, __ATOMIC_RELAXED,
                                    __HIP_MEMORY_SCOPE_AGENT);
    } else if (kind == 0) {
      x = xorshift(x);
      __hip_atomic_fetch_add(E + (x % words), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else if (kind == 1) {
      x = xorshift(x);
      __hip_atomic_fetch_add((unsigned long long*)E + (x % (words / 2)), 0x100000001ull, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (i % run == 0) {
        x = xorshift(x);
        base = x % (words - run);
      }
      __hip_atomic_fetch_add(E + base + (i % run), 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (kind == 4 && acc == 0x5A5A5A5A5A5A5A5Aull) E[0] = 1u;  // keeps the returned values live
}

}  // namespace
}  // namespace rtmc

using namespace rtmc;

extern "C" int rtmc_probe_atomics(uint32_t* buf, int64_t words, int32_t kind, int64_t iters, double* requests_per_s,
                                  void* stream) {
  clear_error();
  RTMC_REQUIRE(buf != nullptr && requests_per_s != nullptr, "buf and requests_per_s must be given");
  RTMC_REQUIRE(words >= 16, "words must be >= 16");
  RTMC_REQUIRE(kind >= 0 && kind <= 4, "kind must be 0..4");
  RTMC_REQUIRE(iters >= 1 && iters <= 100000, "iters must be in [1, 100000]");
  hipStream_t st = (hipStream_t)stream;
  int dev = 0, ncu = 0;
  RTMC_HIP(hipGetDevice(&dev));
  RTMC_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const unsigned grid = (unsigned)ncu * 8u, block = 256u;
  hipEvent_t a, b;
  RTMC_HIP(hipEventCreate(&a));
  RTMC_HIP(hipEventCreate(&b));
  float ms = 0.f;
  for (int rep = 0; rep < 2; ++rep) {  // the first launch warms up
    RTMC_HIP(hipEventRecord(a, st));
    hipLaunchKernelGGL(k_probe, dim3(grid), dim3(block), 0, st, buf, (uint64_t)words, kind, iters);
    RTMC_HIP(hipGetLastError());
    RTMC_HIP(hipEventRecord(b, st));
    RTMC_HIP(hipEventSynchronize(b));
    RTMC_HIP(hipEventElapsedTime(&ms, a, b));
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *requests_per_s = (double)grid * block * (double)iters / (ms * 1e-3);
  return RTMC_OK;
}
