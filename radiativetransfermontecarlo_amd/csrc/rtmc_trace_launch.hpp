// rtmc_trace_launch.hpp -- compile-time specialisation of k_trace: histogram
// placement, process, acoustic/elastic, with/without boundaries (the
// event-driven process has none), forward or adjoint (acoustic only).
// Included by one translation unit per dimension and flight precision.
#pragma once

#include "rtmc_trace.hpp"

namespace rtmc {

template <int DIM, bool LDSH, bool QUANT, bool AC, bool BND, typename R, int SINK>
static int launch_one(const LaunchParams& p, const ProbDev* probs, size_t smem, unsigned grid, unsigned block,
                      hipStream_t st) {
  auto kern = k_trace<DIM, LDSH, QUANT, AC, BND, R, SINK>;
  if (smem > 64 * 1024)
    RTMC_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), smem, st, p, probs);
  RTMC_HIP(hipGetLastError());
  return RTMC_OK;
}

template <int DIM, bool LDSH, typename R, int SINK>
static int launch_proc(const LaunchParams& p, const ProbDev* probs, bool quant, bool ac, bool bnd, size_t smem,
                       unsigned grid, unsigned block, hipStream_t st) {
  if (SINK == 1) {
    if (quant) {
      if (bnd) return launch_one<DIM, LDSH, true, true, true, R, 1>(p, probs, smem, grid, block, st);
      return launch_one<DIM, LDSH, true, true, false, R, 1>(p, probs, smem, grid, block, st);
    }
    return launch_one<DIM, LDSH, false, true, false, R, 1>(p, probs, smem, grid, block, st);
  } else {
    if (quant) {
      if (ac) {
        if (bnd) return launch_one<DIM, LDSH, true, true, true, R, 0>(p, probs, smem, grid, block, st);
        return launch_one<DIM, LDSH, true, true, false, R, 0>(p, probs, smem, grid, block, st);
      }
      if (bnd) return launch_one<DIM, LDSH, true, false, true, R, 0>(p, probs, smem, grid, block, st);
      return launch_one<DIM, LDSH, true, false, false, R, 0>(p, probs, smem, grid, block, st);
    }
    if (ac) return launch_one<DIM, LDSH, false, true, false, R, 0>(p, probs, smem, grid, block, st);
    return launch_one<DIM, LDSH, false, false, false, R, 0>(p, probs, smem, grid, block, st);
  }
}

template <int DIM, typename R>
int launch_trace(const LaunchParams& p, const ProbDev* probs, bool ldsh, bool quant, bool ac, bool bnd, int sink,
                 size_t smem, unsigned grid, unsigned block, hipStream_t st) {
  if (sink == 1) {
    RTMC_REQUIRE(ac, "the adjoint supports acoustic media only");
    if (ldsh) return launch_proc<DIM, true, R, 1>(p, probs, quant, ac, bnd, smem, grid, block, st);
    return launch_proc<DIM, false, R, 1>(p, probs, quant, ac, bnd, smem, grid, block, st);
  }
  if (ldsh) return launch_proc<DIM, true, R, 0>(p, probs, quant, ac, bnd, smem, grid, block, st);
  return launch_proc<DIM, false, R, 0>(p, probs, quant, ac, bnd, smem, grid, block, st);
}

}  // namespace rtmc
