// rtmc_trace_launch.hpp -- compile-time specialisation of k_trace: histogram
// placement (LDS, direct atomics, time-major scratch with pairs or with runs
// of three), process, acoustic/elastic, with/without boundaries (the
// event-driven process has none), forward or adjoint (acoustic only), and
// the direction variable (dir_z or the angle to the position vector), and the
// source of the scattering angles (full tables in global memory, or compressed
// tables in LDS; dir_z instances only).
// Included by one translation unit per dimension and flight precision.
#pragma once

#include "rtmc_trace.hpp"

namespace rtmc {

template <int DIM, int HM, bool QUANT, bool AC, bool BND, typename R, int SINK, bool POS, bool LT>
static int launch_one(const LaunchParams& p, const ProbDev* probs, size_t smem, unsigned grid, unsigned block,
                      hipStream_t st) {
  auto kern = k_trace<DIM, HM, QUANT, AC, BND, R, SINK, POS, LT>;
  if (smem > 64 * 1024)
    RTMC_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(block), smem, st, p, probs);
  RTMC_HIP(hipGetLastError());
  return RTMC_OK;
}

template <int DIM, int HM, typename R, int SINK, bool POS, bool LT>
static int launch_proc(const LaunchParams& p, const ProbDev* probs, bool quant, bool ac, bool bnd, size_t smem,
                       unsigned grid, unsigned block, hipStream_t st) {
  if constexpr (SINK == 1) {  // the adjoint instances exist for the placements it uses only
    if (quant) {
      if (bnd) return launch_one<DIM, HM, true, true, true, R, 1, POS, LT>(p, probs, smem, grid, block, st);
      return launch_one<DIM, HM, true, true, false, R, 1, POS, LT>(p, probs, smem, grid, block, st);
    }
    return launch_one<DIM, HM, false, true, false, R, 1, POS, LT>(p, probs, smem, grid, block, st);
  } else {
    if (quant) {
      if (ac) {
        if (bnd) return launch_one<DIM, HM, true, true, true, R, 0, POS, LT>(p, probs, smem, grid, block, st);
        return launch_one<DIM, HM, true, true, false, R, 0, POS, LT>(p, probs, smem, grid, block, st);
      }
      if (bnd) return launch_one<DIM, HM, true, false, true, R, 0, POS, LT>(p, probs, smem, grid, block, st);
      return launch_one<DIM, HM, true, false, false, R, 0, POS, LT>(p, probs, smem, grid, block, st);
    }
    if (ac) return launch_one<DIM, HM, false, true, false, R, 0, POS, LT>(p, probs, smem, grid, block, st);
    return launch_one<DIM, HM, false, false, false, R, 0, POS, LT>(p, probs, smem, grid, block, st);
  }
}

template <int DIM, typename R, bool POS, bool LT>
static int launch_dirvar(const LaunchParams& p, const ProbDev* probs, int hm, bool quant, bool ac, bool bnd,
                         int sink, size_t smem, unsigned grid, unsigned block, hipStream_t st) {
  if (sink == 1) {
    RTMC_REQUIRE(ac, "the adjoint supports acoustic media only");
    if (hm == HM_LDS) return launch_proc<DIM, HM_LDS, R, 1, POS, LT>(p, probs, quant, ac, bnd, smem, grid, block, st);
    return launch_proc<DIM, HM_GLOBAL, R, 1, POS, LT>(p, probs, quant, ac, bnd, smem, grid, block, st);
  }
  if (hm == HM_LDS) return launch_proc<DIM, HM_LDS, R, 0, POS, LT>(p, probs, quant, ac, bnd, smem, grid, block, st);
  if (hm == HM_TMAJOR)
    return launch_proc<DIM, HM_TMAJOR, R, 0, POS, LT>(p, probs, quant, ac, bnd, smem, grid, block, st);
  if constexpr (!POS) {  // runs of three are built for dir_z only
    if (hm == HM_TMAJOR3)
      return launch_proc<DIM, HM_TMAJOR3, R, 0, POS, LT>(p, probs, quant, ac, bnd, smem, grid, block, st);
  }
  RTMC_REQUIRE(hm != HM_TMAJOR3, "runs of three are built for the dir_z direction variable only");
  return launch_proc<DIM, HM_GLOBAL, R, 0, POS, LT>(p, probs, quant, ac, bnd, smem, grid, block, st);
}

// The direction variable measured against the position vector is a separate
// instance (POS): it is binned at every deposit instead of once per event.
template <int DIM, typename R>
int launch_trace(const LaunchParams& p, const ProbDev* probs, int hm, bool quant, bool ac, bool bnd, int sink,
                 size_t smem, unsigned grid, unsigned block, hipStream_t st) {
  if (p.dirvar == RTMC_DIRVAR_POSITION) {
    RTMC_REQUIRE(p.lt_cap == 0, "compressed angle tables are built for the dir_z direction variable only");
    return launch_dirvar<DIM, R, true, false>(p, probs, hm, quant, ac, bnd, sink, smem, grid, block, st);
  }
  if (p.lt_cap > 0) return launch_dirvar<DIM, R, false, true>(p, probs, hm, quant, ac, bnd, sink, smem, grid, block, st);
  return launch_dirvar<DIM, R, false, false>(p, probs, hm, quant, ac, bnd, sink, smem, grid, block, st);
}

}  // namespace rtmc
