// rtmc_common.hpp -- host-side helpers shared by the librtmc translation
// units: error reporting across the C ABI, argument checks, and conversion
// of the ABI descriptors into the structs the kernels read.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/rtmc.h"
#include "rtmc_device.hpp"

namespace rtmc {

void set_error(const char* fmt, ...);
void clear_error();
const char* last_error();

// Returns RTMC_ERR_HIP (and records the message) when a HIP call fails.
#define RTMC_HIP(call)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ::rtmc::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                      \
      return RTMC_ERR_HIP;                                                              \
    }                                                                                   \
  } while (0)

#define RTMC_REQUIRE(cond, ...)        \
  do {                                 \
    if (!(cond)) {                     \
      ::rtmc::set_error(__VA_ARGS__);  \
      return RTMC_ERR_ARG;             \
    }                                  \
  } while (0)

// Per-device scratch of the library: a few words for reductions and error
// flags of the function-level kernels.  Allocated once per device.
struct Scratch {
  double* sum;   // 1 double
  int* flag;     // 4 ints
};
int get_scratch(Scratch* s);

int check_geometry(const rtmc_geometry* g);
int check_table(const rtmc_table* t, const char* what);
int check_material(const rtmc_material* m, const rtmc_geometry* g);
MatDev to_matdev(const rtmc_material* m);
SlotDraws to_slotdraws(const rtmc_draws* d, int64_t n);
int check_draws(const rtmc_draws* d, int n_slots, const char* what);

}  // namespace rtmc
