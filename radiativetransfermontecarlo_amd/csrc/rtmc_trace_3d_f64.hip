// rtmc_trace_3d_f64.hip -- k_trace instances for 3-D problems with double flight arithmetic.
#include "rtmc_trace_launch.hpp"

namespace rtmc {
template int launch_trace<3, double>(const LaunchParams&, const ProbDev*, int, bool, bool, bool, int, size_t,
                                      unsigned, unsigned, hipStream_t);
}  // namespace rtmc
