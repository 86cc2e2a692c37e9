// Acquisition boundaries of the batched form: one thread stores the device's
// constant-rate wall clock, so that the times between acquisitions are read
// the same way after an eager call and after a graph replay (a captured
// kernel writes fresh stamps on every replay).
#include "internal.hpp"

namespace mtd {

__global__ void k_stamp(unsigned long long* __restrict__ out) {
  *out = (unsigned long long)wall_clock64();
}

hipError_t launch_stamp(unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, out);
  return hipGetLastError();
}

}  // namespace mtd
