// RKF45 stage kernels, FP32 bulk precision.
#include "heom_stage_launch.h"

namespace phi {

hipError_t LaunchRkf45F32(const RhsArgs& a, int stage, const void* in, const double* rho,
                         const Rkf45Bufs<float>& b, const Rkf45Coefs& c, hipStream_t s) {
  if (stage == 0) return LaunchRkf45T<float, double>(a, 0, (const double*)in, rho, b, c, s);
  return LaunchRkf45T<float, float>(a, stage, (const float*)in, rho, b, c, s);
}

}  // namespace phi
