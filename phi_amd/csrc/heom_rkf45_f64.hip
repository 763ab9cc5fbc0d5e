// RKF45 stage kernels, FP64 bulk precision.
#include "heom_stage_launch.h"

namespace phi {

hipError_t LaunchRkf45F64(const RhsArgs& a, int stage, const void* in, const double* rho,
                         const Rkf45Bufs<double>& b, const Rkf45Coefs& c, hipStream_t s) {
  if (stage == 0) return LaunchRkf45T<double, double>(a, 0, (const double*)in, rho, b, c, s);
  return LaunchRkf45T<double, double>(a, stage, (const double*)in, rho, b, c, s);
}

}  // namespace phi
