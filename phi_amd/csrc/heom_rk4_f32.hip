// RK4 stage kernels, FP32 bulk precision (stage 0 reads the FP64 state).
#include "heom_stage_launch.h"

namespace phi {

hipError_t LaunchRk4F32(const RhsArgs& a, int stage, bool horner, const void* in, double* rho,
                        double* acc, void* yout, double dt, hipStream_t s) {
  if (stage == 0)
    return LaunchRk4T<float, double>(a, 0, horner, (const double*)in, rho, acc, (float*)yout,
                                     dt, s);
  return LaunchRk4T<float, float>(a, stage, horner, (const float*)in, rho, acc, (float*)yout,
                                  dt, s);
}

}  // namespace phi
