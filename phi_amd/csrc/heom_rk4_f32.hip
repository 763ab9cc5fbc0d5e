// RK4 stage kernels, FP32 bulk precision (stage 0 reads the FP64 state).
#include "heom_stage_launch.h"

namespace phi {

hipError_t LaunchRk4F32(const RhsArgs& a, int stage, int form, const void* in, double* rho,
                        void* aux, void* yout, double dt, hipStream_t s) {
  if (stage == 0)
    return LaunchRk4T<float, double>(a, 0, form, (const double*)in, rho, aux, (float*)yout,
                                     dt, s);
  return LaunchRk4T<float, float>(a, stage, form, (const float*)in, rho, aux, (float*)yout,
                                  dt, s);
}

}  // namespace phi
