// RK4 stage kernels, FP64 bulk precision.
#include "heom_stage_launch.h"

namespace phi {

hipError_t LaunchRk4F64(const RhsArgs& a, int stage, bool horner, const void* in, double* rho,
                        double* acc, void* yout, double dt, hipStream_t s) {
  return LaunchRk4T<double, double>(a, stage, horner, (const double*)in, rho, acc,
                                    (double*)yout, dt, s);
}

}  // namespace phi
