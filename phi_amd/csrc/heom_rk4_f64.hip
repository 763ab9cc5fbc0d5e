// RK4 stage kernels, FP64 bulk precision.
#include "heom_stage_launch.h"

namespace phi {

hipError_t LaunchRk4F64(const RhsArgs& a, int stage, int form, const void* in, double* rho,
                        void* aux, void* yout, double dt, hipStream_t s) {
  return LaunchRk4T<double, double>(a, stage, form, (const double*)in, rho, aux,
                                    (double*)yout, dt, s);
}

}  // namespace phi
