// Launch templates of the stage kernels, shared by the per-precision
// translation units.
#ifndef PHI_AMD_HEOM_STAGE_LAUNCH_H_
#define PHI_AMD_HEOM_STAGE_LAUNCH_H_

#include "heom_device.h"

namespace phi {

inline int Blocks(const RhsArgs& a) { return (a.num_nodes + a.tile - 1) / a.tile; }

// The launch shape of a precision: nodes and threads per block chosen on the
// host for that precision's LDS budget.
inline RhsArgs Shape(const RhsArgs& a, bool fp32, int* threads) {
  RhsArgs b = a;
  b.tile = a.col_tile[fp32 ? 1 : 0];
  *threads = a.col_threads[fp32 ? 1 : 0];
  return b;
}

template <typename TB, typename TI, int NM>
inline void LaunchRk4NM(const RhsArgs& a, int stage, const TI* in, double* rho, double* acc,
                        TB* yout, double dt, int threads, hipStream_t s) {
  const size_t sh = RhsSharedBytes(a, sizeof(TB) == 4);
  const dim3 grid(Blocks(a)), block(threads);
  switch (stage) {
    case 0: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, NM, 0>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    case 1: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, NM, 1>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    case 2: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, NM, 2>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    default: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, NM, 3>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
  }
}

template <typename TB, typename TI>
inline hipError_t LaunchRk4T(const RhsArgs& a0, int stage, const TI* in, double* rho, double* acc,
                             TB* yout, double dt, hipStream_t s) {
  int threads;
  const RhsArgs a = Shape(a0, sizeof(TB) == 4, &threads);
  switch (a.col_nm) {
    case 1: LaunchRk4NM<TB, TI, 1>(a, stage, in, rho, acc, yout, dt, threads, s); break;
    case 8: LaunchRk4NM<TB, TI, 8>(a, stage, in, rho, acc, yout, dt, threads, s); break;
    case 16: LaunchRk4NM<TB, TI, 16>(a, stage, in, rho, acc, yout, dt, threads, s); break;
    case 24: LaunchRk4NM<TB, TI, 24>(a, stage, in, rho, acc, yout, dt, threads, s); break;
    case 32: LaunchRk4NM<TB, TI, 32>(a, stage, in, rho, acc, yout, dt, threads, s); break;
    default: LaunchRk4NM<TB, TI, 0>(a, stage, in, rho, acc, yout, dt, threads, s); break;
  }
  return hipGetLastError();
}

template <typename TB, typename TI, int NM>
inline void LaunchRkf45NM(const RhsArgs& a, int stage, const TI* in, const double* rho,
                          const Rkf45Bufs<TB>& b, const Rkf45Coefs& c, int threads,
                          hipStream_t s) {
  const size_t sh = RhsSharedBytes(a, sizeof(TB) == 4);
  const dim3 grid(Blocks(a)), block(threads);
  switch (stage) {
    case 0: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, NM, 0>), grid, block, sh, s, a, in, rho, b, c); break;
    case 1: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, NM, 1>), grid, block, sh, s, a, in, rho, b, c); break;
    case 2: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, NM, 2>), grid, block, sh, s, a, in, rho, b, c); break;
    case 3: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, NM, 3>), grid, block, sh, s, a, in, rho, b, c); break;
    case 4: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, NM, 4>), grid, block, sh, s, a, in, rho, b, c); break;
    default: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, NM, 5>), grid, block, sh, s, a, in, rho, b, c); break;
  }
}

template <typename TB, typename TI>
inline hipError_t LaunchRkf45T(const RhsArgs& a0, int stage, const TI* in, const double* rho,
                               const Rkf45Bufs<TB>& b, const Rkf45Coefs& c, hipStream_t s) {
  int threads;
  const RhsArgs a = Shape(a0, sizeof(TB) == 4, &threads);
  switch (a.col_nm) {
    case 1: LaunchRkf45NM<TB, TI, 1>(a, stage, in, rho, b, c, threads, s); break;
    case 8: LaunchRkf45NM<TB, TI, 8>(a, stage, in, rho, b, c, threads, s); break;
    case 16: LaunchRkf45NM<TB, TI, 16>(a, stage, in, rho, b, c, threads, s); break;
    case 24: LaunchRkf45NM<TB, TI, 24>(a, stage, in, rho, b, c, threads, s); break;
    case 32: LaunchRkf45NM<TB, TI, 32>(a, stage, in, rho, b, c, threads, s); break;
    default: LaunchRkf45NM<TB, TI, 0>(a, stage, in, rho, b, c, threads, s); break;
  }
  return hipGetLastError();
}


}  // namespace phi

#endif  // PHI_AMD_HEOM_STAGE_LAUNCH_H_
