// Launch templates of the stage kernels, shared by the per-precision
// translation units.
#ifndef PHI_AMD_HEOM_STAGE_LAUNCH_H_
#define PHI_AMD_HEOM_STAGE_LAUNCH_H_

#include "heom_device.h"
#include "heom_stream.h"

namespace phi {

inline int Blocks(const RhsArgs& a) { return (a.num_nodes + a.tile - 1) / a.tile; }

// Blocks of a persistent kernel that are resident at once on the device
// (occupancy per CU times the CU count), cached per kernel and shape.
int ResidentBlocks(const void* kernel, int threads, size_t lds);

// One launch of the persistent stage kernel: `elems` elements per thread
// (1..4) and `mps` modes per site (1..2) select the instantiation.
template <typename TB, typename TI, int EPT, int MPS, bool ENS, class EPI>
inline hipError_t LaunchStreamKE(const RhsArgs& a, const TI* in, const EPI& epi, int threads,
                                 int grid, hipStream_t s) {
  const size_t sh = dev::StreamLdsBytes(a.N, a.D, a.L, a.tile, sizeof(TB));
  const void* k = reinterpret_cast<const void*>(&dev::stream_stage<TB, TI, EPT, MPS, ENS, EPI>);
  const int ntiles = ENS ? a.ntiles : (a.num_nodes + a.tile - 1) / a.tile;
  // An ensemble block prefetches a new member's H fragments two per thread.
  if (ENS && dev::StreamFragCount(a.N) > 2 * threads) return hipErrorInvalidConfiguration;
  int g = grid > 0 ? grid : ResidentBlocks(k, threads, sh);
  if (g > ntiles) g = ntiles;
  if (g < 1) g = 1;
  hipLaunchKernelGGL((dev::stream_stage<TB, TI, EPT, MPS, ENS, EPI>), dim3(g), dim3(threads), sh,
                     s, a, in, epi);
  return hipGetLastError();
}

// Ensembles (a descriptor list) and single hierarchies are separate
// instantiations, so the single-hierarchy loop carries no descriptor code.
template <typename TB, typename TI, int EPT, int MPS, class EPI>
inline hipError_t LaunchStreamK(const RhsArgs& a, const TI* in, const EPI& epi, int threads,
                                int grid, hipStream_t s) {
  if (a.tiles) return LaunchStreamKE<TB, TI, EPT, MPS, true, EPI>(a, in, epi, threads, grid, s);
  return LaunchStreamKE<TB, TI, EPT, MPS, false, EPI>(a, in, epi, threads, grid, s);
}

template <typename TB, typename TI, int MPS, class EPI>
inline hipError_t LaunchStreamE(const RhsArgs& a, const TI* in, const EPI& epi, int threads,
                                int grid, hipStream_t s) {
  const int ept = (a.tile * a.NE + threads - 1) / threads;
  switch (ept) {
    case 1: return LaunchStreamK<TB, TI, 1, MPS, EPI>(a, in, epi, threads, grid, s);
    case 2: return LaunchStreamK<TB, TI, 2, MPS, EPI>(a, in, epi, threads, grid, s);
    case 3: return LaunchStreamK<TB, TI, 3, MPS, EPI>(a, in, epi, threads, grid, s);
    default: return LaunchStreamK<TB, TI, 4, MPS, EPI>(a, in, epi, threads, grid, s);
  }
}

template <typename TB, typename TI, class EPI>
inline hipError_t LaunchStream(const RhsArgs& a0, const TI* in, const EPI& epi, hipStream_t s) {
  const bool fp32 = sizeof(TB) == 4;
  RhsArgs a = a0;
  a.tile = a0.col_tile[fp32 ? 1 : 0];
  const int threads = a0.col_threads[fp32 ? 1 : 0];
  const int grid = a0.col_grid[fp32 ? 1 : 0];
  if (a.mps <= 1) return LaunchStreamE<TB, TI, 1, EPI>(a, in, epi, threads, grid, s);
  return LaunchStreamE<TB, TI, 2, EPI>(a, in, epi, threads, grid, s);
}

// RK4 stage through the persistent kernels (stage 0 reads the FP64 state).
// form: kRk4Classic (aux = FP64 accumulator), kRk4Nested, kRk4Increment
// (aux = K, bulk precision), kRk4ClassicBulk (aux = bulk accumulator, then
// the bulk copy of rho).
template <typename TB>
inline hipError_t LaunchRk4Stream(const RhsArgs& a, int stage, int form, const void* in,
                                  double* rho, void* aux, TB* yout, double dt, hipStream_t s) {
  const double* x = (const double*)in;
  const TB* y = (const TB*)in;
  double* acc = (double*)aux;
  if (form == kRk4Increment) {
    const TB* kk = (const TB*)aux;
    switch (stage) {
      case 0: return LaunchStream<TB, double>(a, x, dev::EpiIncr<TB, 0>{rho, kk, yout, dt}, s);
      case 1: return LaunchStream<TB, TB>(a, y, dev::EpiIncr<TB, 1>{rho, kk, yout, dt}, s);
      case 2: return LaunchStream<TB, TB>(a, y, dev::EpiIncr<TB, 2>{rho, kk, yout, dt}, s);
      default: return LaunchStream<TB, TB>(a, y, dev::EpiIncr<TB, 3>{rho, kk, yout, dt}, s);
    }
  }
  if (form == kRk4ClassicBulk) {
    // aux = accumulator (first half) and bulk copy of rho (second half)
    TB* ab = (TB*)aux;
    TB* rb = ab + 2 * (size_t)a.num_nodes * a.NE;
    switch (stage) {
      case 0: return LaunchStream<TB, double>(a, x, dev::EpiClassicBulk<TB, 0, true>{rho, ab, rb, yout, dt}, s);
      case 1: return LaunchStream<TB, TB>(a, y, dev::EpiClassicBulk<TB, 1, false>{rho, ab, rb, yout, dt}, s);
      case 2: return LaunchStream<TB, TB>(a, y, dev::EpiClassicBulk<TB, 2, false>{rho, ab, rb, yout, dt}, s);
      default: return LaunchStream<TB, TB>(a, y, dev::EpiClassicBulk<TB, 3, false>{rho, ab, rb, yout, dt}, s);
    }
  }
  if (form == kRk4Nested) {
    switch (stage) {
      case 0: return LaunchStream<TB, double>(a, x, dev::EpiNested<TB, 0, true>{rho, yout, dt / 4}, s);
      case 1: return LaunchStream<TB, TB>(a, y, dev::EpiNested<TB, 1, false>{rho, yout, dt / 3}, s);
      case 2: return LaunchStream<TB, TB>(a, y, dev::EpiNested<TB, 2, false>{rho, yout, dt / 2}, s);
      default: return LaunchStream<TB, TB>(a, y, dev::EpiNested<TB, 3, false>{rho, yout, dt}, s);
    }
  }
  switch (stage) {
    case 0: return LaunchStream<TB, double>(a, x, dev::EpiClassic<TB, 0, true>{rho, acc, yout, dt}, s);
    case 1: return LaunchStream<TB, TB>(a, y, dev::EpiClassic<TB, 1, false>{rho, acc, yout, dt}, s);
    case 2: return LaunchStream<TB, TB>(a, y, dev::EpiClassic<TB, 2, false>{rho, acc, yout, dt}, s);
    default: return LaunchStream<TB, TB>(a, y, dev::EpiClassic<TB, 3, false>{rho, acc, yout, dt}, s);
  }
}

// The launch shape of a precision: nodes and threads per block chosen on the
// host for that precision's LDS budget.
inline RhsArgs Shape(const RhsArgs& a, bool fp32, int* threads) {
  RhsArgs b = a;
  if (a.col_nm == 2) {
    b.tile = a.v1_tile[fp32 ? 1 : 0];
    *threads = 256;
  } else {
    b.tile = a.col_tile[fp32 ? 1 : 0];
    *threads = a.col_threads[fp32 ? 1 : 0];
  }
  return b;
}

template <typename TB, typename TI, int NM, bool HORNER>
inline void LaunchRk4NM(const RhsArgs& a, int stage, const TI* in, double* rho, double* acc,
                        TB* yout, double dt, int threads, hipStream_t s) {
  const size_t sh = RhsSharedBytes(a, sizeof(TB) == 4);
  const dim3 grid(Blocks(a)), block(threads);
  switch (stage) {
    case 0: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, NM, 0, HORNER>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    case 1: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, NM, 1, HORNER>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    case 2: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, NM, 2, HORNER>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    default: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, NM, 3, HORNER>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
  }
}

// `form` selects the stage order (the one-shot kernels have the classic and
// the nested one, see rk4_stage).
template <typename TB, typename TI>
inline hipError_t LaunchRk4T(const RhsArgs& a0, int stage, int form, const TI* in, double* rho,
                             void* aux, TB* yout, double dt, hipStream_t s) {
  if (a0.col_nm == 2) return LaunchRk4Stream<TB>(a0, stage, form, in, rho, aux, yout, dt, s);
  if (form == kRk4Increment || form == kRk4ClassicBulk) return hipErrorInvalidValue;
  const bool horner = form == kRk4Nested;
  double* acc = (double*)aux;
  int threads;
  const RhsArgs a = Shape(a0, sizeof(TB) == 4, &threads);
  if (a.col_nm >= 1) {
    if (horner) LaunchRk4NM<TB, TI, 1, true>(a, stage, in, rho, acc, yout, dt, threads, s);
    else LaunchRk4NM<TB, TI, 1, false>(a, stage, in, rho, acc, yout, dt, threads, s);
  } else {
    if (horner) LaunchRk4NM<TB, TI, 0, true>(a, stage, in, rho, acc, yout, dt, threads, s);
    else LaunchRk4NM<TB, TI, 0, false>(a, stage, in, rho, acc, yout, dt, threads, s);
  }
  return hipGetLastError();
}

// RKF45 stage through the persistent kernels (stage 0 reads the FP64 state).
template <typename TB>
inline hipError_t LaunchRkf45Stream(const RhsArgs& a, int stage, const void* in, const double* rho,
                                    const Rkf45Bufs<TB>& b, const Rkf45Coefs& c, hipStream_t s) {
  const double* x = (const double*)in;
  const TB* y = (const TB*)in;
  switch (stage) {
    case 0: return LaunchStream<TB, double>(a, x, dev::EpiRkf45<TB, 0, true>{b, rho, c}, s);
    case 1: return LaunchStream<TB, TB>(a, y, dev::EpiRkf45<TB, 1, false>{b, rho, c}, s);
    case 2: return LaunchStream<TB, TB>(a, y, dev::EpiRkf45<TB, 2, false>{b, rho, c}, s);
    case 3: return LaunchStream<TB, TB>(a, y, dev::EpiRkf45<TB, 3, false>{b, rho, c}, s);
    case 4: return LaunchStream<TB, TB>(a, y, dev::EpiRkf45<TB, 4, false>{b, rho, c}, s);
    default: return LaunchStream<TB, TB>(a, y, dev::EpiRkf45<TB, 5, false>{b, rho, c}, s);
  }
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
  if (a0.col_nm == 2) return LaunchRkf45Stream<TB>(a0, stage, in, rho, b, c, s);
  int threads;
  const RhsArgs a = Shape(a0, sizeof(TB) == 4, &threads);
  if (a.col_nm >= 1) LaunchRkf45NM<TB, TI, 1>(a, stage, in, rho, b, c, threads, s);
  else LaunchRkf45NM<TB, TI, 0>(a, stage, in, rho, b, c, threads, s);
  return hipGetLastError();
}

}  // namespace phi

#endif  // PHI_AMD_HEOM_STAGE_LAUNCH_H_
