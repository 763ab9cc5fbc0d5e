// Kernels of the explicit-k RK4 / RKF45 path (adaptive hierarchy truncation,
// PHI's FilterTolerance, hierarchy_updater.cc:256-304,
// hierarchy_integrator.cc:2358-2385, 2828-2851; diagonal/full coupling
// operators; time-local truncation).
//
// With the filter on, every node carries an activity flag, fixed during a
// step.  The RHS of a stage is PHI's base Get with the adaptive DoNext /
// DoPrevious (the general RHS kernel with the flags, heom_kernels.h):
// every node keeps its own terms and the terms of inactive neighbours are
// skipped.  After each accepted step a node is switched off when
// max |rho_n(i,j)| < tol and on when it is > tol.
#ifndef PHI_AMD_FILTER_KERNELS_H_
#define PHI_AMD_FILTER_KERNELS_H_

#include <hip/hip_runtime.h>

#include <cstddef>

namespace phi {

// out = rho + c[0] k[0] + ... + c[nk-1] k[nk-1] (left to right, nk <= 6).
hipError_t LaunchFilterCombine(const double* rho, const double* const* k, const double* c, int nk,
                               double* out, size_t n, hipStream_t s);
// RK4 update in PHI's order: rho += c * (k0 + 2 k1 + 2 k2 + k3).
hipError_t LaunchFilterRk4Final(double* rho, const double* const* k, double c, size_t n,
                                hipStream_t s);
// RKF45 end of an attempt: y4 = rho + c4 . (k0, k2, k3, k4),
// y5 = rho + c5 . (k0, k2, k3, k4, k5); out = y5; *err = max |y4 - y5| (as the
// bits of a non-negative double, atomically maxed; the caller zeroes it).
hipError_t LaunchFilterRkf45Final(const double* rho, const double* const* k, const double* c4,
                                  const double* c5, double* out, unsigned long long* err,
                                  size_t n, hipStream_t s);
// Flag update after a step: off when max |rho_n| < tol, on when > tol.
hipError_t LaunchUpdateActive(const double* rho, unsigned char* active, int ne, int nodes,
                              double tol, hipStream_t s);

}  // namespace phi

#endif  // PHI_AMD_FILTER_KERNELS_H_
