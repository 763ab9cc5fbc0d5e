// Kernels of the adaptive hierarchy truncation (PHI's FilterTolerance,
// hierarchy_updater.h:217-245, hierarchy_updater.cc:256-304,
// hierarchy_integrator.cc:2358-2385, 2828-2851).
//
// With the filter on, every node carries an activity flag, fixed during a
// step.  The RHS of a stage is the plain operator applied to the stage input
// with the ADOs of inactive nodes zeroed (that drops an inactive node's own
// terms and every inactive neighbour's terms), and an inactive node's k
// keeps the previous contents of its per-stage buffer and adds the neighbour
// terms to them: PHI's adaptive Get skips DoUnitary, the call that
// overwrites drho.  After each accepted step a node is switched off when
// max |rho_n(i,j)| < tol and on when it is > tol.
#ifndef PHI_AMD_FILTER_KERNELS_H_
#define PHI_AMD_FILTER_KERNELS_H_

#include <hip/hip_runtime.h>

#include <cstddef>

namespace phi {

// y = x with the ADOs of inactive nodes zeroed (N*N = ne complex per node).
hipError_t LaunchMaskNodes(const double* x, const unsigned char* active, double* y, int ne,
                           int nodes, hipStream_t s);
// k = kraw on active nodes, k += kraw on inactive ones.
hipError_t LaunchFilterK(const double* kraw, double* k, const unsigned char* active, int ne,
                         int nodes, hipStream_t s);
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
