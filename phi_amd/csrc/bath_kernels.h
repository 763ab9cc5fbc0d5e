// RHS of PHI's DiagonalBath and FullBath coupling models
// (hierarchy_updater.cc:10-127 with the operators of
// Hierarchy::FillDiagonalBathOperators / FillFullBathOperators,
// hierarchy.cc:157-225), for every node, FP64, any complex input:
//   y_n = -(i/hbar)(H x_n - x_n H^dagger) - mu_n x_n + coupling terms,
// diagonal (operator b_m = diag entries of bath m, mode j of bath m = j/K):
//   - sum_m alpha_m (b_m[r] - b_m[c])^2 x(r,c)
//   + sum_{j: n+e_j} -i (b_m[r] - b_m[c]) x_{n+e_j}(r,c)
//   + sum_{j: n-e_j} n_j (-i)(b_m[r] c_j - b_m[c] conj(c_j)) x_{n-e_j}(r,c)
// (PHI's DiagonalBath uses neither ADO normalisation factor);
// full (operator F_m, N x N column-major):
//   - sum_m alpha_m [F_m, [F_m, x]]
//   + sum_{j: n+e_j} -i s_j (F_m x' - x' F_m)      s_j = next_s(j, n_j)
//   + sum_{j: n-e_j} -(row_j F_m x' + col_j x' F_m) row/col = prev_row/col(j, n_j).
#ifndef PHI_AMD_BATH_KERNELS_H_
#define PHI_AMD_BATH_KERNELS_H_

#include <hip/hip_runtime.h>

#include "heom_kernels.h"

namespace phi {

struct BathArgs {
  int model;              // 1 diagonal, 2 full
  int M, K;               // baths, terms per bath (mode j belongs to bath j / K)
  const double2* ops;     // M * N (diagonal) or M * N * N (full)
  const double* alpha;    // M, PHI's matsubara_prefactor per bath
  const double2* corr;    // D correlation-function constants c_j
};

hipError_t LaunchRhsBath(const RhsArgs& a, const BathArgs& b, const double* x, double* y,
                         hipStream_t s);

// Time-local truncation term of PHI's DoTimelocal (hierarchy_updater.cc:
// 28-43, 94-107, 170-213), added onto y for the nodes [first, first+count)
// (the deepest tier):
//   X_m = Q_m x_n - x_n Q_m^dagger,   y_n -= sum_m [V_m, X_m],
// V_m the projector on site[m] (model 0), diag(ops_m) (1) or ops_m (2).
struct TimeLocalArgs {
  int model;              // 0 site projectors, 1 diagonal, 2 full
  int M;
  int first, count;       // node range
  const double2* Q;       // M * N * N, column-major
  const double2* ops;     // M * N (diagonal) or M * N * N (full)
  const int* site;        // M (site projectors)
};

hipError_t LaunchTimeLocal(int N, const TimeLocalArgs& t, const double* x, double* y,
                           hipStream_t s);

}  // namespace phi

#endif  // PHI_AMD_BATH_KERNELS_H_
