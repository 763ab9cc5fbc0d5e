// Launchers of the closure-network kernels (mlp_kernels.hip).
#ifndef PHI_AMD_MLP_KERNELS_H_
#define PHI_AMD_MLP_KERNELS_H_

#include <hip/hip_runtime.h>

namespace phi {
namespace mlp {

// C(i, j) = sum_k A(i, k) B(k, j) (+ bias[j]) for i < M, j < N, k < K with
// A(i, k) = A[i*sa_i + k*sa_k], B(k, j) = B[k*sb_k + j*sb_j], C(i, j) at
// C[i*sc_i + j].  gelu: Z = pre-activation, C = GELU(Z); accumulate: C +=.
struct GemmArgs {
  int M, N, K;
  const float* A;
  long long sa_i, sa_k;
  const float* B;
  long long sb_k, sb_j;
  float* C;
  long long sc_i;
  const float* bias;
  float* Z;
  int gelu;
  int accumulate;
};

}  // namespace mlp

hipError_t LaunchGemmBf16(const mlp::GemmArgs& g, hipStream_t s);
hipError_t LaunchGeluBackward(const float* dH, const float* Z, float* dZ, long long n,
                              hipStream_t s);
hipError_t LaunchColumnSums(const float* X, int rows, int cols, long long ld, float* out,
                            hipStream_t s);
hipError_t LaunchMseGrad(const float* Y, const float* T, float* dY, long long n, double* partial,
                         int blocks, hipStream_t s);
hipError_t LaunchAdam(float* p, const float* g, float* m, float* v, long long n, float lr_t,
                      float beta1, float beta2, float eps_t, hipStream_t s);
hipError_t LaunchGatherRows(const float* src, const int* index, int rows, int cols, float* dst,
                            hipStream_t s);

}  // namespace phi

#endif  // PHI_AMD_MLP_KERNELS_H_
