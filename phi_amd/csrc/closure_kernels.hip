// Data movement of the learned truncation closure: network inputs from the
// ADOs of the truncation tier, training targets from a deeper hierarchy,
// and the predicted next-tier ADOs written into the "virtual" neighbour
// slots that the stage kernels gather from.
//
// Row r = m * nb + b is boundary node b (local id first + b) of member m.
// Input row:  [Re rho_n / s (N^2, column-major), Im rho_n / s (N^2), n / L (D)]
// Target/output row: for j = 0 .. D-1: [Re rho_{n+e_j} / s (N^2), Im (N^2)]
// with s = max(max |rho_n|, 1e-30) per row.
#include <hip/hip_runtime.h>

#include "closure_kernels.h"

namespace phi {
namespace closure {

template <typename TI>
struct C2T;
template <>
struct C2T<float> { typedef float2 type; };
template <>
struct C2T<double> { typedef double2 type; };

// One wavefront per row.
template <typename TI>
__global__ void __launch_bounds__(256) inputs(const TI* __restrict__ in,
                                              const int* __restrict__ moff, int members,
                                              int first, int nb, int ne, int D,
                                              const float* __restrict__ bmi,
                                              float* __restrict__ X, float* __restrict__ S) {
  typedef typename C2T<TI>::type V;
  const int lane = threadIdx.x & 63;
  const long long row = (long long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (row >= (long long)members * nb) return;
  const int m = (int)(row / nb), b = (int)(row - (long long)m * nb);
  const V* src = reinterpret_cast<const V*>(in) + ((size_t)moff[m] + first + b) * ne;
  float mx = 0.f;
  for (int e = lane; e < ne; e += 64) {
    const V v = src[e];
    mx = fmaxf(mx, fmaxf(fabsf((float)v.x), fabsf((float)v.y)));
  }
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  const float s = fmaxf(mx, 1e-30f);
  const float inv = 1.0f / s;
  const int n_in = 2 * ne + D;
  float* x = X + row * n_in;
  for (int e = lane; e < ne; e += 64) {
    const V v = src[e];
    x[e] = (float)v.x * inv;
    x[ne + e] = (float)v.y * inv;
  }
  for (int j = lane; j < D; j += 64) x[2 * ne + j] = bmi[(size_t)b * D + j];
  if (lane == 0) S[row] = s;
}

// Targets from the FP64 state of a deeper hierarchy: children[b*D + j] =
// local id of n + e_j.
__global__ void __launch_bounds__(256) targets(const double* __restrict__ rho,
                                               const int* __restrict__ moff, int members,
                                               int nb, int ne, int D,
                                               const int* __restrict__ children,
                                               const float* __restrict__ S,
                                               float* __restrict__ T) {
  const long long row = blockIdx.x;
  if (row >= (long long)members * nb) return;
  const int m = (int)(row / nb), b = (int)(row - (long long)m * nb);
  const float inv = 1.0f / S[row];
  const int n_out = 2 * ne * D;
  float* t = T + row * n_out;
  for (int e = threadIdx.x; e < ne * D; e += blockDim.x) {
    const int j = e / ne, w = e - j * ne;
    const int child = children[(size_t)b * D + j];
    const double2 v =
        reinterpret_cast<const double2*>(rho)[((size_t)moff[m] + child) * ne + w];
    t[(size_t)j * 2 * ne + w] = (float)v.x * inv;
    t[(size_t)j * 2 * ne + ne + w] = (float)v.y * inv;
  }
}

// Predicted next-tier ADOs -> virtual slots: virtual node vbase + b*D + j of
// member m holds rho_{n+e_j} of boundary node b.
template <typename TO>
__global__ void __launch_bounds__(256) scatter(const float* __restrict__ Y,
                                               const float* __restrict__ S,
                                               const int* __restrict__ moff, int members,
                                               int nb, int ne, int D, int vbase,
                                               TO* __restrict__ out) {
  typedef typename C2T<TO>::type V;
  const long long row = blockIdx.x;
  if (row >= (long long)members * nb) return;
  const int m = (int)(row / nb), b = (int)(row - (long long)m * nb);
  const float s = S[row];
  const float* y = Y + row * (size_t)(2 * ne * D);
  V* dst = reinterpret_cast<V*>(out) + ((size_t)moff[m] + vbase + (size_t)b * D) * ne;
  for (int e = threadIdx.x; e < ne * D; e += blockDim.x) {
    const int j = e / ne, w = e - j * ne;
    V v;
    v.x = (TO)(y[(size_t)j * 2 * ne + w] * s);
    v.y = (TO)(y[(size_t)j * 2 * ne + ne + w] * s);
    dst[(size_t)j * ne + w] = v;
  }
}

}  // namespace closure

hipError_t LaunchClosureInputs(const void* in, bool in_fp32, const int* moff, int members,
                               int first, int nb, int ne, int D, const float* bmi, float* X,
                               float* S, hipStream_t s) {
  const long long rows = (long long)members * nb;
  const unsigned blocks = (unsigned)((rows + 3) / 4);
  if (in_fp32)
    hipLaunchKernelGGL(closure::inputs<float>, dim3(blocks), dim3(256), 0, s, (const float*)in,
                       moff, members, first, nb, ne, D, bmi, X, S);
  else
    hipLaunchKernelGGL(closure::inputs<double>, dim3(blocks), dim3(256), 0, s, (const double*)in,
                       moff, members, first, nb, ne, D, bmi, X, S);
  return hipGetLastError();
}

hipError_t LaunchClosureTargets(const double* rho, const int* moff, int members, int nb, int ne,
                                int D, const int* children, const float* S, float* T,
                                hipStream_t s) {
  const long long rows = (long long)members * nb;
  hipLaunchKernelGGL(closure::targets, dim3((unsigned)rows), dim3(256), 0, s, rho, moff, members,
                     nb, ne, D, children, S, T);
  return hipGetLastError();
}

hipError_t LaunchClosureScatter(const float* Y, const float* S, const int* moff, int members,
                                int nb, int ne, int D, int vbase, void* out, bool out_fp32,
                                hipStream_t s) {
  const long long rows = (long long)members * nb;
  if (out_fp32)
    hipLaunchKernelGGL(closure::scatter<float>, dim3((unsigned)rows), dim3(256), 0, s, Y, S,
                       moff, members, nb, ne, D, vbase, (float*)out);
  else
    hipLaunchKernelGGL(closure::scatter<double>, dim3((unsigned)rows), dim3(256), 0, s, Y, S,
                       moff, members, nb, ne, D, vbase, (double*)out);
  return hipGetLastError();
}

}  // namespace phi
