// Data movement of the learned truncation closure: network inputs from the
// ADOs of the truncation tier, training targets from a deeper hierarchy,
// and the predicted next-tier ADOs written into the "virtual" neighbour
// slots that the stage kernels gather from.
//
// Row r = m * nb + b is boundary node b (local id first + b) of member m.
// Input row:  [Re rho_n / s (N^2, column-major), Im rho_n / s (N^2), n / L (D)]
// Target/output row: for j = 0 .. D-1: [Re rho_{n+e_j} / s (N^2), Im (N^2)]
// with s = the largest |Re| or |Im| entry of rho_n (at least 1e-30), per row.
#include <hip/hip_runtime.h>

#include "closure_kernels.h"

namespace phi {
namespace closure {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ unsigned short to_bf16(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

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

// --- Fused closure pass ------------------------------------------------------------
// One block of 256 threads takes kRows rows end to end: network inputs from
// the stage input (as `inputs`), the three layers on the matrix cores
// (v_mfma_f32_16x16x32_bf16; activations in LDS as bfloat16, weights packed
// as bfloat16 rows streamed from L2), and the scatter of the outputs into
// the virtual slots (as `scatter`).  Same arithmetic as the unfused path:
// operands rounded to bfloat16, float32 sums in the same K order, bias and
// GELU in float32.
constexpr int kRows = 32;

// One layer for the block's rows: Out = act(In W^T + b).  Each wavefront
// computes column chunks of 64 (4 MFMA tiles) for all kRows rows (2 tiles).
// W is (ncols_padded x K) bfloat16 row-major with K a multiple of 32.
template <class EPI>
__device__ __forceinline__ void fused_layer(const unsigned short* In, int stride_in,
                                            const unsigned short* __restrict__ W, int K,
                                            int ncols, const float* __restrict__ bias,
                                            const EPI& epi) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int q = lane >> 4, c = lane & 15;
  for (int chunk = wave; chunk * 64 < ncols; chunk += 4) {
    floatx4 acc[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = (floatx4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 32) {
      bf16x8 av[2], bv[4];
#pragma unroll
      for (int a = 0; a < 2; ++a)
        av[a] = *reinterpret_cast<const bf16x8*>(In + (16 * a + c) * stride_in + k0 + 8 * q);
#pragma unroll
      for (int b = 0; b < 4; ++b)
        bv[b] = *reinterpret_cast<const bf16x8*>(W + (size_t)(chunk * 64 + 16 * b + c) * K + k0 +
                                                 8 * q);
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * a + 4 * q + r;
          const int col = chunk * 64 + 16 * b + c;
          if (col < ncols) epi(row, col, acc[a][b][r] + bias[col]);
        }
  }
}

struct ToLds {
  unsigned short* out;
  int stride;
  __device__ __forceinline__ void operator()(int row, int col, float v) const {
    out[row * stride + col] = to_bf16(gelu(v));
  }
};

template <typename TO>
struct ToSlots {
  TO* slab;
  const int* rowbase;  // per block row: first virtual slot's element offset, or -1
  const float* srow;
  int ne;
  __device__ __forceinline__ void operator()(int row, int col, float v) const {
    const int base = rowbase[row];
    if (base < 0) return;
    const int j = col / (2 * ne), w = col - j * 2 * ne;
    const int part = w >= ne, e = w - part * ne;
    // Complex element e of slot j: real part at 2*(...), imaginary at +1.
    slab[2 * ((size_t)base + (size_t)j * ne + e) + part] = (TO)(v * srow[row]);
  }
};

template <typename TI>
__global__ void __launch_bounds__(256) fused(FusedNet net, TI* slab, const int* __restrict__ moff,
                                             int members, int first, int nb, int ne, int D,
                                             const float* __restrict__ bmi, int vbase) {
  typedef typename C2T<TI>::type V;
  extern __shared__ __align__(16) unsigned short lds[];
  __shared__ float srow[kRows];
  __shared__ int rowbase[kRows];
  const int stride = net.width + 8;
  unsigned short* A0 = lds;
  unsigned short* A1 = lds + kRows * stride;
  const long long R = (long long)members * nb;
  const long long row0 = (long long)blockIdx.x * kRows;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // Inputs: one wavefront per row.
  for (int r = wave; r < kRows; r += 4) {
    const long long row = row0 + r;
    unsigned short* x = A0 + r * stride;
    if (row < R) {
      const int m = (int)(row / nb), b = (int)(row - (long long)m * nb);
      const V* src = reinterpret_cast<const V*>(slab) + ((size_t)moff[m] + first + b) * ne;
      float mx = 0.f;
      for (int e = lane; e < ne; e += 64) {
        const V v = src[e];
        mx = fmaxf(mx, fmaxf(fabsf((float)v.x), fabsf((float)v.y)));
      }
      for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
      const float s = fmaxf(mx, 1e-30f);
      const float inv = 1.0f / s;
      for (int e = lane; e < ne; e += 64) {
        const V v = src[e];
        x[e] = to_bf16((float)v.x * inv);
        x[ne + e] = to_bf16((float)v.y * inv);
      }
      for (int j = lane; j < D; j += 64) x[2 * ne + j] = to_bf16(bmi[(size_t)b * D + j]);
      for (int k = 2 * ne + D + lane; k < net.k1; k += 64) x[k] = 0;
      if (lane == 0) {
        srow[r] = s;
        rowbase[r] = (int)(((size_t)moff[m] + vbase + (size_t)b * D) * ne);
      }
    } else {
      for (int k = lane; k < net.k1; k += 64) x[k] = 0;
      if (lane == 0) {
        srow[r] = 0.f;
        rowbase[r] = -1;
      }
    }
  }
  __syncthreads();
  fused_layer(A0, stride, net.w1, net.k1, net.hid, net.b1, ToLds{A1, stride});
  __syncthreads();
  fused_layer(A1, stride, net.w2, net.hid, net.hid, net.b2, ToLds{A0, stride});
  __syncthreads();
  fused_layer(A0, stride, net.w3, net.hid, net.n_out, net.b3,
              ToSlots<TI>{slab, rowbase, srow, ne});
}

}  // namespace closure

size_t FusedClosureLds(const FusedNet& net) {
  return (size_t)2 * closure::kRows * (net.width + 8) * sizeof(unsigned short);
}

hipError_t LaunchClosureFused(const FusedNet& net, void* slab, bool fp32, const int* moff,
                              int members, int first, int nb, int ne, int D, const float* bmi,
                              int vbase, hipStream_t s) {
  const long long rows = (long long)members * nb;
  const unsigned blocks = (unsigned)((rows + closure::kRows - 1) / closure::kRows);
  const size_t lds = FusedClosureLds(net);
  if (fp32)
    hipLaunchKernelGGL(closure::fused<float>, dim3(blocks), dim3(256), lds, s, net, (float*)slab,
                       moff, members, first, nb, ne, D, bmi, vbase);
  else
    hipLaunchKernelGGL(closure::fused<double>, dim3(blocks), dim3(256), lds, s, net,
                       (double*)slab, moff, members, first, nb, ne, D, bmi, vbase);
  return hipGetLastError();
}

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
