// Kernels of the truncation-closure network (gfx950): bfloat16 matrix-core
// GEMMs with float32 accumulation, GELU, the squared-error loss, Adam, and
// the fused inference pass that feeds the HEOM stage kernels.
//
// Storage: float32 everywhere in HBM (weights (out, in) row-major,
// activations (batch, features) row-major); GEMM operands are rounded to
// bfloat16 (nearest even) when staged into LDS, products accumulate in
// float32 on v_mfma_f32_16x16x32_bf16.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_act.h"
#include "mlp_kernels.h"

namespace phi {
namespace mlp {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

using act::gelu;
using act::gelu_grad;
using act::to_bf16;

// --- Strided GEMM -----------------------------------------------------------------
// C(i, j) = sum_k A(i, k) B(k, j), A(i, k) = A[i*sa_i + k*sa_k],
// B(k, j) = B[k*sb_k + j*sb_j].  Block tile 64 x 64, K step 32, four
// wavefronts of 32 x 32 (2 x 2 MFMA tiles of 16 x 16).  Lane l of
// v_mfma_f32_16x16x32_bf16 holds A[l%16][8(l/16) .. 8(l/16)+7] and
// B[8(l/16) .. +7][l%16]; output register r holds D[4(l/16) + r][l%16].
constexpr int kBM = 64, kBN = 64, kBK = 32, kPad = 8;

__global__ void __launch_bounds__(256) gemm_bf16(GemmArgs g) {
  __shared__ __align__(16) unsigned short As[kBM][kBK + kPad];
  __shared__ __align__(16) unsigned short Bs[kBN][kBK + kPad];  // Bs[j][k]
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int q = lane >> 4, c = lane & 15;
  const int i0 = blockIdx.y * kBM, j0 = blockIdx.x * kBN;
  floatx4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = (floatx4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += kBK) {
    // Stage the tiles (zero outside the matrices).
#pragma unroll
    for (int t = 0; t < (kBM * kBK) / 256; ++t) {
      const int e = tid + t * 256;
      const int r = e / kBK, kk = e % kBK;
      const int i = i0 + r, k = k0 + kk;
      float v = 0.f;
      if (i < g.M && k < g.K) v = g.A[(long long)i * g.sa_i + (long long)k * g.sa_k];
      As[r][kk] = to_bf16(v);
    }
#pragma unroll
    for (int t = 0; t < (kBN * kBK) / 256; ++t) {
      const int e = tid + t * 256;
      const int r = e / kBK, kk = e % kBK;
      const int j = j0 + r, k = k0 + kk;
      float v = 0.f;
      if (j < g.N && k < g.K) v = g.B[(long long)k * g.sb_k + (long long)j * g.sb_j];
      Bs[r][kk] = to_bf16(v);
    }
    __syncthreads();
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const unsigned short* pa = &As[wm * 32 + a * 16 + c][8 * q];
      const unsigned short* pb = &Bs[wn * 32 + a * 16 + c][8 * q];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        af[a][t] = __builtin_bit_cast(__bf16, pa[t]);
        bfr[a][t] = __builtin_bit_cast(__bf16, pb[t]);
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    __syncthreads();
  }
  // Epilogue.
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wm * 32 + a * 16 + 4 * q + r;
        const int j = j0 + wn * 32 + b * 16 + c;
        if (i < g.M && j < g.N) {
          float v = acc[a][b][r];
          if (g.bias) v += g.bias[j];
          const long long o = (long long)i * g.sc_i + j;
          if (g.gelu) {
            g.Z[o] = v;
            g.C[o] = gelu(v);
          } else if (g.accumulate) {
            g.C[o] += v;
          } else {
            g.C[o] = v;
          }
        }
      }
}

// dZ = dH * GELU'(Z), elementwise.
__global__ void gelu_backward(const float* __restrict__ dH, const float* __restrict__ Z,
                              float* __restrict__ dZ, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dZ[i] = dH[i] * gelu_grad(Z[i]);
}

// out[j] = sum_i X[i*ld + j] over rows, in row order (deterministic).
__global__ void column_sums(const float* __restrict__ X, int rows, int cols, long long ld,
                            float* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cols) return;
  float s = 0.f;
  for (int i = 0; i < rows; ++i) s += X[(long long)i * ld + j];
  out[j] = s;
}

// dY = 2 (Y - T) / n; partial[block] = sum of (Y - T)^2 over the block's
// grid-stride range (summed by the host in block order).
__global__ void __launch_bounds__(256) mse_grad(const float* __restrict__ Y,
                                                const float* __restrict__ T,
                                                float* __restrict__ dY, long long n,
                                                double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0;
  const float scale = 2.0f / (float)n;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float d = Y[i] - T[i];
    dY[i] = scale * d;
    s += (double)d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam on flat float32 arrays (bias-corrected step sizes from the host).
__global__ void adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                     float* __restrict__ v, long long n, float lr_t, float beta1, float beta2,
                     float eps_t) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr_t * mi / (sqrtf(vi) + eps_t);
  }
}

// rows[i] = src[index[i]] for rows of `cols` floats (minibatch gather).
__global__ void gather_rows(const float* __restrict__ src, const int* __restrict__ index,
                            int rows, int cols, float* __restrict__ dst) {
  const long long n = (long long)rows * cols;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e / cols), c = (int)(e - (long long)r * cols);
    dst[e] = src[(long long)index[r] * cols + c];
  }
}

}  // namespace mlp

// --- Host launchers ---------------------------------------------------------------

hipError_t LaunchGemmBf16(const mlp::GemmArgs& g, hipStream_t s) {
  const dim3 grid((g.N + mlp::kBN - 1) / mlp::kBN, (g.M + mlp::kBM - 1) / mlp::kBM);
  hipLaunchKernelGGL(mlp::gemm_bf16, grid, dim3(256), 0, s, g);
  return hipGetLastError();
}

hipError_t LaunchGeluBackward(const float* dH, const float* Z, float* dZ, long long n,
                              hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(mlp::gelu_backward, dim3((unsigned)blocks), dim3(256), 0, s, dH, Z, dZ, n);
  return hipGetLastError();
}

hipError_t LaunchColumnSums(const float* X, int rows, int cols, long long ld, float* out,
                            hipStream_t s) {
  hipLaunchKernelGGL(mlp::column_sums, dim3((cols + 255) / 256), dim3(256), 0, s, X, rows, cols,
                     ld, out);
  return hipGetLastError();
}

hipError_t LaunchMseGrad(const float* Y, const float* T, float* dY, long long n, double* partial,
                         int blocks, hipStream_t s) {
  hipLaunchKernelGGL(mlp::mse_grad, dim3(blocks), dim3(256), 0, s, Y, T, dY, n, partial);
  return hipGetLastError();
}

hipError_t LaunchAdam(float* p, const float* g, float* m, float* v, long long n, float lr_t,
                      float beta1, float beta2, float eps_t, hipStream_t s) {
  long long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(mlp::adam, dim3((unsigned)blocks), dim3(256), 0, s, p, g, m, v, n, lr_t,
                     beta1, beta2, eps_t);
  return hipGetLastError();
}

hipError_t LaunchGatherRows(const float* src, const int* index, int rows, int cols, float* dst,
                            hipStream_t s) {
  long long n = (long long)rows * cols;
  long long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(mlp::gather_rows, dim3((unsigned)blocks), dim3(256), 0, s, src, index, rows,
                     cols, dst);
  return hipGetLastError();
}

}  // namespace phi
