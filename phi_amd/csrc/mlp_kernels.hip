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
// The next K step's operands load into registers while the current one is
// multiplied from LDS (two LDS buffers, one barrier per step); the loads are
// unconditional (an element outside the matrices reads element 0 and is
// replaced by zero), so all of a step's loads are in flight at once.
constexpr int kBM = 64, kBN = 64, kBK = 32, kPad = 8;
constexpr int kPer = kBM * kBK / 256;  // elements per thread and operand

__global__ void __launch_bounds__(256) gemm_bf16(GemmArgs g) {
  __shared__ __align__(16) unsigned short As[2][kBM][kBK + kPad];
  __shared__ __align__(16) unsigned short Bs[2][kBN][kBK + kPad];  // Bs[j][k]
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
  // This thread's elements of a step, (tile row, k) = (tid / 32 + 8 t,
  // tid % 32) when the operand is contiguous along k, else (tid % 64,
  // tid / 64 + 4 t): consecutive lanes read consecutive addresses either way.
  const bool a_rows = g.sa_k != 1, b_rows = g.sb_k != 1;
  int ar[kPer], ak[kPer], br[kPer], bk[kPer];
#pragma unroll
  for (int t = 0; t < kPer; ++t) {
    ar[t] = a_rows ? (tid & 63) : (tid >> 5) + 8 * t;
    ak[t] = a_rows ? (tid >> 6) + 4 * t : (tid & (kBK - 1));
    br[t] = b_rows ? (tid & 63) : (tid >> 5) + 8 * t;
    bk[t] = b_rows ? (tid >> 6) + 4 * t : (tid & (kBK - 1));
  }
  float ra[kPer], rb[kPer];
  auto load = [&](int k0) {
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
      const int i = i0 + ar[t], ka = k0 + ak[t], j = j0 + br[t], kb = k0 + bk[t];
      const bool oa = i < g.M && ka < g.K, ob = j < g.N && kb < g.K;
      const float va = g.A[oa ? (long long)i * g.sa_i + (long long)ka * g.sa_k : 0];
      const float vb = g.B[ob ? (long long)kb * g.sb_k + (long long)j * g.sb_j : 0];
      ra[t] = oa ? va : 0.f;
      rb[t] = ob ? vb : 0.f;
    }
  };
  auto put = [&](int buf) {
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
      As[buf][ar[t]][ak[t]] = to_bf16(ra[t]);
      Bs[buf][br[t]][bk[t]] = to_bf16(rb[t]);
    }
  };
  load(0);
  put(0);
  __syncthreads();
  const int nsteps = (g.K + kBK - 1) / kBK;
  for (int st = 0; st < nsteps; ++st) {
    const int cur = st & 1;
    load((st + 1) * kBK);  // past the end: every element is outside, zero
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      af[a] = *reinterpret_cast<const bf16x8*>(&As[cur][wm * 32 + a * 16 + c][8 * q]);
      bfr[a] = *reinterpret_cast<const bf16x8*>(&Bs[cur][wn * 32 + a * 16 + c][8 * q]);
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    if (st + 1 < nsteps) put(cur ^ 1);
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

// out[j] = sum_i X[i*ld + j] over rows, in a fixed order (deterministic):
// thread (g, c) of a block sums rows g, g + 16, ... of column
// 16 blockIdx.x + c in row order (8 loads in flight), then the 16 row
// groups' sums are added in group order: the same order on every launch.
constexpr int kSumCols = 16, kSumGroups = 16;

__global__ void __launch_bounds__(256) column_sums(const float* __restrict__ X, int rows,
                                                   int cols, long long ld,
                                                   float* __restrict__ out) {
  __shared__ float part[kSumGroups][kSumCols];
  const int c = threadIdx.x % kSumCols, grp = threadIdx.x / kSumCols;
  const int j = blockIdx.x * kSumCols + c;
  const int jc = j < cols ? j : cols - 1;  // loads stay in the matrix
  float s = 0.f;
  int i = grp;
  for (; i + 7 * kSumGroups < rows; i += 8 * kSumGroups) {
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = X[(long long)(i + u * kSumGroups) * ld + jc];
#pragma unroll
    for (int u = 0; u < 8; ++u) s += v[u];
  }
  for (; i < rows; i += kSumGroups) s += X[(long long)i * ld + jc];
  part[grp][c] = s;
  __syncthreads();
  if (grp == 0 && j < cols) {
    float t = part[0][c];
    for (int q = 1; q < kSumGroups; ++q) t += part[q][c];
    out[j] = t;
  }
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
  hipLaunchKernelGGL(mlp::column_sums, dim3((cols + mlp::kSumCols - 1) / mlp::kSumCols),
                     dim3(256), 0, s, X, rows, cols, ld, out);
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
