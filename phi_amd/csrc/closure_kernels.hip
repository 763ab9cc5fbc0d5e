// Data movement of the learned truncation closure: network inputs from the
// ADOs of the truncation tier, training targets from a deeper hierarchy,
// and the predicted next-tier ADOs written into the "virtual" neighbour
// slots that the stage kernels gather from.
//
// Row r = m * nb + b is boundary node b (local id first + b) of member m.
// Input row:  [Re rho_n / s (N^2, column-major), Im rho_n / s (N^2), n / L (D)]
// Target/output row: for j = 0 .. D-1: [Re rho_{n+e_j} / s (N^2), Im (N^2)]
// with s = the largest |Re| or |Im| entry of rho_n (at least 1e-30), per row.
// The stage kernels read a next-tier neighbour n + e_j of n only in column
// site(j) (element (a, b) of node n gathers element (a, b) of n + e_j for
// the modes j of site b), so at run time only those outputs are evaluated
// (the fused pass) or written (the scatter): 2 N D of the 2 N^2 D; the rest
// of a virtual slot is never read.
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

// Network outputs -> virtual slots: virtual node vbase + b*D + j of member m
// holds rho_{n+e_j} of boundary node b, of which column site[j] (the read
// one) is written; the guard checks those outputs.
template <typename TO>
__global__ void __launch_bounds__(256) scatter(const float* __restrict__ Y,
                                               const float* __restrict__ S,
                                               const int* __restrict__ moff, int members,
                                               int nb, int N, int D,
                                               const int* __restrict__ site, int vbase,
                                               TO* __restrict__ out, ClosureGuard guard) {
  typedef typename C2T<TO>::type V;
  __shared__ int bad;
  const long long row = blockIdx.x;
  if (row >= (long long)members * nb) return;
  const int m = (int)(row / nb), b = (int)(row - (long long)m * nb);
  const int ne = N * N;
  const float s = S[row];
  const float* y = Y + row * (size_t)(2 * ne * D);
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < 2 * N * D; e += blockDim.x) {
    const int j = e / (2 * N), k = e - j * 2 * N, part = k / N, w = k - part * N;
    if (!(fabsf(y[(size_t)j * 2 * ne + part * ne + site[j] * N + w]) <= guard.ymax)) bad = 1;
  }
  __syncthreads();
  const float keep = bad ? 0.f : 1.f;
  V* dst = reinterpret_cast<V*>(out) + ((size_t)moff[m] + vbase + (size_t)b * D) * ne;
  for (int e = threadIdx.x; e < N * D; e += blockDim.x) {
    const int j = e / N, w = e - j * N;
    const size_t o = (size_t)j * 2 * ne + (size_t)site[j] * N + w;
    V v;
    v.x = (TO)(keep * y[o] * s);
    v.y = (TO)(keep * y[o + ne] * s);
    dst[(size_t)j * ne + (size_t)site[j] * N + w] = v;
  }
  if (threadIdx.x == 0 && bad) atomicAdd(guard.fallen, 1ULL);
}

// --- Fused closure pass ------------------------------------------------------------
// One block of 512 threads (8 wavefronts) takes kRows = 128 rows end to
// end: network inputs from the stage input (as `inputs`), the three layers
// on the matrix cores (v_mfma_f32_16x16x32_bf16; activations in LDS as
// bfloat16, each hidden layer's output written over its input), and the
// scatter of the outputs into the virtual slots (as `scatter`).  Same
// arithmetic as the unfused path: operands rounded to bfloat16, float32
// sums in the same K order, bias and GELU in float32.
//
// A layer runs over groups of 256 output columns (the packed weights have
// 256-row multiples, zero padded).  Within a group, wavefront w owns the
// 64 x 64 output tile of row half w & 1 and column chunk w >> 1.  The
// group's weights stream through LDS in K steps of 32: each thread loads
// two 16-byte pieces of every step (coalesced, two steps ahead in
// registers) and stores them in MFMA B-fragment order, piece (row, k-octet
// o) at fragment (row / 16) * 64 + 16 o + row % 16, so a wavefront reads
// each of its four fragments as one contiguous 1 KB block and every weight
// is fetched once per block.
constexpr int kRows = 128;
constexpr int kThreads = 512;
constexpr int kGroup = 256;
constexpr int kStepK = 32;
constexpr int kStageFrags = kGroup * kStepK / 8;  // 16-byte fragments per stage (16 KB)
constexpr size_t kStageBytes = (size_t)2 * kStageFrags * 16;

template <class EPI>
__device__ __forceinline__ void fused_layer(const unsigned short* act, int stride,
                                            const unsigned short* __restrict__ W, int K,
                                            int ncols, const float* __restrict__ bias,
                                            bf16x8* stage, const EPI& epi) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane >> 4, c = lane & 15;
  const int rh = wave & 1, cw = wave >> 1;
  const int nsteps = K / kStepK;  // even: K is a multiple of 64
  // This thread's two pieces of every stage.
  int poff[2], pfrag[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int pc = tid + kThreads * j, row = pc >> 2, oct = pc & 3;
    poff[j] = row * K + 8 * oct;
    pfrag[j] = (row >> 4) * 64 + 16 * oct + (row & 15);
  }
  const unsigned short* arow = act + (64 * rh + c) * stride + 8 * q;
  for (int g0 = 0; g0 < ncols; g0 += kGroup) {
    const unsigned short* Wg = W + (size_t)g0 * K;
    const bool busy = g0 + 64 * cw < ncols;  // uniform over the wavefront
    floatx4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = (floatx4){0.f, 0.f, 0.f, 0.f};
    auto load = [&](bf16x8* r, int s) {
#pragma unroll
      for (int j = 0; j < 2; ++j)
        r[j] = *reinterpret_cast<const bf16x8*>(Wg + poff[j] + s * kStepK);
    };
    auto put = [&](const bf16x8* r, int s) {
      bf16x8* st = stage + (s & 1) * kStageFrags;
#pragma unroll
      for (int j = 0; j < 2; ++j) st[pfrag[j]] = r[j];
    };
    auto compute = [&](int s) {
      if (!busy) return;
      const bf16x8* st = stage + (s & 1) * kStageFrags + 4 * cw * 64 + lane;
      bf16x8 bv[4], av[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = st[64 * b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
        av[a] = *reinterpret_cast<const bf16x8*>(arow + 16 * a * stride + s * kStepK);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a], bv[b], acc[a][b], 0, 0, 0);
    };
    // Step s computes from stage s & 1 while step s + 1 waits in registers
    // to be stored into the other stage and step s + 2 loads (steps past
    // the end reload the last one: no branch around loads).
    bf16x8 r0[2], r1[2];
    load(r0, 0);
    load(r1, 1);
    put(r0, 0);
    __syncthreads();
    for (int s = 0; s < nsteps; s += 2) {
      load(r0, min(s + 2, nsteps - 1));
      compute(s);
      put(r1, s + 1);
      __syncthreads();
      load(r1, min(s + 3, nsteps - 1));
      compute(s + 1);
      if (s + 2 < nsteps) put(r0, s + 2);
      __syncthreads();
    }
    if (busy) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int col = g0 + 64 * cw + 16 * b + c;
        if (col < ncols) {
          const float bc = bias[col];
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) epi(64 * rh + 16 * a + 4 * q + r, col, acc[a][b][r] + bc);
        }
      }
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

// Outputs straight into the virtual slots: the packed output neurons are
// the network's outputs the stage kernels read, in slot order (packed output
// 2 (j N + w) + part is element w of column site(j) of slot j, real or
// imaginary part), and slotoff[col] is the float offset of packed output
// col from the row's first slot.
template <typename TO>
struct ToSlots {
  TO* slab;
  const long long* rowbase;  // per block row: first virtual slot's real offset, or -1
  const float* srow;
  int* rowbad;               // per block row: an output failed the guard
  const int* slotoff;
  float ymax;
  __device__ __forceinline__ void operator()(int row, int col, float v) const {
    const long long base = rowbase[row];
    if (base < 0) return;
    slab[base + slotoff[col]] = (TO)(v * srow[row]);
    if (!(fabsf(v) <= ymax)) rowbad[row] = 1;
  }
};

template <typename TI>
__global__ void __launch_bounds__(kThreads) fused(FusedNet net, TI* slab,
                                                  const int* __restrict__ moff, int members,
                                                  int first, int nb, int ne, int D,
                                                  const float* __restrict__ bmi, int vbase,
                                                  const int* __restrict__ slotoff,
                                                  ClosureGuard guard) {
  typedef typename C2T<TI>::type V;
  extern __shared__ __align__(16) unsigned short lds[];
  __shared__ float srow[kRows];
  __shared__ long long rowbase[kRows];
  __shared__ long long rowsrc[kRows];
  __shared__ unsigned smax[kRows];
  __shared__ int rowbad[kRows];
  const int stride = net.width + 8;
  // The activations (each layer's output over its input), then the two
  // weight stages.
  unsigned short* A0 = lds;
  bf16x8* stage = reinterpret_cast<bf16x8*>(lds + kRows * stride);
  const long long R = (long long)members * nb;
  const long long row0 = (long long)blockIdx.x * kRows;
  const int tid = threadIdx.x;
  // Inputs, for all the block's rows at once (one trip to memory for the
  // member offsets, one for the ADOs): each row's source node and slots.
  if (tid < kRows) {
    const long long row = row0 + tid;
    smax[tid] = 0u;
    rowbad[tid] = 0;
    if (row < R) {
      const int m = (int)(row / nb), b = (int)(row - (long long)m * nb);
      const long long mo = moff[m];
      rowsrc[tid] = (mo + first + b) * ne;
      rowbase[tid] = 2 * ((mo + vbase + (long long)b * D) * ne);
    } else {
      rowsrc[tid] = -1;
      rowbase[tid] = -1;
    }
  }
  __syncthreads();
  // Elements (row, e) of the tile, several per thread: largest |Re|, |Im|
  // of each row (non-negative floats order as their bits).
  const int tot = kRows * ne;
  constexpr int kPer = 16;  // enough for ne <= 64 (N <= 8)
  V v[kPer];
#pragma unroll
  for (int t = 0; t < kPer; ++t) {
    const int i = tid + t * kThreads;
    const int r = i / ne;
    const long long src = i < tot ? rowsrc[r] : -1;
    v[t] = src >= 0 ? reinterpret_cast<const V*>(slab)[src + (i - r * ne)] : V{};
  }
#pragma unroll
  for (int t = 0; t < kPer; ++t) {
    const int i = tid + t * kThreads;
    if (i < tot) {
      const float mx = fmaxf(fabsf((float)v[t].x), fabsf((float)v[t].y));
      atomicMax(&smax[i / ne], __float_as_uint(mx));
    }
  }
  __syncthreads();
  if (tid < kRows) srow[tid] = fmaxf(__uint_as_float(smax[tid]), 1e-30f);
  __syncthreads();
#pragma unroll
  for (int t = 0; t < kPer; ++t) {
    const int i = tid + t * kThreads;
    if (i < tot) {
      const int r = i / ne, e = i - r * ne;
      const float inv = 1.0f / srow[r];
      unsigned short* x = A0 + r * stride;
      x[e] = to_bf16((float)v[t].x * inv);
      x[ne + e] = to_bf16((float)v[t].y * inv);
    }
  }
  // Boundary multi-index columns and the zero padding up to k1.
  const int extra = net.k1 - 2 * ne;
  for (int i = tid; i < kRows * extra; i += kThreads) {
    const int r = i / extra, k = i - r * extra;
    const long long row = row0 + r;
    unsigned short val = 0;
    if (k < D && row < R) {
      const int b = (int)(row % nb);
      val = to_bf16(bmi[(size_t)b * D + k]);
    }
    A0[r * stride + 2 * ne + k] = val;
  }
  __syncthreads();
  // Hidden layers write their output over their input: every wavefront has
  // finished reading it at the layer's last barrier.  One copy of the layer
  // code serves both.
#pragma unroll 1
  for (int layer = 0; layer < 2; ++layer) {
    fused_layer(A0, stride, layer ? net.w2 : net.w1, layer ? net.hid : net.k1, net.hid,
                layer ? net.b2 : net.b1, stage, ToLds{A0, stride});
    __syncthreads();
  }
  fused_layer(A0, stride, net.w3, net.hid, net.n_out, net.b3, stage,
              ToSlots<TI>{slab, rowbase, srow, rowbad, slotoff, guard.ymax});
  // Rows whose outputs failed the guard fall back to plain truncation: zero
  // slots (a second write after every wavefront's first one).
  __syncthreads();
  int nbad = 0;
  for (int r = 0; r < kRows; ++r) {
    if (!rowbad[r]) continue;
    ++nbad;
    for (int col = tid; col < net.n_out; col += kThreads) slab[rowbase[r] + slotoff[col]] = (TI)0;
  }
  if (tid == 0 && nbad) atomicAdd(guard.fallen, (unsigned long long)nbad);
}

}  // namespace closure

size_t FusedClosureLds(const FusedNet& net) {
  return (size_t)closure::kRows * (net.width + 8) * sizeof(unsigned short) +
         closure::kStageBytes;
}

bool FusedClosureFits(const FusedNet& net) {
  // The activations and two weight stages in one CU's 160 KB of LDS, with
  // the block's per-row tables; K steps in pairs of 32; hidden layers of
  // one column group (their output overwrites their input).
  return FusedClosureLds(net) + closure::kRows * 32 <= (size_t)160 * 1024 &&
         net.k1 % 64 == 0 && net.hid % 64 == 0 && net.hid <= closure::kGroup;
}

hipError_t LaunchClosureFused(const FusedNet& net, void* slab, bool fp32, const int* moff,
                              int members, int first, int nb, int ne, int D, const float* bmi,
                              int vbase, const int* slotoff, const ClosureGuard& guard,
                              hipStream_t s) {
  const long long rows = (long long)members * nb;
  const unsigned blocks = (unsigned)((rows + closure::kRows - 1) / closure::kRows);
  const size_t lds = FusedClosureLds(net);
  if (!FusedClosureFits(net) || ne > 64) return hipErrorInvalidConfiguration;
  const void* k = fp32 ? reinterpret_cast<const void*>(&closure::fused<float>)
                       : reinterpret_cast<const void*>(&closure::fused<double>);
  hipError_t e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (fp32)
    hipLaunchKernelGGL(closure::fused<float>, dim3(blocks), dim3(closure::kThreads), lds, s, net,
                       (float*)slab, moff, members, first, nb, ne, D, bmi, vbase, slotoff, guard);
  else
    hipLaunchKernelGGL(closure::fused<double>, dim3(blocks), dim3(closure::kThreads), lds, s, net,
                       (double*)slab, moff, members, first, nb, ne, D, bmi, vbase, slotoff, guard);
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
                                int nb, int N, int D, const int* site, int vbase, void* out,
                                bool out_fp32, const ClosureGuard& guard, hipStream_t s) {
  const long long rows = (long long)members * nb;
  if (out_fp32)
    hipLaunchKernelGGL(closure::scatter<float>, dim3((unsigned)rows), dim3(256), 0, s, Y, S,
                       moff, members, nb, N, D, site, vbase, (float*)out, guard);
  else
    hipLaunchKernelGGL(closure::scatter<double>, dim3((unsigned)rows), dim3(256), 0, s, Y, S,
                       moff, members, nb, N, D, site, vbase, (double*)out, guard);
  return hipGetLastError();
}

}  // namespace phi
