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
#include "mlp_act.h"

namespace phi {
namespace closure {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

using act::gelu;
using act::to_bf16;

// Largest x over the wavefront, in every lane: rotations by 1, 2, 4 and 8
// within each row of 16 lanes (DPP row_ror, no LDS), then the four rows'
// results read as scalars.
__device__ __forceinline__ float wave_max(float x) {
#define PHI_ROR(n)                                                                          \
  x = fmaxf(x, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(                      \
                   0, __builtin_bit_cast(int, x), 0x120 + (n), 0xf, 0xf, false)))
  PHI_ROR(1);
  PHI_ROR(2);
  PHI_ROR(4);
  PHI_ROR(8);
#undef PHI_ROR
  const int b = __builtin_bit_cast(int, x);
  return fmaxf(fmaxf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 0)),
                     __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 16))),
               fmaxf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 32)),
                     __builtin_bit_cast(float, __builtin_amdgcn_readlane(b, 48))));
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
// A layer runs over groups of output columns (the packed weights have
// 256-row multiples, zero padded).  The 8 wavefronts split the block's 128
// rows into RW row groups and a group's columns into 8 / RW chunks of 64:
// RW = 2 (64 x 64 tiles, groups of 256 columns) for the hidden layers, RW =
// 4 (32 x 64 tiles, groups of 128) for an output layer of at most 128
// columns, so that no wavefront idles.  Tiles are computed transposed
// (weights as the MFMA's A operand, activations as B): a lane's four
// accumulator entries are then four consecutive output columns of one row,
// stored as one 8-byte LDS write.  The group's weights stream through LDS
// in K steps of 64 (two MFMA k-halves of 32, one barrier per step): each
// thread loads 16-byte pieces of every step (coalesced, two steps ahead in
// registers) and stores piece (row, k-octet o) at slot 16 o + (row % 16 ^
// 2 o % 16) of its 16-row block of 128 slots, so that a wavefront reads each
// of its fragments as one 1 KB block, the 16 pieces that 16 consecutive
// lanes store (2 rows x 8 octets) fall on distinct banks, and every weight
// is fetched once per block.
constexpr int kRows = 128;
constexpr int kThreads = 512;
constexpr int kGroup = 256;
constexpr int kStepK = 64;
constexpr int kOcts = kStepK / 8;                 // k-octets per row and step
constexpr int kStageFrags = kGroup * kOcts;       // 16-byte fragments per stage (32 KB)
constexpr size_t kStageBytes = (size_t)2 * kStageFrags * 16;

__device__ __forceinline__ int Slot(int row, int oct) {
  return (row >> 4) * (16 * kOcts) + 16 * oct + ((row & 15) ^ ((2 * oct) & 15));
}

template <int RW, class EPI>
__device__ __forceinline__ void fused_layer(const unsigned short* act, int stride,
                                            const unsigned short* __restrict__ W, int K,
                                            int ncols, const float* __restrict__ bias,
                                            bf16x8* stage, const EPI& epi) {
  constexpr int kCW = 8 / RW;                           // column chunks of 64
  constexpr int kGrp = 64 * kCW;                        // columns per group
  constexpr int kNA = kRows / RW / 16;                  // 16-row blocks per wavefront
  constexpr int kPieces = kGrp * kOcts / kThreads;      // 16-byte pieces per thread and step
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int q = lane >> 4, c = lane & 15;
  const int rq = wave % RW, cw = wave / RW;
  const int nsteps = K / kStepK;  // K is a multiple of 64
  // This thread's pieces of every stage.
  int poff[kPieces], pslot[kPieces];
#pragma unroll
  for (int j = 0; j < kPieces; ++j) {
    const int pc = tid + kThreads * j, row = pc / kOcts, oct = pc % kOcts;
    poff[j] = row * K + 8 * oct;
    pslot[j] = Slot(row, oct);
  }
  const int wrow0 = (kRows / RW) * rq;
  const unsigned short* arow = act + (wrow0 + c) * stride + 8 * q;
  // This lane's fragment of k-half h of a 16-row block: octet 4 h + q, row c.
  const int myslot[2] = {Slot(c, q), Slot(c, 4 + q)};
  for (int g0 = 0; g0 < ncols; g0 += kGrp) {
    const unsigned short* Wg = W + (size_t)g0 * K;
    const bool busy = g0 + 64 * cw < ncols;  // uniform over the wavefront
    // The epilogue's loads, issued before the K loop hides their latency
    // and unconditional so that they need no wait of their own (columns
    // past ncols read the zero padding of the packed biases and rows).
    floatx4 bc[4];
    typename EPI::Pre pre[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int col0 = g0 + 64 * cw + 16 * b + 4 * q;
      bc[b] = *reinterpret_cast<const floatx4*>(bias + col0);
      pre[b] = epi.pre(col0);
    }
    floatx4 acc[kNA][4];
#pragma unroll
    for (int a = 0; a < kNA; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = (floatx4){0.f, 0.f, 0.f, 0.f};
    auto load = [&](bf16x8* r, int s) {
#pragma unroll
      for (int j = 0; j < kPieces; ++j)
        r[j] = *reinterpret_cast<const bf16x8*>(Wg + poff[j] + s * kStepK);
    };
    auto put = [&](const bf16x8* r, int s) {
      bf16x8* st = stage + (s & 1) * kStageFrags;
#pragma unroll
      for (int j = 0; j < kPieces; ++j) st[pslot[j]] = r[j];
    };
    auto compute = [&](int s) {
      if (!busy) return;
      const bf16x8* st = stage + (s & 1) * kStageFrags + 4 * cw * (16 * kOcts);
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        bf16x8 bv[4], av[kNA];
#pragma unroll
        for (int b = 0; b < 4; ++b) bv[b] = st[b * (16 * kOcts) + myslot[h]];
#pragma unroll
        for (int a = 0; a < kNA; ++a)
          av[a] = *reinterpret_cast<const bf16x8*>(arow + 16 * a * stride + s * kStepK + 32 * h);
#pragma unroll
        for (int a = 0; a < kNA; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b)
            acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bv[b], av[a], acc[a][b], 0, 0, 0);
      }
    };
    // Step s computes from stage s & 1 while step s + 1 waits in registers
    // to be stored into the other stage and step s + 2 loads (steps past
    // the end reload the last one: no branch around loads).  The step
    // count is uniform over the block, so are the branches around barriers.
    bf16x8 r0[kPieces], r1[kPieces];
    load(r0, 0);
    load(r1, min(1, nsteps - 1));
    put(r0, 0);
    __syncthreads();
    for (int s = 0; s < nsteps; s += 2) {
      load(r0, min(s + 2, nsteps - 1));
      compute(s);
      if (s + 1 < nsteps) put(r1, s + 1);
      __syncthreads();
      if (s + 1 < nsteps) {
        load(r1, min(s + 3, nsteps - 1));
        compute(s + 1);
        if (s + 2 < nsteps) put(r0, s + 2);
        __syncthreads();
      }
    }
    if (busy) {
      // Entry r of acc[a][b]: row wrow0 + 16 a + c, column col0 + r.
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int col0 = g0 + 64 * cw + 16 * b + 4 * q;
        if (col0 < ncols) {
#pragma unroll
          for (int a = 0; a < kNA; ++a) epi(wrow0 + 16 * a + c, col0, acc[a][b] + bc[b], pre[b]);
        }
      }
    }
  }
}

// Hidden layers: GELU, bfloat16, four columns in one 8-byte LDS write.
struct ToLds {
  struct Pre {};
  unsigned short* out;
  int stride;
  __device__ __forceinline__ Pre pre(int) const { return {}; }
  __device__ __forceinline__ void operator()(int row, int col0, floatx4 v, Pre) const {
    bf16x4 o;
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = (__bf16)gelu(v[r]);
    *reinterpret_cast<bf16x4*>(out + row * stride + col0) = o;
  }
};

// Outputs straight into the virtual slots: the packed output neurons are
// the network's outputs the stage kernels read, in slot order (packed output
// 2 (j N + w) + part is element w of column site(j) of slot j, real or
// imaginary part), and slotoff[col] is the float offset of packed output
// col from the row's first slot, so an even column and the next one are the
// real and imaginary part of one complex element (one vector store; the
// packed count 2 N D is even).
template <typename TO>
struct ToSlots {
  typedef typename C2T<TO>::type V;
  TO* slab;
  const long long* rowbase;  // per block row: first virtual slot's real offset, or -1
  const float* srow;
  int* rowbad;               // per block row: an output failed the guard
  const int* slotoff;
  int ncols;
  float ymax;
  // Slot offsets of the pairs (col0, col0 + 1) and (col0 + 2, col0 + 3)
  // (indices clamped into the table; a pair past ncols is not stored).
  typedef int2 Pre;
  __device__ __forceinline__ Pre pre(int col0) const {
    return int2{slotoff[min(col0, ncols - 2)], slotoff[min(col0 + 2, ncols - 2)]};
  }
  __device__ __forceinline__ void operator()(int row, int col0, floatx4 v, Pre off) const {
    const long long base = rowbase[row];
    if (base < 0) return;
    const float s = srow[row];
    bool bad = false;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int col = col0 + 2 * h;
      if (col < ncols) {
        V z;
        z.x = (TO)(v[2 * h] * s);
        z.y = (TO)(v[2 * h + 1] * s);
        *reinterpret_cast<V*>(slab + base + (h ? off.y : off.x)) = z;
        bad |= !(fabsf(v[2 * h]) <= ymax) || !(fabsf(v[2 * h + 1]) <= ymax);
      }
    }
    if (bad) rowbad[row] = 1;
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
  constexpr int kRowsPerWave = kRows / (kThreads / 64);
  extern __shared__ __align__(16) unsigned short lds[];
  __shared__ float srow[kRows];
  __shared__ long long rowbase[kRows];
  __shared__ long long rowsrc[kRows];
  __shared__ int rowb[kRows];
  __shared__ int rowbad[kRows];
  const int stride = net.width + 8;
  // The activations (each layer's output over its input), then the two
  // weight stages.
  unsigned short* A0 = lds;
  bf16x8* stage = reinterpret_cast<bf16x8*>(lds + kRows * stride);
  const long long R = (long long)members * nb;
  const long long row0 = (long long)blockIdx.x * kRows;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // Each row's source node, boundary node and slots.
  if (tid < kRows) {
    const long long row = row0 + tid;
    rowbad[tid] = 0;
    if (row < R) {
      const int m = (int)(row / nb), b = (int)(row - (long long)m * nb);
      const long long mo = moff[m];
      rowsrc[tid] = (mo + first + b) * ne;
      rowbase[tid] = 2 * ((mo + vbase + (long long)b * D) * ne);
      rowb[tid] = b;
    } else {
      rowsrc[tid] = -1;
      rowbase[tid] = -1;
      rowb[tid] = -1;
    }
  }
  __syncthreads();
  // Inputs: wavefront w takes rows w * 16 .. w * 16 + 15, lane e element e
  // of each (ne <= 64) and boundary multi-index entry e (D <= 64); a row's
  // scale s (largest |Re|, |Im|) is a wavefront maximum.  Then the scaled
  // elements and the multi-index and zero padding columns up to k1.  The
  // loads are unconditional (lanes and rows without data read element 0 and
  // drop it), so all of them are in flight at once: a load under a branch
  // gets its own wait for memory.
  V v[kRowsPerWave];
  float bm[kRowsPerWave];
#pragma unroll
  for (int t = 0; t < kRowsPerWave; ++t) {
    const int r = wave * kRowsPerWave + t;
    const long long src = rowsrc[r];
    const int b = rowb[r];
    const bool has_v = src >= 0 && lane < ne, has_b = b >= 0 && lane < D;
    const V x = reinterpret_cast<const V*>(slab)[has_v ? src + lane : 0];
    const float y = bmi[has_b ? (size_t)b * D + lane : 0];
    v[t] = has_v ? x : V{};
    bm[t] = has_b ? y : 0.f;
  }
  const int extra = net.k1 - 2 * ne;
#pragma unroll
  for (int t = 0; t < kRowsPerWave; ++t) {
    const int r = wave * kRowsPerWave + t;
    const float sc = fmaxf(wave_max(fmaxf(fabsf((float)v[t].x), fabsf((float)v[t].y))), 1e-30f);
    const float inv = 1.0f / sc;
    if (lane == 0) srow[r] = sc;
    unsigned short* x = A0 + r * stride;
    if (lane < ne) {
      x[lane] = to_bf16((float)v[t].x * inv);
      x[ne + lane] = to_bf16((float)v[t].y * inv);
    }
    for (int k = lane; k < extra; k += 64) x[2 * ne + k] = k < D ? to_bf16(bm[t]) : 0;
  }
  __syncthreads();
  // Hidden layers write their output over their input: every wavefront has
  // finished reading it at the layer's last barrier.  One copy of the layer
  // code serves both.
#pragma unroll 1
  for (int layer = 0; layer < 2; ++layer) {
    fused_layer<2>(A0, stride, layer ? net.w2 : net.w1, layer ? net.hid : net.k1, net.hid,
                   layer ? net.b2 : net.b1, stage, ToLds{A0, stride});
    __syncthreads();
  }
  const ToSlots<TI> to_slots{slab, rowbase, srow, rowbad, slotoff, net.n_out, guard.ymax};
  if (net.n_out <= 128)
    fused_layer<4>(A0, stride, net.w3, net.hid, net.n_out, net.b3, stage, to_slots);
  else
    fused_layer<2>(A0, stride, net.w3, net.hid, net.n_out, net.b3, stage, to_slots);
  // Rows whose outputs failed the guard fall back to plain truncation: zero
  // slots (a second write after every wavefront's first one).
  __syncthreads();
  const int nbad = __syncthreads_count(tid < kRows && rowbad[tid]);
  if (nbad) {
    for (int i = tid; i < kRows * net.n_out; i += kThreads) {
      const int r = i / net.n_out, col = i - r * net.n_out;
      if (rowbad[r]) slab[rowbase[r] + slotoff[col]] = (TI)0;
    }
    if (tid == 0) atomicAdd(guard.fallen, (unsigned long long)nbad);
  }
}

}  // namespace closure

size_t FusedClosureLds(const FusedNet& net) {
  return (size_t)closure::kRows * (net.width + 8) * sizeof(unsigned short) +
         closure::kStageBytes;
}

bool FusedClosureFits(const FusedNet& net) {
  // The activations and two weight stages in one CU's 160 KB of LDS, with
  // the block's per-row tables; K steps of 64; hidden layers of one column
  // group (their output overwrites their input); outputs in (real,
  // imaginary) pairs.
  return FusedClosureLds(net) + closure::kRows * 32 <= (size_t)160 * 1024 &&
         net.k1 % closure::kStepK == 0 && net.hid % closure::kStepK == 0 &&
         net.hid <= closure::kGroup && net.n_out >= 2 && net.n_out % 2 == 0;
}

hipError_t LaunchClosureFused(const FusedNet& net, void* slab, bool fp32, const int* moff,
                              int members, int first, int nb, int ne, int D, const float* bmi,
                              int vbase, const int* slotoff, const ClosureGuard& guard,
                              hipStream_t s) {
  const long long rows = (long long)members * nb;
  const unsigned blocks = (unsigned)((rows + closure::kRows - 1) / closure::kRows);
  const size_t lds = FusedClosureLds(net);
  if (!FusedClosureFits(net) || ne > 64 || D > 64) return hipErrorInvalidConfiguration;
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
