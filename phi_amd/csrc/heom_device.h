// Device templates of the HEOM stage kernels (included by the per-precision
// translation units heom_rk4_*.hip / heom_rkf45_*.hip and heom_kernels.hip).
#ifndef PHI_AMD_HEOM_DEVICE_H_
#define PHI_AMD_HEOM_DEVICE_H_

#include <hip/hip_runtime.h>

#include <cstdint>

#include "heom_kernels.h"

// State layout in HBM: one slab per buffer, node-major, each node an N x N
// complex matrix stored column-major with interleaved (re, im):
//   buf[(node * N*N + col * N + row) * 2 + {0,1}]
// ADOs are Hermitian (every term of the IndependentBath HEOM maps Hermitian
// matrices to Hermitian matrices), so the time-stepping kernels use
//   k = Y' + Y'^dagger,
//   Y' = -(i/hbar) H Y - mu/2 Y - (Delta_a + Delta_b)/2 (a != b) Y + A,
//   A(:, m) = sum_{modes j of site m} i s_j Y_{n+e_j}(:, m) - col_j Y_{n-e_j}(:, m)
// which is one N x N GEMM per node and a gather of one column per neighbour.
// It equals PHI's IndependentBath::Get (hierarchy_updater.h:36-75,
// hierarchy_updater.cc:129-144, 216-254) on Hermitian input, and keeps the
// state exactly Hermitian.  The general kernel (rhs_general) applies the
// operator to any complex input (PHI's Get semantics verbatim) and serves
// the steady-state solver and the per-call RHS entry point.
//
// Precision: TB is the bulk type (double, or float for FP32-bulk).  The RK
// state, its accumulator, the RKF45 error norm and all reductions are FP64.
namespace phi {
namespace dev {

template <typename T> struct C2;
template <> struct C2<float> { typedef float2 type; };
template <> struct C2<double> { typedef double2 type; };

template <typename T>
__device__ __forceinline__ typename C2<T>::type mk(T x, T y) {
  typename C2<T>::type r;
  r.x = x;
  r.y = y;
  return r;
}

// Loads interleaved complex element e of a slab of real type TI as TB.
template <typename TB, typename TI>
__device__ __forceinline__ typename C2<TB>::type ld(const TI* __restrict__ p, size_t e) {
  typedef typename C2<TI>::type VI;
  const VI v = reinterpret_cast<const VI*>(p)[e];
  return mk<TB>((TB)v.x, (TB)v.y);
}

template <typename T>
__device__ __forceinline__ void st(T* __restrict__ p, size_t e, double re, double im) {
  typedef typename C2<T>::type V;
  V v;
  v.x = (T)re;
  v.y = (T)im;
  reinterpret_cast<V*>(p)[e] = v;
}

// Division by a launch constant d (N or N*N): q = (n * m) >> 32 with
// m = ceil(2^32 / d) is exact for n * d < 2^32, which holds for every
// in-block element index here (n < 2^20, d < 2^12).
__device__ __forceinline__ int fdiv(int n, unsigned long long m) {
  return (int)(((unsigned long long)(unsigned)n * m) >> 32);
}

// --- Hermitian RHS on a tile of nodes -----------------------------------------
// A block owns the nodes [base, base + nt).  Its LDS holds, per launch:
//   H (N*N), the tile's stage input Y and the work buffer W (nt*N*N each),
//   the edge-prefactor tables (D*L), Delta (N), mu/2 (nt), the site->mode
//   CSR and the tile's neighbour ids and occupations (nt*D each).
// Every one of these loads, and the epilogue operands each thread keeps in
// registers, is issued before the first barrier; the only dependent global
// round trip after it is the neighbour-column gather.
template <typename TB>
struct Lds {
  typedef typename C2<TB>::type V;
  V* H;
  V* Y;
  V* W;
  V* pcol;
  TB* ns;
  TB* delta;
  TB* mu2;
  int* sp;
  int* sm;
  int* nx;
  int* pv;
  uint8_t* oc;
};

__host__ __device__ inline size_t Align16(size_t x) { return (x + 15) & ~(size_t)15; }

// Byte size of the LDS layout above (host and device agree on it).
// The matrix-core path keeps H in registers and has no LDS copy (with_h).
__host__ __device__ inline size_t LdsBytes(int N, int D, int L, int tile, size_t real,
                                           bool with_h = true) {
  const size_t NE = (size_t)N * N, V = 2 * real;
  const size_t NP = (size_t)N * (N + 1);
  size_t b = V * ((with_h ? NE : 0) + NE * tile + NP * tile + (size_t)D * L);
  b = Align16(b + real * ((size_t)D * L + N + tile));
  b = Align16(b + sizeof(int) * ((size_t)N + 1 + D + 2 * (size_t)tile * D));
  b = Align16(b + (size_t)tile * D);
  return b;
}

// W holds Y' with column stride N + 1 (element (a, b) of tile node t at
// t*N*(N+1) + b*(N+1) + a), which spreads the transposed reads of the
// epilogue over the LDS banks.
template <typename TB>
__device__ __forceinline__ Lds<TB> carve(unsigned char* smem, int N, int D, int L, int tile,
                                         bool with_h = true) {
  typedef typename C2<TB>::type V;
  const int NE = N * N;
  Lds<TB> s;
  s.H = reinterpret_cast<V*>(smem);
  s.Y = s.H + (with_h ? NE : 0);
  s.W = s.Y + NE * tile;
  s.pcol = s.W + N * (N + 1) * tile;
  s.ns = reinterpret_cast<TB*>(s.pcol + D * L);
  s.delta = s.ns + D * L;
  s.mu2 = s.delta + N;
  unsigned char* p = reinterpret_cast<unsigned char*>(s.mu2 + tile);
  p = smem + Align16(p - smem);
  s.sp = reinterpret_cast<int*>(p);
  s.sm = s.sp + N + 1;
  s.nx = s.sm + D;
  s.pv = s.nx + tile * D;
  p = reinterpret_cast<unsigned char*>(s.pv + tile * D);
  p = smem + Align16(p - smem);
  s.oc = p;
  return s;
}

// Phase 0: cooperative loads of everything the tile needs into LDS.
template <typename TB, typename TI>
__device__ __forceinline__ void load_tile(const RhsArgs& A, const TI* __restrict__ in, int base,
                                          int nt, const Lds<TB>& s, bool with_h = true) {
  const int N = A.N, NE = A.NE, D = A.D, L = A.L;
  const int tid = threadIdx.x, nth = blockDim.x;
  if (with_h)
    for (int e = tid; e < NE; e += nth)
      s.H[e] = ld<TB, double>(reinterpret_cast<const double*>(A.H), e);
  const size_t off = (size_t)base * NE;
  for (int e = tid; e < nt * NE; e += nth) s.Y[e] = ld<TB, TI>(in, off + e);
  for (int e = tid; e < D * L; e += nth) {
    s.pcol[e] = ld<TB, double>(reinterpret_cast<const double*>(A.prev_col), e);
    s.ns[e] = (TB)A.next_s[e];
  }
  for (int e = tid; e < N; e += nth) s.delta[e] = (TB)A.delta[e];
  for (int e = tid; e < nt; e += nth) s.mu2[e] = (TB)(0.5 * A.mu[base + e]);
  for (int e = tid; e <= N; e += nth) s.sp[e] = A.site_ptr[e];
  for (int e = tid; e < D; e += nth) s.sm[e] = A.site_modes[e];
  const size_t o = (size_t)base * D;
  for (int e = tid; e < nt * D; e += nth) {
    s.nx[e] = A.next[o + e];
    s.pv[e] = A.prev[o + e];
    s.oc[e] = A.occ[o + e];
  }
}

// Phase 1: W(a,b) = Y'(a,b) = -(i/hbar) (H Y)(a,b) - (mu/2 + (Delta_a +
// Delta_b)/2 [a != b]) Y(a,b) + A(a,b), where column b of A gathers, over the
// modes j coupled to site b, i s_j Y_{n+e_j}(a,b) - col_j Y_{n-e_j}(a,b).
// Consecutive threads take consecutive elements, so the gathers coalesce.
template <typename TB, typename TI>
__device__ __forceinline__ void rhs_phase(const RhsArgs& A, const TI* __restrict__ in, int nt,
                                          const Lds<TB>& s) {
  typedef typename C2<TB>::type V;
  const int N = A.N, NE = A.NE, D = A.D, L = A.L;
  const int NP = N * (N + 1);
  const TB ih = (TB)A.inv_hbar;
  for (int e = threadIdx.x; e < nt * NE; e += blockDim.x) {
    const int t = fdiv(e, A.mNE);
    const int r = e - t * NE;
    const int b = fdiv(r, A.mN);
    const int a = r - b * N;
    TB ax = 0, ay = 0;
    if (!A.vacuum) {
      for (int q = s.sp[b]; q < s.sp[b + 1]; ++q) {
        const int j = s.sm[q];
        const int k = s.oc[t * D + j];
        const int nx = s.nx[t * D + j];
        if (nx >= 0) {
          const TB sc = s.ns[j * L + k];
          const V y = ld<TB, TI>(in, (size_t)nx * NE + r);
          ax -= sc * y.y;
          ay += sc * y.x;
        }
        const int pv = s.pv[t * D + j];
        if (pv >= 0) {
          const V c = s.pcol[j * L + k];
          const V y = ld<TB, TI>(in, (size_t)pv * NE + r);
          ax -= c.x * y.x - c.y * y.y;
          ay -= c.x * y.y + c.y * y.x;
        }
      }
    }
    const V* yc = s.Y + t * NE + b * N;
    TB xr = 0, xi = 0;
    for (int c = 0; c < N; ++c) {
      const V h = s.H[a + c * N];
      const V y = yc[c];
      xr += h.x * y.x - h.y * y.y;
      xi += h.x * y.y + h.y * y.x;
    }
    const V y = s.Y[e];
    TB f = s.mu2[t];
    if (a != b) f += (TB)0.5 * (s.delta[a] + s.delta[b]);
    s.W[t * NP + b * (N + 1) + a] = mk<TB>(ax + ih * xi - f * y.x, ay - ih * xr - f * y.y);
  }
}

// Phase 1, matrix-core form (N <= 32).  The elementwise part
//   W(a,b) = A(a,b) - (mu/2 + (Delta_a + Delta_b)/2 [a != b]) Y(a,b)
// is formed first with coalesced gathers; then each wavefront multiplies
// H (M = N rows, padded to 16 or 32) into 16-column slices of the tile's
// columns (node, b) and adds -(i/hbar) H Y into W.  Lane l of a 16x16x4
// MFMA holds A[l%16][l/16] and B[l/16][l%16]; its output register r holds
// D[4(l/16) + r][l%16] for v_mfma_f32_16x16x4f32 and D[l/16 + 4r][l%16] for
// v_mfma_f64_16x16x4f64.  The K dimension is laid out so that lane group
// g = l/16 covers rows g*KS .. g*KS + KS - 1 (KS = ceil(N/4)), the same
// permutation on both operands.  H fragments are read once from global
// memory into registers; real H (the common case) takes one product per
// real/imaginary part, complex H two.
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

template <typename TB> struct Mfma;
template <> struct Mfma<float> {
  typedef floatx4 acc;
  static __device__ __forceinline__ acc zero() { return (floatx4){0.f, 0.f, 0.f, 0.f}; }
  static __device__ __forceinline__ acc mac(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};
template <> struct Mfma<double> {
  typedef doublex4 acc;
  static __device__ __forceinline__ acc zero() { return (doublex4){0., 0., 0., 0.}; }
  static __device__ __forceinline__ acc mac(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};

template <typename TB, typename TI>
__device__ __forceinline__ void rhs_phase_mfma(const RhsArgs& A, const TI* __restrict__ in,
                                               int nt, const Lds<TB>& s) {
  typedef typename C2<TB>::type V;
  typedef Mfma<TB> M;
  const int N = A.N, NE = A.NE, D = A.D, L = A.L;
  const int NP = N * (N + 1);
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;          // k group / output row group
  const int c16 = lane & 15;        // column within the slice / H row
  const int KS = (N + 3) >> 2;      // k steps
  const int MT = (N + 15) >> 4;     // 16-row tiles of H (1 or 2)
  // H fragments hf[mt][s] = H[16 mt + c16][g KS + s] (zero outside N),
  // issued before the elementwise pass so their latency overlaps it.
  TB hr[2][8], hi[2][8];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int st = 0; st < 8; ++st) {
      const int row = 16 * mt + c16, k = g * KS + st;
      double2 h = make_double2(0., 0.);
      if (mt < MT && st < KS && row < N && k < N) h = A.H[row + k * N];
      hr[mt][st] = (TB)h.x;
      hi[mt][st] = (TB)h.y;
    }
  // (1) elementwise terms
  for (int e = threadIdx.x; e < nt * NE; e += blockDim.x) {
    const int t = fdiv(e, A.mNE);
    const int r = e - t * NE;
    const int b = fdiv(r, A.mN);
    const int a = r - b * N;
    TB ax = 0, ay = 0;
    if (!A.vacuum) {
      for (int q = s.sp[b]; q < s.sp[b + 1]; ++q) {
        const int j = s.sm[q];
        const int k = s.oc[t * D + j];
        const int nx = s.nx[t * D + j];
        if (nx >= 0) {
          const TB sc = s.ns[j * L + k];
          const V y = ld<TB, TI>(in, (size_t)nx * NE + r);
          ax -= sc * y.y;
          ay += sc * y.x;
        }
        const int pv = s.pv[t * D + j];
        if (pv >= 0) {
          const V c = s.pcol[j * L + k];
          const V y = ld<TB, TI>(in, (size_t)pv * NE + r);
          ax -= c.x * y.x - c.y * y.y;
          ay -= c.x * y.y + c.y * y.x;
        }
      }
    }
    const V y = s.Y[e];
    TB f = s.mu2[t];
    if (a != b) f += (TB)0.5 * (s.delta[a] + s.delta[b]);
    s.W[t * NP + b * (N + 1) + a] = mk<TB>(ax - f * y.x, ay - f * y.y);
  }
  __syncthreads();
  // (2) -(i/hbar) H Y on the matrix cores
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const TB ih = (TB)A.inv_hbar;
  const bool realh = A.real_h != 0;
  const int ncols = nt * N;
  for (int slice = wave; slice * 16 < ncols; slice += nwaves) {
    const int j = slice * 16 + c16;            // tile column of this lane
    const bool jok = j < ncols;
    const int tj = jok ? fdiv(j, A.mN) : 0;
    const int bj = j - tj * N;
    const V* ycol = s.Y + tj * NE + bj * N;
    typename M::acc xr[2], xi[2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      xr[mt] = M::zero();
      xi[mt] = M::zero();
    }
#pragma unroll
    for (int st = 0; st < 8; ++st) {
      if (st < KS) {
        const int k = g * KS + st;
        V y = mk<TB>(0, 0);
        if (jok && k < N) y = ycol[k];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
          if (mt < MT) {
            xr[mt] = M::mac(hr[mt][st], y.x, xr[mt]);
            xi[mt] = M::mac(hr[mt][st], y.y, xi[mt]);
            if (!realh) {
              xr[mt] = M::mac(-hi[mt][st], y.y, xr[mt]);
              xi[mt] = M::mac(hi[mt][st], y.x, xi[mt]);
            }
          }
        }
      }
    }
    if (jok) {
      V* wcol = s.W + tj * NP + bj * (N + 1);
#pragma unroll
      for (int mt = 0; mt < 2; ++mt) {
        if (mt < MT) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = 16 * mt + M::row(g, r);
            if (row < N) {
              V w = wcol[row];
              w.x += ih * xi[mt][r];
              w.y -= ih * xr[mt][r];
              wcol[row] = w;
            }
          }
        }
      }
    }
  }
}

// k(a, b) = Y'(a, b) + conj(Y'(b, a)), in double.
template <typename TB>
__device__ __forceinline__ double2 kval(const RhsArgs& A, const Lds<TB>& s, int e) {
  typedef typename C2<TB>::type V;
  const int N = A.N, NE = A.NE;
  const int t = fdiv(e, A.mNE);
  const int r = e - t * NE;
  const int b = fdiv(r, A.mN);
  const int a = r - b * N;
  const int NP = N * (N + 1);
  const V p = s.W[t * NP + b * (N + 1) + a];
  const V q = s.W[t * NP + a * (N + 1) + b];
  double2 k;
  k.x = (double)(p.x + q.x);
  k.y = (double)(p.y - q.y);
  return k;
}

// Elements each thread prefetches for the epilogue (tiles hold at most
// 256 * kPre elements; larger tiles load the rest in the epilogue).
constexpr int kPre = 4;

__device__ __forceinline__ double2 ld2(const double* p, size_t e) {
  return reinterpret_cast<const double2*>(p)[e];
}

// Streaming accesses: elements a kernel reads or writes exactly once (the
// FP64 state and accumulators, the stage output).  The non-temporal hint
// lets L2 evict these lines first, which keeps resident the stage input
// that neighbouring nodes gather from.
typedef double V2dNt __attribute__((ext_vector_type(2)));
typedef float V2fNt __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double2 ld2_nt(const double* p, size_t e) {
  const V2dNt v = __builtin_nontemporal_load(reinterpret_cast<const V2dNt*>(p) + e);
  return make_double2(v.x, v.y);
}

// Element e of a bulk slab in its own precision, non-temporal.
template <typename TB>
__device__ __forceinline__ typename C2<TB>::type ldv_nt(const TB* p, size_t e);

template <>
__device__ __forceinline__ double2 ldv_nt<double>(const double* p, size_t e) {
  return ld2_nt(p, e);
}

template <>
__device__ __forceinline__ float2 ldv_nt<float>(const float* p, size_t e) {
  const V2fNt v = __builtin_nontemporal_load(reinterpret_cast<const V2fNt*>(p) + e);
  return make_float2(v.x, v.y);
}

template <typename T>
__device__ __forceinline__ void st_nt(T* __restrict__ p, size_t e, double re, double im);

template <>
__device__ __forceinline__ void st_nt<double>(double* __restrict__ p, size_t e, double re,
                                              double im) {
  V2dNt v;
  v.x = re;
  v.y = im;
  __builtin_nontemporal_store(v, reinterpret_cast<V2dNt*>(p) + e);
}

template <>
__device__ __forceinline__ void st_nt<float>(float* __restrict__ p, size_t e, double re,
                                             double im) {
  V2fNt v;
  v.x = (float)re;
  v.y = (float)im;
  __builtin_nontemporal_store(v, reinterpret_cast<V2fNt*>(p) + e);
}

// A bulk element in its own precision, non-temporal.
template <typename TB>
__device__ __forceinline__ void stv_nt(TB* __restrict__ p, size_t e, typename C2<TB>::type v);

template <>
__device__ __forceinline__ void stv_nt<double>(double* __restrict__ p, size_t e, double2 v) {
  st_nt<double>(p, e, v.x, v.y);
}

template <>
__device__ __forceinline__ void stv_nt<float>(float* __restrict__ p, size_t e, float2 v) {
  V2fNt w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<V2fNt*>(p) + e);
}

// --- RK4 -------------------------------------------------------------------
// Two stage orders of the same step.  Classic (PHI's form,
// hierarchy_integrator.cc:2714-2825), buffers rho, acc (FP64), Yin/Yout (TB):
//   s0: acc = rho + dt/6 k, Yout = rho + dt/2 k      (input rho)
//   s1: acc += dt/3 k,      Yout = rho + dt/2 k
//   s2: acc += dt/3 k,      Yout = rho + dt k
//   s3: rho = acc + dt/6 k
// Nested (HORNER): for the time-independent linear RHS the RK4 step is the
// degree-4 Taylor polynomial of dt L, evaluated as
//   v1 = rho + dt/4 L rho, v2 = rho + dt/3 L v1, v3 = rho + dt/2 L v2,
//   rho = rho + dt L v3,
// which needs no accumulator: a stage reads rho and v and writes v (rho in
// the last one), N^2 (80 + 6 s) bytes per node-step instead of
// N^2 (160 + 6 s).  The two orders agree to rounding.
template <typename TB, typename TI, int NM, int STAGE, bool HORNER>
__global__ void __launch_bounds__(256) rk4_stage(RhsArgs A, const TI* __restrict__ in,
                                                 double* __restrict__ rho,
                                                 double* __restrict__ acc,
                                                 TB* __restrict__ yout, double dt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NE = A.NE;
  const int base = blockIdx.x * A.tile;
  const int nt = min(A.tile, A.num_nodes - base);
  if (nt <= 0) return;
  const Lds<TB> s = carve<TB>(smem, A.N, A.D, A.L, A.tile, NM != 1);
  const int tot = nt * NE;
  const size_t off = (size_t)base * NE;
  load_tile<TB, TI>(A, in, base, nt, s, NM != 1);
  // Epilogue operands stream in chunks of kPre elements per thread, the loads
  // of chunk c+1 in flight while chunk c is combined; chunk 0 is issued
  // with the tile loads.
  constexpr bool kRho = HORNER || STAGE <= 2;
  constexpr bool kAcc = !HORNER && STAGE >= 1;
  const int chunk = kPre * blockDim.x;
  double2 pr[kPre], pa[kPre];
  auto fetch = [&](int e0, double2* r, double2* q) {
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int e = e0 + threadIdx.x + i * blockDim.x;
      if (e < tot) {
        if (kRho) r[i] = ld2(rho, off + e);
        if (kAcc) q[i] = ld2(acc, off + e);
      }
    }
  };
  fetch(0, pr, pa);
  __syncthreads();
  if constexpr (NM == 1) {
    rhs_phase_mfma<TB, TI>(A, in, nt, s);
  } else {
    rhs_phase<TB, TI>(A, in, nt, s);
  }
  __syncthreads();
  for (int e0 = 0; e0 < tot; e0 += chunk) {
    double2 nr[kPre], na[kPre];
    if (e0 + chunk < tot) fetch(e0 + chunk, nr, na);
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int e = e0 + threadIdx.x + i * blockDim.x;
      if (e >= tot) break;
      const double2 k = kval<TB>(A, s, e);
      const size_t g = off + e;
      const double2 p = pr[i], q = pa[i];
      if (HORNER) {
        if (STAGE < 3) {
          const double c = dt / (4 - STAGE);
          st<TB>(yout, g, p.x + c * k.x, p.y + c * k.y);
        } else {
          st<double>(rho, g, p.x + dt * k.x, p.y + dt * k.y);
        }
      } else if (STAGE == 0) {
        st<double>(acc, g, p.x + (dt / 6) * k.x, p.y + (dt / 6) * k.y);
        st<TB>(yout, g, p.x + (dt / 2) * k.x, p.y + (dt / 2) * k.y);
      } else if (STAGE == 1 || STAGE == 2) {
        st<double>(acc, g, q.x + (dt / 3) * k.x, q.y + (dt / 3) * k.y);
        const double c = STAGE == 1 ? dt / 2 : dt;
        st<TB>(yout, g, p.x + c * k.x, p.y + c * k.y);
      } else {
        st<double>(rho, g, q.x + (dt / 6) * k.x, q.y + (dt / 6) * k.y);
      }
    }
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      if (kRho) pr[i] = nr[i];
      if (kAcc) pa[i] = na[i];
    }
  }
}

// --- RKF45 -------------------------------------------------------------------
// Stage s (0..4) computes k_s = RHS(stage input) and writes the next stage
// input  tmp = rho + sum_{i<=s} b_{s+1,i} k_i  (TB), left to right like PHI
// (hierarchy_integrator.cc:2019-2253).  Stage 5 forms
//   y4 = rho + c50 k0 + c52 k2 + c53 k3 + c54 k4
//   y5 = rho + c60 k0 + c62 k2 + c63 k3 + c64 k4 + c65 k5
// writes y5 (FP64) to rho_new and folds max |y4 - y5| into *err
// (hierarchy_integrator.cc:2317-2328).  k's are stored in TB.
__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* p, double v) {
  // Non-negative doubles order like their bit patterns.
  atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

template <typename TB, typename TI, int NM, int STAGE>
__global__ void __launch_bounds__(256) rkf45_stage(RhsArgs A, const TI* __restrict__ in,
                                                   const double* __restrict__ rho,
                                                   Rkf45Bufs<TB> B, Rkf45Coefs cf) {
  typedef typename C2<TB>::type V;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[256 / 64];
  const int NE = A.NE;
  const int base = blockIdx.x * A.tile;
  const int nt = min(A.tile, A.num_nodes - base);
  if (nt <= 0) return;
  const Lds<TB> s = carve<TB>(smem, A.N, A.D, A.L, A.tile, NM != 1);
  const int tot = nt * NE;
  const size_t off = (size_t)base * NE;
  load_tile<TB, TI>(A, in, base, nt, s, NM != 1);
  double2 pr[kPre];
  auto prefetch = [&]() {
#pragma unroll
    for (int i = 0; i < kPre; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      if (e < tot) pr[i] = ld2(rho, off + e);
    }
  };
  prefetch();
  __syncthreads();
  if constexpr (NM == 1) {
    rhs_phase_mfma<TB, TI>(A, in, nt, s);
  } else {
    rhs_phase<TB, TI>(A, in, nt, s);
  }
  __syncthreads();
  double emax = 0;
  for (int i = 0, e = threadIdx.x; e < tot; ++i, e += blockDim.x) {
    const double2 k = kval<TB>(A, s, e);
    const size_t g = off + e;
    const double2 p = i < kPre ? pr[i] : ld2(rho, g);
    if (STAGE < 5) {
      st<TB>(B.k[STAGE], g, k.x, k.y);
      // The TB value of k_STAGE as stored, so later stages see the same k.
      const double ksx = (double)(TB)k.x, ksy = (double)(TB)k.y;
      double sx = p.x, sy = p.y;
      for (int j = 0; j < STAGE; ++j) {
        const V kj = ld<TB, TB>(B.k[j], g);
        sx += (double)kj.x * cf.b[STAGE][j];
        sy += (double)kj.y * cf.b[STAGE][j];
      }
      sx += ksx * cf.b[STAGE][STAGE];
      sy += ksy * cf.b[STAGE][STAGE];
      st<TB>(B.tmp_out, g, sx, sy);
    } else {
      const V k0 = ld<TB, TB>(B.k[0], g);
      const V k2 = ld<TB, TB>(B.k[2], g);
      const V k3 = ld<TB, TB>(B.k[3], g);
      const V k4 = ld<TB, TB>(B.k[4], g);
      const double k5x = (double)(TB)k.x, k5y = (double)(TB)k.y;
      const double y4x = p.x + (double)k0.x * cf.c4[0] + (double)k2.x * cf.c4[1] +
                         (double)k3.x * cf.c4[2] + (double)k4.x * cf.c4[3];
      const double y4y = p.y + (double)k0.y * cf.c4[0] + (double)k2.y * cf.c4[1] +
                         (double)k3.y * cf.c4[2] + (double)k4.y * cf.c4[3];
      const double y5x = p.x + (double)k0.x * cf.c5[0] + (double)k2.x * cf.c5[1] +
                         (double)k3.x * cf.c5[2] + (double)k4.x * cf.c5[3] + k5x * cf.c5[4];
      const double y5y = p.y + (double)k0.y * cf.c5[0] + (double)k2.y * cf.c5[1] +
                         (double)k3.y * cf.c5[2] + (double)k4.y * cf.c5[3] + k5y * cf.c5[4];
      st<double>(B.rho_new, g, y5x, y5y);
      const double d = hypot(y4x - y5x, y4y - y5y);
      emax = d > emax ? d : emax;
    }
  }
  if (STAGE == 5) {
    for (int o = 32; o > 0; o >>= 1) {
      const double v = __shfl_xor(emax, o, 64);
      emax = v > emax ? v : emax;
    }
    const int w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) red[w] = emax;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int i = 1; i < (int)(blockDim.x / 64); ++i) m = red[i] > m ? red[i] : m;
      atomic_max_nonneg(B.err, m);
    }
  }
}

}  // namespace dev
}  // namespace phi

#endif  // PHI_AMD_HEOM_DEVICE_H_
