// HIP kernels of the MI355X HEOM integrator (gfx950).
//
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
#include <hip/hip_runtime.h>

#include <cstdint>

#include "heom_kernels.h"

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
__host__ __device__ inline size_t LdsBytes(int N, int D, int L, int tile, size_t real) {
  const size_t NE = (size_t)N * N, V = 2 * real;
  size_t b = V * (NE + 2 * NE * tile + (size_t)D * L);
  b = Align16(b + real * ((size_t)D * L + N + tile));
  b = Align16(b + sizeof(int) * ((size_t)N + 1 + D + 2 * (size_t)tile * D));
  b = Align16(b + (size_t)tile * D);
  return b;
}

template <typename TB>
__device__ __forceinline__ Lds<TB> carve(unsigned char* smem, int N, int D, int L, int tile) {
  typedef typename C2<TB>::type V;
  const int NE = N * N;
  Lds<TB> s;
  s.H = reinterpret_cast<V*>(smem);
  s.Y = s.H + NE;
  s.W = s.Y + NE * tile;
  s.pcol = s.W + NE * tile;
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
                                          int nt, const Lds<TB>& s) {
  const int N = A.N, NE = A.NE, D = A.D, L = A.L;
  const int tid = threadIdx.x, nth = blockDim.x;
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
template <typename TB, typename TI>
__device__ __forceinline__ void rhs_phase(const RhsArgs& A, const TI* __restrict__ in, int nt,
                                          const Lds<TB>& s) {
  typedef typename C2<TB>::type V;
  const int N = A.N, NE = A.NE, D = A.D, L = A.L;
  const TB ih = (TB)A.inv_hbar;
  for (int e = threadIdx.x; e < nt * NE; e += blockDim.x) {
    const int t = e / NE;
    const int r = e - t * NE;
    const int b = r / N;
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
    s.W[e] = mk<TB>(ax + ih * xi - f * y.x, ay - ih * xr - f * y.y);
  }
}

// k(a, b) = Y'(a, b) + conj(Y'(b, a)), in double.
template <typename TB>
__device__ __forceinline__ double2 kval(const Lds<TB>& s, int e, int NE, int N) {
  typedef typename C2<TB>::type V;
  const int t = e / NE;
  const int r = e - t * NE;
  const int b = r / N;
  const int a = r - b * N;
  const V p = s.W[e];
  const V q = s.W[t * NE + a * N + b];
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

// --- RK4 -------------------------------------------------------------------
// Buffers: rho (FP64), acc (FP64), Yin/Yout (TB).  Stage 0 reads rho as the
// RHS input.  Update order:
//   s0: acc = rho + dt/6 k, Yout = rho + dt/2 k
//   s1: acc += dt/3 k,      Yout = rho + dt/2 k
//   s2: acc += dt/3 k,      Yout = rho + dt k
//   s3: rho = acc + dt/6 k
// which is PHI's rho += dt/6 (k0 + 2k1 + 2k2 + k3) (hierarchy_integrator.cc:
// 2714-2825) up to rounding.
template <typename TB, typename TI, int STAGE>
__global__ void __launch_bounds__(256) rk4_stage(RhsArgs A, const TI* __restrict__ in,
                                                 double* __restrict__ rho,
                                                 double* __restrict__ acc,
                                                 TB* __restrict__ yout, double dt) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NE = A.NE;
  const int base = blockIdx.x * A.tile;
  const int nt = min(A.tile, A.num_nodes - base);
  if (nt <= 0) return;
  const Lds<TB> s = carve<TB>(smem, A.N, A.D, A.L, A.tile);
  const int tot = nt * NE;
  const size_t off = (size_t)base * NE;
  // Epilogue operands, issued with the tile loads.
  double2 pr[kPre], pa[kPre];
#pragma unroll
  for (int i = 0; i < kPre; ++i) {
    const int e = threadIdx.x + i * blockDim.x;
    if (e < tot) {
      if (STAGE <= 2) pr[i] = ld2(rho, off + e);
      if (STAGE >= 1) pa[i] = ld2(acc, off + e);
    }
  }
  load_tile<TB, TI>(A, in, base, nt, s);
  __syncthreads();
  rhs_phase<TB, TI>(A, in, nt, s);
  __syncthreads();
  for (int i = 0, e = threadIdx.x; e < tot; ++i, e += blockDim.x) {
    const double2 k = kval<TB>(s, e, NE, A.N);
    const size_t g = off + e;
    double2 p, q;
    if (i < kPre) {
      p = pr[i];
      q = pa[i];
    } else {
      if (STAGE <= 2) p = ld2(rho, g);
      if (STAGE >= 1) q = ld2(acc, g);
    }
    if (STAGE == 0) {
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

template <typename TB, typename TI, int STAGE>
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
  const Lds<TB> s = carve<TB>(smem, A.N, A.D, A.L, A.tile);
  const int tot = nt * NE;
  const size_t off = (size_t)base * NE;
  double2 pr[kPre];
#pragma unroll
  for (int i = 0; i < kPre; ++i) {
    const int e = threadIdx.x + i * blockDim.x;
    if (e < tot) pr[i] = ld2(rho, off + e);
  }
  load_tile<TB, TI>(A, in, base, nt, s);
  __syncthreads();
  rhs_phase<TB, TI>(A, in, nt, s);
  __syncthreads();
  double emax = 0;
  for (int i = 0, e = threadIdx.x; e < tot; ++i, e += blockDim.x) {
    const double2 k = kval<TB>(s, e, NE, A.N);
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

// --- General (non-Hermitian) RHS: PHI's Get on any complex input ------------
// One thread per output element; every term is evaluated elementwise as in
// hierarchy_updater.h:64-75 and hierarchy_updater.cc:129-144, 216-254:
//   drho = -(i/hbar)(H x - x H^dag) - mu x - (Delta_a + Delta_b)(a != b) x
//          + sum_next p (x' Q_m - Q_m x') - sum_prev (col x'' Q_m + row Q_m x'')
// With trace_row set, element (0,0) of node 0 is replaced by trace(x_0)
// (the steady-state operator of this integrator).
__global__ void __launch_bounds__(256) rhs_general(RhsArgs A, const double* __restrict__ x,
                                                   double* __restrict__ y, int trace_row) {
  const size_t tot = (size_t)A.num_nodes * A.NE;
  const int N = A.N, NE = A.NE, D = A.D, L = A.L;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < tot;
       g += (size_t)gridDim.x * blockDim.x) {
    const int n = (int)(g / NE);
    const int r = (int)(g - (size_t)n * NE);
    const int b = r / N;
    const int a = r - b * N;
    const double2* xn = reinterpret_cast<const double2*>(x) + (size_t)n * NE;
    // The unshifted H: the mean-diagonal shift only cancels in the
    // Hermitian form H x - (H x)^dagger, not in H x - x H^dagger for general x.
    const double2* H = A.Hraw;
    const double2* Hd = A.Hadj;
    double hr = 0, hi = 0;
    for (int c = 0; c < N; ++c) {
      const double2 h = H[a + c * N], v = xn[b * N + c];
      hr += h.x * v.x - h.y * v.y;
      hi += h.x * v.y + h.y * v.x;
      const double2 w = xn[a + c * N], hd = Hd[c + b * N];
      hr -= w.x * hd.x - w.y * hd.y;
      hi -= w.x * hd.y + w.y * hd.x;
    }
    // Unitary part with the unshifted H: -(i/hbar)(Hx - xH^dag).
    double ox = A.inv_hbar * hi, oy = -A.inv_hbar * hr;
    if (!A.vacuum) {
      const double2 v = xn[r];
      double f = A.mu[n];
      if (a != b) f += A.delta[a] + A.delta[b];
      ox -= f * v.x;
      oy -= f * v.y;
      // Column b gathers (site b), row a gathers (site a).
      for (int side = 0; side < 2; ++side) {
        const int site = side == 0 ? b : a;
        for (int q = A.site_ptr[site]; q < A.site_ptr[site + 1]; ++q) {
          const int j = A.site_modes[q];
          const int k = A.occ[(size_t)n * D + j];
          const int nx = A.next[(size_t)n * D + j];
          if (nx >= 0) {
            const double s = A.next_s[j * L + k];
            const double2 u = reinterpret_cast<const double2*>(x)[(size_t)nx * NE + r];
            // column: + i s u ; row: - i s u
            const double sg = side == 0 ? s : -s;
            ox -= sg * u.y;
            oy += sg * u.x;
          }
          const int pv = A.prev[(size_t)n * D + j];
          if (pv >= 0) {
            const double2 c = side == 0 ? A.prev_col[j * L + k] : A.prev_row[j * L + k];
            const double2 u = reinterpret_cast<const double2*>(x)[(size_t)pv * NE + r];
            ox -= c.x * u.x - c.y * u.y;
            oy -= c.x * u.y + c.y * u.x;
          }
        }
      }
    }
    if (trace_row && g == 0) {
      ox = 0;
      oy = 0;
      for (int c = 0; c < N; ++c) {
        ox += xn[c * (N + 1)].x;
        oy += xn[c * (N + 1)].y;
      }
    }
    reinterpret_cast<double2*>(y)[g] = make_double2(ox, oy);
  }
}

// --- Small utility kernels ----------------------------------------------------

// dst[slot] = node 0 of src (N*N complex), slot read from *counter, then
// ++*counter.  One block.
__global__ void snapshot_node0(const double* __restrict__ src, double* __restrict__ hist,
                               int* __restrict__ counter, int ne, int capacity) {
  const int slot = *counter;
  if (slot < capacity) {
    for (int e = threadIdx.x; e < ne; e += blockDim.x)
      reinterpret_cast<double2*>(hist)[(size_t)slot * ne + e] =
          reinterpret_cast<const double2*>(src)[e];
  }
  __syncthreads();
  if (threadIdx.x == 0) *counter = slot + 1;
}

__global__ void fill_u64(unsigned long long* p, unsigned long long v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename TB>
__global__ void cast_to(const double* __restrict__ src, TB* __restrict__ dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (TB)src[i];
}

// Deterministic FP64 complex reductions for the steady-state solver:
// partial[block] = sum over the block's grid-stride range, combined in a
// fixed order by a single block (two passes, no atomics).
// mode 0: sum conj(u) v  (BLAS zdotc); mode 1: sum |u|^2 (v unused).
__global__ void __launch_bounds__(256) dot_partial(const double* __restrict__ u,
                                                   const double* __restrict__ v, size_t n,
                                                   int mode, double* __restrict__ partial) {
  __shared__ double sr[256], si[256];
  double ar = 0, ai = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 p = reinterpret_cast<const double2*>(u)[i];
    if (mode == 0) {
      const double2 q = reinterpret_cast<const double2*>(v)[i];
      ar += p.x * q.x + p.y * q.y;
      ai += p.x * q.y - p.y * q.x;
    } else {
      ar += p.x * p.x + p.y * p.y;
    }
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr[0];
    partial[2 * blockIdx.x + 1] = si[0];
  }
}

__global__ void __launch_bounds__(256) dot_final(const double* __restrict__ partial, int nb,
                                                 double* __restrict__ out) {
  __shared__ double sr[256], si[256];
  double ar = 0, ai = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) {
    ar += partial[2 * i];
    ai += partial[2 * i + 1];
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sr[threadIdx.x] += sr[threadIdx.x + s];
      si[threadIdx.x] += si[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sr[0];
    out[1] = si[0];
  }
}

// y = sum_t c_t x_t over up to 4 complex vectors (c on the host), in place
// allowed for y == x_0.
__global__ void __launch_bounds__(256) zlincomb(double* __restrict__ y, size_t n, int nv,
                                                const double* x0, const double* x1,
                                                const double* x2, const double* x3, double2 c0,
                                                double2 c1, double2 c2, double2 c3) {
  const double* xs[4] = {x0, x1, x2, x3};
  const double2 cs[4] = {c0, c1, c2, c3};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double ar = 0, ai = 0;
    for (int t = 0; t < nv; ++t) {
      const double2 v = reinterpret_cast<const double2*>(xs[t])[i];
      ar += cs[t].x * v.x - cs[t].y * v.y;
      ai += cs[t].x * v.y + cs[t].y * v.x;
    }
    reinterpret_cast<double2*>(y)[i] = make_double2(ar, ai);
  }
}

}  // namespace dev

// --- Host-side launchers -------------------------------------------------------

size_t RhsSharedBytes(const RhsArgs& a, bool fp32) {
  return dev::LdsBytes(a.N, a.D, a.L, a.tile, fp32 ? sizeof(float) : sizeof(double));
}

static int Blocks(const RhsArgs& a) { return (a.num_nodes + a.tile - 1) / a.tile; }

template <typename TB, typename TI>
static hipError_t LaunchRk4T(const RhsArgs& a, int stage, const TI* in, double* rho, double* acc,
                             TB* yout, double dt, hipStream_t s) {
  const size_t sh = RhsSharedBytes(a, sizeof(TB) == 4);
  const dim3 grid(Blocks(a)), block(256);
  switch (stage) {
    case 0: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, 0>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    case 1: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, 1>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    case 2: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, 2>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
    default: hipLaunchKernelGGL((dev::rk4_stage<TB, TI, 3>), grid, block, sh, s, a, in, rho, acc, yout, dt); break;
  }
  return hipGetLastError();
}

hipError_t LaunchRk4Stage(const RhsArgs& a, bool fp32, int stage, const void* in, double* rho,
                          double* acc, void* yout, double dt, hipStream_t s) {
  if (stage == 0) {
    if (fp32) return LaunchRk4T<float, double>(a, 0, (const double*)in, rho, acc, (float*)yout, dt, s);
    return LaunchRk4T<double, double>(a, 0, (const double*)in, rho, acc, (double*)yout, dt, s);
  }
  if (fp32) return LaunchRk4T<float, float>(a, stage, (const float*)in, rho, acc, (float*)yout, dt, s);
  return LaunchRk4T<double, double>(a, stage, (const double*)in, rho, acc, (double*)yout, dt, s);
}

template <typename TB, typename TI>
static hipError_t LaunchRkf45T(const RhsArgs& a, int stage, const TI* in, const double* rho,
                               const Rkf45Bufs<TB>& b, const Rkf45Coefs& c, hipStream_t s) {
  const size_t sh = RhsSharedBytes(a, sizeof(TB) == 4);
  const dim3 grid(Blocks(a)), block(256);
  switch (stage) {
    case 0: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, 0>), grid, block, sh, s, a, in, rho, b, c); break;
    case 1: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, 1>), grid, block, sh, s, a, in, rho, b, c); break;
    case 2: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, 2>), grid, block, sh, s, a, in, rho, b, c); break;
    case 3: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, 3>), grid, block, sh, s, a, in, rho, b, c); break;
    case 4: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, 4>), grid, block, sh, s, a, in, rho, b, c); break;
    default: hipLaunchKernelGGL((dev::rkf45_stage<TB, TI, 5>), grid, block, sh, s, a, in, rho, b, c); break;
  }
  return hipGetLastError();
}

hipError_t LaunchRkf45Stage(const RhsArgs& a, bool fp32, int stage, const void* in,
                            const double* rho, const Rkf45Bufs<float>& bf,
                            const Rkf45Bufs<double>& bd, const Rkf45Coefs& c, hipStream_t s) {
  if (stage == 0) {
    if (fp32) return LaunchRkf45T<float, double>(a, 0, (const double*)in, rho, bf, c, s);
    return LaunchRkf45T<double, double>(a, 0, (const double*)in, rho, bd, c, s);
  }
  if (fp32) return LaunchRkf45T<float, float>(a, stage, (const float*)in, rho, bf, c, s);
  return LaunchRkf45T<double, double>(a, stage, (const double*)in, rho, bd, c, s);
}

hipError_t LaunchRhsGeneral(const RhsArgs& a, const double* x, double* y, int trace_row,
                            hipStream_t s) {
  const size_t tot = (size_t)a.num_nodes * a.NE;
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(dev::rhs_general, dim3(blocks), dim3(256), 0, s, a, x, y, trace_row);
  return hipGetLastError();
}

hipError_t LaunchSnapshot(const double* src, double* hist, int* counter, int ne, int capacity,
                          hipStream_t s) {
  hipLaunchKernelGGL(dev::snapshot_node0, dim3(1), dim3(256), 0, s, src, hist, counter, ne, capacity);
  return hipGetLastError();
}

hipError_t LaunchFillU64(unsigned long long* p, unsigned long long v, int n, hipStream_t s) {
  hipLaunchKernelGGL(dev::fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t LaunchCastToFloat(const double* src, float* dst, size_t n, hipStream_t s) {
  hipLaunchKernelGGL(dev::cast_to<float>, dim3(1024), dim3(256), 0, s, src, dst, n);
  return hipGetLastError();
}

hipError_t LaunchDot(const double* u, const double* v, size_t n, int mode, double* partial,
                     int nblocks, double* out, hipStream_t s) {
  hipLaunchKernelGGL(dev::dot_partial, dim3(nblocks), dim3(256), 0, s, u, v, n, mode, partial);
  hipLaunchKernelGGL(dev::dot_final, dim3(1), dim3(256), 0, s, partial, nblocks, out);
  return hipGetLastError();
}

hipError_t LaunchLincomb(double* y, size_t n, int nv, const double* const* x, const double2* c,
                         hipStream_t s) {
  const double* xs[4] = {nullptr, nullptr, nullptr, nullptr};
  double2 cs[4] = {make_double2(0, 0), make_double2(0, 0), make_double2(0, 0), make_double2(0, 0)};
  for (int t = 0; t < nv; ++t) {
    xs[t] = x[t];
    cs[t] = c[t];
  }
  int blocks = (int)((n + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(dev::zlincomb, dim3(blocks), dim3(256), 0, s, y, n, nv, xs[0], xs[1], xs[2],
                     xs[3], cs[0], cs[1], cs[2], cs[3]);
  return hipGetLastError();
}

}  // namespace phi
