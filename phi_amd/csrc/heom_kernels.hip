// HIP kernels of the MI355X HEOM integrator (gfx950): the general RHS, the
// small utility kernels, the reductions of the steady-state solver and the
// stage-kernel dispatch.  The stage kernels themselves are templates in
// heom_device.h, instantiated per method and precision in heom_rk4_f64.hip,
// heom_rk4_f32.hip, heom_rkf45_f64.hip and heom_rkf45_f32.hip.
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "heom_stream.h"

namespace phi {
namespace dev {

// --- General (non-Hermitian) RHS: PHI's Get on any complex input ------------
// One thread per output element; every term is evaluated elementwise as in
// hierarchy_updater.h:64-75 and hierarchy_updater.cc:129-144, 216-254:
//   drho = -(i/hbar)(H x - x H^dag) - mu x - (Delta_a + Delta_b)(a != b) x
//          + sum_next p (x' Q_m - Q_m x') - sum_prev (col x'' Q_m + row Q_m x'')
// With trace_row set, element (0,0) of node 0 is replaced by trace(x_0)
// (the steady-state operator of this integrator).  With act non-null (the
// adaptive truncation's node flags) the terms of inactive neighbours are
// skipped and every node keeps its own terms (PHI's base Get with the
// adaptive DoNext / DoPrevious, hierarchy_updater.cc:256-304).
// One element (a, b) of node n; xn points at the node's ADO, H / Hd at H
// and H^dagger (global memory or LDS copies: same arithmetic either way).
template <typename XP, typename HP>
__device__ __forceinline__ double2 general_elem(const RhsArgs& A, const double* __restrict__ x,
                                                XP xn, HP H, HP Hd, int n, int r, int a, int b,
                                                const unsigned char* __restrict__ act,
                                                bool trace) {
  const int N = A.N, NE = A.NE, D = A.D, L = A.L;
  // The unshifted H: the mean-diagonal shift only cancels in the
  // Hermitian form H x - (H x)^dagger, not in H x - x H^dagger for general x.
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
    // Column b gathers (site b), row a gathers (site a): the modes of the
    // site in site_modes order, from the edge records {next id, prev id,
    // j L + n_j} when the engine keeps them (one load instead of the chain
    // site_ptr -> site_modes -> occ / next / prev), else from the tables.
    for (int side = 0; side < 2; ++side) {
      const int site = side == 0 ? b : a;
      const int q0 = A.edge ? 0 : A.site_ptr[site];
      const int q1 = A.edge ? A.mps : A.site_ptr[site + 1];
      for (int q = q0; q < q1; ++q) {
        int nx, pv, z;
        if (A.edge) {
          const int4 c = A.edge[((size_t)n * N + site) * A.mps + q];
          nx = c.x;
          pv = c.y;
          z = c.z;
        } else {
          const int j = A.site_modes[q];
          const size_t o = (size_t)n * D + j;
          nx = A.next[o];
          pv = A.prev[o];
          z = j * L + A.occ[o];
        }
        if (nx >= 0 && (!act || act[nx])) {
          const double s = A.next_s[z];
          const double2 u = reinterpret_cast<const double2*>(x)[(size_t)nx * NE + r];
          // column: + i s u ; row: - i s u
          const double sg = side == 0 ? s : -s;
          ox -= sg * u.y;
          oy += sg * u.x;
        }
        if (pv >= 0 && (!act || act[pv])) {
          const double2 c = side == 0 ? A.prev_col[z] : A.prev_row[z];
          const double2 u = reinterpret_cast<const double2*>(x)[(size_t)pv * NE + r];
          ox -= c.x * u.x - c.y * u.y;
          oy -= c.x * u.y + c.y * u.x;
        }
      }
    }
  }
  if (trace) {
    ox = 0;
    oy = 0;
    for (int c = 0; c < N; ++c) {
      ox += xn[c * (N + 1)].x;
      oy += xn[c * (N + 1)].y;
    }
  }
  return make_double2(ox, oy);
}

__global__ void __launch_bounds__(256) rhs_general(RhsArgs A, const double* __restrict__ x,
                                                   double* __restrict__ y, int trace_row,
                                                   const unsigned char* __restrict__ act) {
  const size_t tot = (size_t)A.num_nodes * A.NE;
  const int N = A.N, NE = A.NE;
  for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < tot;
       g += (size_t)gridDim.x * blockDim.x) {
    const int n = (int)(g / NE);
    const int r = (int)(g - (size_t)n * NE);
    const int b = r / N;
    const int a = r - b * N;
    const double2* xn = reinterpret_cast<const double2*>(x) + (size_t)n * NE;
    reinterpret_cast<double2*>(y)[g] =
        general_elem(A, x, xn, A.Hraw, A.Hadj, n, r, a, b, act, trace_row && g == 0);
  }
}

// partial[NC * block + c] = sum over the (256-thread) block of a[c], in a
// fixed order (shuffles within each wavefront, then the four wavefront sums
// in order).  Called by every thread of the block.
template <int NC>
__device__ __forceinline__ void block_sum(const double* a, double* __restrict__ partial) {
  __shared__ double sw[4][NC];
  double v[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) v[c] = a[c];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int c = 0; c < NC; ++c) v[c] += __shfl_xor(v[c], o, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) sw[threadIdx.x >> 6][c] = v[c];
  __syncthreads();
  if ((int)threadIdx.x < NC) {
    const int c = threadIdx.x;
    partial[NC * blockIdx.x + c] = ((sw[0][c] + sw[1][c]) + sw[2][c]) + sw[3][c];
  }
}

// The same with H, H^dagger and the ADOs of T nodes per block in LDS (one
// load of each per element instead of 2N): for N*N*(T+2) complex within
// the LDS budget.  DOT fuses the solver's next reduction into the matvec:
// 1 = the block's sum of conj(w) y (2 doubles per block), 3 = the block's
// sums of conj(y) w, |y|^2, conj(w2) y and conj(w2) w (7 doubles per block).
template <int DOT>
__global__ void __launch_bounds__(256) rhs_general_lds(RhsArgs A, const double* __restrict__ x,
                                                       double* __restrict__ y, int trace_row,
                                                       const unsigned char* __restrict__ act,
                                                       int T, const double* __restrict__ w,
                                                       const double* __restrict__ w2,
                                                       double* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, NE = A.NE;
  double2* sH = reinterpret_cast<double2*>(smem);
  double2* sHd = sH + NE;
  double2* sx = sHd + NE;
  for (int e = threadIdx.x; e < NE; e += blockDim.x) {
    sH[e] = A.Hraw[e];
    sHd[e] = A.Hadj[e];
  }
  double acc[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int n0 = blockIdx.x * T; n0 < A.num_nodes; n0 += gridDim.x * T) {
    const int nt = min(T, A.num_nodes - n0);
    __syncthreads();
    for (int e = threadIdx.x; e < nt * NE; e += blockDim.x)
      sx[e] = reinterpret_cast<const double2*>(x)[(size_t)n0 * NE + e];
    __syncthreads();
    for (int e = threadIdx.x; e < nt * NE; e += blockDim.x) {
      const int te = e / NE;
      const int r = e - te * NE;
      const int b = r / N;
      const int a = r - b * N;
      const int n = n0 + te;
      const size_t g = (size_t)n * NE + r;
      // the dot operands load while the element is evaluated
      double2 q = make_double2(0., 0.), h = make_double2(0., 0.);
      if constexpr (DOT != 0) q = reinterpret_cast<const double2*>(w)[g];
      if constexpr (DOT == 3) h = reinterpret_cast<const double2*>(w2)[g];
      const double2 v = general_elem(A, x, sx + te * NE, sH, sHd, n, r, a, b, act,
                                     trace_row && g == 0);
      reinterpret_cast<double2*>(y)[g] = v;
      if constexpr (DOT == 1) {
        acc[0] += q.x * v.x + q.y * v.y;
        acc[1] += q.x * v.y - q.y * v.x;
      } else if constexpr (DOT == 3) {
        acc[0] += v.x * q.x + v.y * q.y;
        acc[1] += v.x * q.y - v.y * q.x;
        acc[2] += v.x * v.x + v.y * v.y;
        acc[3] += h.x * v.x + h.y * v.y;
        acc[4] += h.x * v.y - h.y * v.x;
        acc[5] += h.x * q.x + h.y * q.y;
        acc[6] += h.x * q.y - h.y * q.x;
      }
    }
  }
  if constexpr (DOT == 1) block_sum<2>(acc, partial);
  if constexpr (DOT == 3) block_sum<7>(acc, partial);
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

// As snapshot_node0 for an ensemble: the mean of every member's node 0,
// summed in member order (deterministic).  One block.
// hist[slot] = mean over the members of their node 0 (slot = *counter,
// left unchanged; bump_counter advances it).  One block per element: each
// thread sums members tid, tid + blockDim.x, ... in order, then a fixed-order
// tree over the block, so the sum is the same on every run.
__global__ void __launch_bounds__(256) snapshot_mean(const double* __restrict__ src,
                                                     const int* __restrict__ moff, int members,
                                                     double* __restrict__ hist,
                                                     const int* __restrict__ counter, int ne,
                                                     int capacity) {
  __shared__ double2 red[256];
  const int slot = *counter;
  const int e = blockIdx.x;
  const double2* s = reinterpret_cast<const double2*>(src);
  double x = 0, y = 0;
  for (int m = threadIdx.x; m < members; m += blockDim.x) {
    const double2 v = s[(size_t)moff[m] * ne + e];
    x += v.x;
    y += v.y;
  }
  red[threadIdx.x] = make_double2(x, y);
  __syncthreads();
  for (int w = blockDim.x >> 1; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[threadIdx.x].x += red[threadIdx.x + w].x;
      red[threadIdx.x].y += red[threadIdx.x + w].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && slot < capacity) {
    const double inv = 1.0 / members;
    reinterpret_cast<double2*>(hist)[(size_t)slot * ne + e] =
        make_double2(red[0].x * inv, red[0].y * inv);
  }
}

__global__ void bump_counter(int* counter) { *counter += 1; }

// dst[m] = node 0 of member m (members x ne complex), for one copy to the host.
__global__ void gather_node0(const double* __restrict__ src, const int* __restrict__ moff,
                             int members, int ne, double* __restrict__ dst) {
  const long long n = (long long)members * ne;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int m = (int)(i / ne), e = (int)(i - (long long)m * ne);
    reinterpret_cast<double2*>(dst)[i] =
        reinterpret_cast<const double2*>(src)[(size_t)moff[m] * ne + e];
  }
}

// One block per ADO (grid-stride over nodes): T = a R^dagger, then
// a <- R T, through LDS (three N x N complex matrices).
template <typename T>
__global__ void __launch_bounds__(256) rotate_nodes(T* __restrict__ buf,
                                                    const double2* __restrict__ R, int N,
                                                    int nodes) {
  extern __shared__ __align__(16) unsigned char smem[];
  double2* r = reinterpret_cast<double2*>(smem);
  double2* a = r + N * N;
  double2* t = a + N * N;
  const int NE = N * N;
  for (int e = threadIdx.x; e < NE; e += blockDim.x) r[e] = R[e];
  for (int n = blockIdx.x; n < nodes; n += gridDim.x) {
    T* m = buf + (size_t)n * NE * 2;
    __syncthreads();
    for (int e = threadIdx.x; e < NE; e += blockDim.x)
      a[e] = make_double2((double)m[2 * e], (double)m[2 * e + 1]);
    __syncthreads();
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
      const int i = e % N, j = e / N;  // element (i, j)
      double x = 0, y = 0;
      for (int c = 0; c < N; ++c) {
        const double2 p = a[i + c * N], q = r[j + c * N];  // a(i,c) conj(R(j,c))
        x += p.x * q.x + p.y * q.y;
        y += p.y * q.x - p.x * q.y;
      }
      t[e] = make_double2(x, y);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
      const int i = e % N, j = e / N;
      double x = 0, y = 0;
      for (int c = 0; c < N; ++c) {
        const double2 p = r[i + c * N], q = t[c + j * N];  // R(i,c) T(c,j)
        x += p.x * q.x - p.y * q.y;
        y += p.x * q.y + p.y * q.x;
      }
      m[2 * e] = (T)x;
      m[2 * e + 1] = (T)y;
    }
  }
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

// Fixed-order sum of nb complex partials by a block of 256 threads; the
// result is valid in thread 0.
__device__ __forceinline__ void dot_reduce(const double* partial, int nb, double* re,
                                           double* im) {
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
  *re = sr[0];
  *im = si[0];
}

__global__ void __launch_bounds__(256) dot_final(const double* __restrict__ partial, int nb,
                                                 double* __restrict__ out) {
  double re, im;
  dot_reduce(partial, nb, &re, &im);
  if (threadIdx.x == 0) {
    out[0] = re;
    out[1] = im;
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
    // Fully unrolled so that xs and cs stay in registers (a loop bounded
    // by nv indexes them at run time, which puts them in scratch memory).
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (t < nv) {
        const double2 v = reinterpret_cast<const double2*>(xs[t])[i];
        ar += cs[t].x * v.x - cs[t].y * v.y;
        ai += cs[t].x * v.y + cs[t].y * v.x;
      }
    }
    reinterpret_cast<double2*>(y)[i] = make_double2(ar, ai);
  }
}

// --- BiCGSTAB with device scalars ------------------------------------------------

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a / b = a conj(b) / |b|^2
__device__ __forceinline__ double2 zdiv(double2 a, double2 b) {
  const double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double2 bicg_omega(const BicgScalars* st) {
  return st->tt.x > 0 ? zdiv(st->ts, st->tt) : make_double2(0., 0.);
}

// Sum of component c of nb partials (stride NC), fixed order; every
// thread of the (256-thread) block gets the NC results.  Called by all.
// Thread t sums records t, t + 256, ... (four loads in flight at a time),
// each wavefront folds its 64 sums by shuffles, and the four wavefront sums
// are added in wavefront order: the same order in every block, so every
// block derives the same scalars from them.
template <int NC>
__device__ __forceinline__ void reduce_partials(const double* __restrict__ partial, int nb,
                                                double* out) {
  __shared__ double sw[4][NC];
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0;
  int i = threadIdx.x;
  for (; i + 768 < nb; i += 1024) {
    double v[4][NC];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NC; ++c) v[u][c] = partial[NC * (i + 256 * u) + c];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += v[u][c];
  }
  for (; i < nb; i += 256)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += partial[NC * i + c];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += __shfl_xor(acc[c], o, 64);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int c = 0; c < NC; ++c) sw[threadIdx.x >> 6][c] = acc[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c < NC; ++c) out[c] = ((sw[0][c] + sw[1][c]) + sw[2][c]) + sw[3][c];
}


// Closes an iteration (every thread of every block computes the same
// values; reduce_partials is called by all threads): from the record
// `prev` of the iteration and the partials of |r|^2 its x/r/p update wrote,
// the scalars that start the next iteration (alpha, omega, rho_old = rho,
// rho = the rho_next that update derived, the residual, the count, the stop
// flag).
__device__ __forceinline__ BicgScalars bicg_close(const BicgScalars* prev, const double* xr_part,
                                                  int nb, double tol, int max_iter) {
  BicgScalars c = *prev;
  if (c.done) return c;  // uniform
  double q[1];
  reduce_partials<1>(xr_part, nb, q);
  c.rn2 = make_double2(q[0], 0.);
  const double2 al = zdiv(c.rho_new, c.rv), om = bicg_omega(&c);
  c.alpha = al;
  c.omega = om;
  c.rho_old = c.rho_new;
  c.rho_new = c.rho_next;
  c.res = sqrt(c.rn2.x);
  c.it += 1;
  if (!(c.res > tol) || (om.x == 0 && om.y == 0) || c.it >= max_iter) c.done = 1;
  return c;
}

// s = r - alpha v for the iteration of record cur, after reducing the
// partials of rhat . v into rv (alpha = rho_new / rv; terms added left to
// right).  CLOSE: first close the previous iteration (record prev and its
// |r|^2 partials) into cur.  Block 0 writes cur (the closed scalars and rv);
// every block computes them itself, and nothing in this kernel reads cur.
template <bool CLOSE>
__global__ void __launch_bounds__(256) bicg_s(double* __restrict__ y, size_t n,
                                              const double* __restrict__ x0,
                                              const double* __restrict__ x1,
                                              const BicgScalars* prev, BicgScalars* cur,
                                              const double* rv_part, int nb,
                                              const double* xr_part, int nbx, double* hist,
                                              double tol, int max_iter) {
  double2 rho;
  int done;
  if (CLOSE) {
    const BicgScalars c = bicg_close(prev, xr_part, nbx, tol, max_iter);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      if (!prev->done) hist[c.it] = c.res;
      *cur = c;
    }
    rho = c.rho_new;
    done = c.done;
  } else {
    rho = cur->rho_new;
    done = cur->done;
  }
  if (done) return;  // uniform
  double pr[2];
  reduce_partials<2>(rv_part, nb, pr);
  const double2 rv = make_double2(pr[0], pr[1]);
  if (blockIdx.x == 0 && threadIdx.x == 0) cur->rv = rv;
  const double2 al = zdiv(rho, rv);
  const double2 c1 = make_double2(-al.x, -al.y);
  const double2* a = reinterpret_cast<const double2*>(x0);
  const double2* b = reinterpret_cast<const double2*>(x1);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  // two elements per trip (a thread past n reads element n - 1, unused)
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 2 * stride) {
    const size_t j = i + stride < n ? i + stride : n - 1;
    const double2 u0 = a[i], v0 = b[i], u1 = a[j], v1 = b[j];
    double ar = u0.x, ai = u0.y;
    ar += c1.x * v0.x - c1.y * v0.y;
    ai += c1.x * v0.y + c1.y * v0.x;
    reinterpret_cast<double2*>(y)[i] = make_double2(ar, ai);
    if (i + stride < n) {
      double br = u1.x, bi = u1.y;
      br += c1.x * v1.x - c1.y * v1.y;
      bi += c1.x * v1.y + c1.y * v1.x;
      reinterpret_cast<double2*>(y)[j] = make_double2(br, bi);
    }
  }
}

// The matvec's DOT 3 reductions on their own (when its LDS form does not
// fit): conj(t) s, |t|^2, conj(rhat) t, conj(rhat) s.
__global__ void __launch_bounds__(256) bicg_dots(const double* __restrict__ t,
                                                 const double* __restrict__ sv,
                                                 const double* __restrict__ rhat, size_t n,
                                                 double* __restrict__ partial) {
  double a[7] = {0, 0, 0, 0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 p = reinterpret_cast<const double2*>(t)[i];
    const double2 q = reinterpret_cast<const double2*>(sv)[i];
    const double2 h = reinterpret_cast<const double2*>(rhat)[i];
    a[0] += p.x * q.x + p.y * q.y;
    a[1] += p.x * q.y - p.y * q.x;
    a[2] += p.x * p.x + p.y * p.y;
    a[3] += h.x * p.x + h.y * p.y;
    a[4] += h.x * p.y - h.y * p.x;
    a[5] += h.x * q.x + h.y * q.y;
    a[6] += h.x * q.y - h.y * q.x;
  }
  block_sum<7>(a, partial);
}

// The x/r/p update of the iteration of record st, from the matvec's
// partials of conj(t) s, |t|^2, conj(rhat) t and conj(rhat) s: omega = ts /
// tt (0 if tt = 0), the next rho = conj(rhat) r_new = rhat.s - omega rhat.t
// (linearity: no pass over r_new needed before p), beta = (rho_next / rho)
// (alpha / omega); x += alpha p + omega s, r = s - omega t, p = r + beta
// (p - omega v), and the partials of |r|^2.  Block 0 keeps ts, tt and
// rho_next in st (other blocks read only rho_new, rv and done of it).
// Nothing once st->done.
__global__ void __launch_bounds__(256) bicg_xrp(double* __restrict__ x, double* __restrict__ r,
                                                double* __restrict__ p,
                                                const double* __restrict__ sv,
                                                const double* __restrict__ t,
                                                const double* __restrict__ v, size_t n,
                                                BicgScalars* st, const double* __restrict__ ts_part,
                                                int nb, double* __restrict__ partial) {
  if (st->done) return;  // uniform: every block returns
  double q[7];
  reduce_partials<7>(ts_part, nb, q);
  const double2 ts = make_double2(q[0], q[1]), tt = make_double2(q[2], 0.);
  const double2 rt = make_double2(q[3], q[4]), rs = make_double2(q[5], q[6]);
  const double2 al = zdiv(st->rho_new, st->rv);
  const double2 om = tt.x > 0 ? zdiv(ts, tt) : make_double2(0., 0.);
  const double2 orr = zmul(om, rt);
  const double2 rho_next = make_double2(rs.x - orr.x, rs.y - orr.y);
  const double2 beta = (om.x == 0 && om.y == 0)
                           ? make_double2(0., 0.)
                           : zmul(zdiv(rho_next, st->rho_new), zdiv(al, om));
  const double2 bo = zmul(beta, om);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->ts = ts;
    st->tt = tt;
    st->rho_next = rho_next;
  }
  double acc = 0;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  // two elements per trip (a thread past n reads element n - 1, unused)
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += 2 * stride) {
    const size_t j[2] = {i, i + stride < n ? i + stride : n - 1};
    double2 xv[2], pv[2], s[2], tv[2], vv[2];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      xv[e] = reinterpret_cast<const double2*>(x)[j[e]];
      pv[e] = reinterpret_cast<const double2*>(p)[j[e]];
      s[e] = reinterpret_cast<const double2*>(sv)[j[e]];
      tv[e] = reinterpret_cast<const double2*>(t)[j[e]];
      vv[e] = reinterpret_cast<const double2*>(v)[j[e]];
    }
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      if (e == 1 && !(i + stride < n)) break;
      // x = 1 x + alpha p + omega s ; r = 1 s - omega t (left to right)
      double xr = xv[e].x, xi = xv[e].y;
      xr += al.x * pv[e].x - al.y * pv[e].y;
      xi += al.x * pv[e].y + al.y * pv[e].x;
      xr += om.x * s[e].x - om.y * s[e].y;
      xi += om.x * s[e].y + om.y * s[e].x;
      reinterpret_cast<double2*>(x)[j[e]] = make_double2(xr, xi);
      double rr = s[e].x, ri = s[e].y;
      rr += -om.x * tv[e].x + om.y * tv[e].y;
      ri += -om.x * tv[e].y - om.y * tv[e].x;
      reinterpret_cast<double2*>(r)[j[e]] = make_double2(rr, ri);
      acc += rr * rr + ri * ri;
      // p = r + beta p - (beta omega) v, terms left to right
      double ar = rr, ai = ri;
      ar += beta.x * pv[e].x - beta.y * pv[e].y;
      ai += beta.x * pv[e].y + beta.y * pv[e].x;
      ar += -bo.x * vv[e].x + bo.y * vv[e].y;
      ai += -bo.x * vv[e].y - bo.y * vv[e].x;
      reinterpret_cast<double2*>(p)[j[e]] = make_double2(ar, ai);
    }
  }
  const double a[1] = {acc};
  block_sum<1>(a, partial);
}

// One block: closes the iteration of record prev into record next (the last
// iteration of a batch; within a batch the next s update closes).
__global__ void __launch_bounds__(256) bicg_finish(const BicgScalars* prev, BicgScalars* next,
                                                   const double* xr_part, int nb, double* hist,
                                                   double tol, int max_iter) {
  const BicgScalars c = bicg_close(prev, xr_part, nb, tol, max_iter);
  if (threadIdx.x == 0) {
    if (!prev->done) hist[c.it] = c.res;
    *next = c;
  }
}

}  // namespace dev

// --- Host-side launchers -------------------------------------------------------

int ResidentBlocks(const void* kernel, int threads, size_t lds) {
  struct Key {
    const void* k;
    int threads;
    size_t lds;
    int dev;
    bool operator<(const Key& o) const {
      return std::tie(k, threads, lds, dev) < std::tie(o.k, o.threads, o.lds, o.dev);
    }
  };
  static std::mutex mu;
  static std::map<Key, int> cache;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const Key key{kernel, threads, lds, dev};
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 1;
  const int blocks = per_cu * cus;
  cache[key] = blocks;
  return blocks;
}

size_t TileLdsBytes(int N, int D, int L, int tile, bool fp32, bool with_h) {
  return dev::LdsBytes(N, D, L, tile, fp32 ? sizeof(float) : sizeof(double), with_h);
}

size_t RhsSharedBytes(const RhsArgs& a, bool fp32) {
  return dev::LdsBytes(a.N, a.D, a.L, a.tile, fp32 ? sizeof(float) : sizeof(double),
                       a.col_nm == 0);
}

size_t StreamTileLdsBytes(int N, int D, int L, int tile, bool fp32) {
  return dev::StreamLdsBytes(N, D, L, tile, fp32 ? sizeof(float) : sizeof(double));
}

hipError_t LaunchRk4F64(const RhsArgs& a, int stage, int form, const void* in, double* rho,
                        void* aux, void* yout, double dt, hipStream_t s);
hipError_t LaunchRk4F32(const RhsArgs& a, int stage, int form, const void* in, double* rho,
                        void* aux, void* yout, double dt, hipStream_t s);
hipError_t LaunchRkf45F64(const RhsArgs& a, int stage, const void* in, const double* rho,
                         const Rkf45Bufs<double>& b, const Rkf45Coefs& c, hipStream_t s);
hipError_t LaunchRkf45F32(const RhsArgs& a, int stage, const void* in, const double* rho,
                         const Rkf45Bufs<float>& b, const Rkf45Coefs& c, hipStream_t s);

hipError_t LaunchRk4Stage(const RhsArgs& a, bool fp32, int stage, int form, const void* in,
                          double* rho, void* aux, void* yout, double dt, hipStream_t s) {
  return fp32 ? LaunchRk4F32(a, stage, form, in, rho, aux, yout, dt, s)
              : LaunchRk4F64(a, stage, form, in, rho, aux, yout, dt, s);
}

hipError_t LaunchRkf45Stage(const RhsArgs& a, bool fp32, int stage, const void* in,
                            const double* rho, const Rkf45Bufs<float>& bf,
                            const Rkf45Bufs<double>& bd, const Rkf45Coefs& c, hipStream_t s) {
  return fp32 ? LaunchRkf45F32(a, stage, in, rho, bf, c, s)
              : LaunchRkf45F64(a, stage, in, rho, bd, c, s);
}

// Nodes per block and blocks of rhs_general_lds; blocks = 0 when H,
// H^dagger and one node do not fit in 64 KB of LDS.
static void RhsLdsShape(const RhsArgs& a, int* T_out, int* blocks_out) {
  int T = a.NE > 0 ? 256 / a.NE : 0;
  if (T < 1) T = 1;
  while (T > 1 && (size_t)(T + 2) * a.NE * sizeof(double2) > 64 * 1024) --T;
  long long blocks = 0;
  if ((size_t)(T + 2) * a.NE * sizeof(double2) <= 64 * 1024) {
    blocks = ((long long)a.num_nodes + T - 1) / T;
    if (blocks > 16384) blocks = 16384;
    if (blocks < 1) blocks = 1;
  }
  *T_out = T;
  *blocks_out = (int)blocks;
}

// The fused form keeps the plain form's blocks: capping them at 512 (the
// consumers' partial count otherwise) cost the 7-site matvec 9.6 -> 16.7 us.
int RhsGeneralDotBlocks(const RhsArgs& a) {
  int T, blocks;
  RhsLdsShape(a, &T, &blocks);
  return blocks;
}

hipError_t LaunchRhsGeneralDot(const RhsArgs& a, const double* x, double* y, int trace_row,
                               int dot, const double* w, const double* w2, double* partial,
                               hipStream_t s) {
  int T, blocks;
  RhsLdsShape(a, &T, &blocks);
  if (blocks == 0 || (dot != 1 && dot != 3)) return hipErrorInvalidValue;
  const size_t sh = (size_t)(T + 2) * a.NE * sizeof(double2);
  if (dot == 1)
    hipLaunchKernelGGL(dev::rhs_general_lds<1>, dim3((unsigned)blocks), dim3(256), sh, s, a, x,
                       y, trace_row, (const unsigned char*)nullptr, T, w, w2, partial);
  else
    hipLaunchKernelGGL(dev::rhs_general_lds<3>, dim3((unsigned)blocks), dim3(256), sh, s, a, x,
                       y, trace_row, (const unsigned char*)nullptr, T, w, w2, partial);
  return hipGetLastError();
}

hipError_t LaunchRhsGeneral(const RhsArgs& a, const double* x, double* y, int trace_row,
                            hipStream_t s, const unsigned char* active) {
  // LDS form when H, H^dagger and at least one node fit in 64 KB.
  int T, blocks;
  RhsLdsShape(a, &T, &blocks);
  if (blocks > 0) {
    const size_t sh = (size_t)(T + 2) * a.NE * sizeof(double2);
    hipLaunchKernelGGL(dev::rhs_general_lds<0>, dim3((unsigned)blocks), dim3(256), sh, s, a, x,
                       y, trace_row, active, T, (const double*)nullptr, (const double*)nullptr,
                       (double*)nullptr);
    return hipGetLastError();
  }
  const size_t tot = (size_t)a.num_nodes * a.NE;
  int gb = (int)((tot + 255) / 256);
  if (gb > 65536) gb = 65536;
  hipLaunchKernelGGL(dev::rhs_general, dim3(gb), dim3(256), 0, s, a, x, y, trace_row, active);
  return hipGetLastError();
}

hipError_t LaunchSnapshot(const double* src, double* hist, int* counter, int ne, int capacity,
                          hipStream_t s) {
  hipLaunchKernelGGL(dev::snapshot_node0, dim3(1), dim3(256), 0, s, src, hist, counter, ne, capacity);
  return hipGetLastError();
}

hipError_t LaunchSnapshotMean(const double* src, const int* member_offsets, int members,
                              double* hist, int* counter, int ne, int capacity, hipStream_t s) {
  hipLaunchKernelGGL(dev::snapshot_mean, dim3(ne), dim3(256), 0, s, src, member_offsets, members,
                     hist, counter, ne, capacity);
  hipLaunchKernelGGL(dev::bump_counter, dim3(1), dim3(1), 0, s, counter);
  return hipGetLastError();
}

hipError_t LaunchGatherNode0(const double* src, const int* member_offsets, int members, int ne,
                             double* dst, hipStream_t s) {
  const long long n = (long long)members * ne;
  const int blocks = (int)std::min<long long>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(dev::gather_node0, dim3(blocks), dim3(256), 0, s, src, member_offsets,
                     members, ne, dst);
  return hipGetLastError();
}

hipError_t LaunchFillU64(unsigned long long* p, unsigned long long v, int n, hipStream_t s) {
  hipLaunchKernelGGL(dev::fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, v, n);
  return hipGetLastError();
}

hipError_t LaunchRotateNodes(void* buf, bool fp32, const double2* R, int N, int nodes,
                             hipStream_t s) {
  if (N < 1 || N > 32 || nodes < 1) return hipErrorInvalidValue;
  const size_t sh = 3 * (size_t)N * N * sizeof(double2);
  const int grid = nodes < 65535 ? nodes : 65535;
  if (fp32)
    hipLaunchKernelGGL(dev::rotate_nodes<float>, dim3(grid), dim3(256), sh, s, (float*)buf, R, N,
                       nodes);
  else
    hipLaunchKernelGGL(dev::rotate_nodes<double>, dim3(grid), dim3(256), sh, s, (double*)buf, R,
                       N, nodes);
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

hipError_t LaunchDotPartial(const double* u, const double* v, size_t n, double* partial,
                            int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(dev::dot_partial, dim3(nblocks), dim3(256), 0, s, u, v, n, 0, partial);
  return hipGetLastError();
}

// The BiCGSTAB update kernels run at most kBicgBlocks blocks (each block
// first reduces the partials of the dot product it needs; fewer blocks
// re-read them fewer times).
constexpr int kBicgBlocks = 512;

hipError_t LaunchBicgS(bool close, double* y, size_t n, const double* r, const double* v,
                       const BicgScalars* prev, BicgScalars* cur, const double* rv_part, int nb,
                       const double* xr_part, int nbx, double* hist, double tol, int max_iter,
                       hipStream_t s) {
  int blocks = (int)((n + 511) / 512);
  if (blocks > kBicgBlocks) blocks = kBicgBlocks;
  if (close)
    hipLaunchKernelGGL(dev::bicg_s<true>, dim3(blocks), dim3(256), 0, s, y, n, r, v, prev, cur,
                       rv_part, nb, xr_part, nbx, hist, tol, max_iter);
  else
    hipLaunchKernelGGL(dev::bicg_s<false>, dim3(blocks), dim3(256), 0, s, y, n, r, v, prev, cur,
                       rv_part, nb, xr_part, nbx, hist, tol, max_iter);
  return hipGetLastError();
}

hipError_t LaunchBicgDots(const double* t, const double* sv, const double* rhat, size_t n,
                          double* partial7, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(dev::bicg_dots, dim3(nblocks), dim3(256), 0, s, t, sv, rhat, n, partial7);
  return hipGetLastError();
}

hipError_t LaunchBicgXRP(double* x, double* r, double* p, const double* sv, const double* t,
                         const double* v, size_t n, BicgScalars* st, const double* ts_part,
                         int nb_ts, double* partial, int nblocks, hipStream_t s) {
  hipLaunchKernelGGL(dev::bicg_xrp, dim3(nblocks), dim3(256), 0, s, x, r, p, sv, t, v, n, st,
                     ts_part, nb_ts, partial);
  return hipGetLastError();
}

hipError_t LaunchBicgFinish(const BicgScalars* prev, BicgScalars* next, const double* xr_part,
                            int nblocks, double* hist, double tol, int max_iter, hipStream_t s) {
  hipLaunchKernelGGL(dev::bicg_finish, dim3(1), dim3(256), 0, s, prev, next, xr_part, nblocks,
                     hist, tol, max_iter);
  return hipGetLastError();
}

namespace dev {

__device__ __forceinline__ bool zzero(double2 a) { return a.x == 0 && a.y == 0; }

// rho0 = -omega rho0 (start of a cycle), then the first BiCG step's beta
// from rho1 = dot[3] = (rtilde, r_0).
__device__ void bicgl_start(BicglScalars* S) {
  const double2 m = zmul(S->omega, S->rho0);
  S->rho0 = make_double2(-m.x, -m.y);
  const double2 rho1 = S->dot[3];
  if (zzero(S->rho0)) {
    S->broke = 1;
    S->stop = 1;
    return;
  }
  S->beta = zdiv(zmul(S->alpha, rho1), S->rho0);
  S->rho0 = rho1;
}

__device__ void bicgl_solve(BicglScalars* S, int ell) {
  S->g[ell] = S->gp[ell];
  S->omega = S->g[ell];
  for (int a = ell - 1; a >= 1; --a) {
    double2 sum = make_double2(0., 0.);
    for (int b = a + 1; b <= ell; ++b) {
      const double2 t = zmul(S->tau[a][b], S->g[b]);
      sum.x += t.x;
      sum.y += t.y;
    }
    S->g[a] = make_double2(S->gp[a].x - sum.x, S->gp[a].y - sum.y);
  }
  for (int a = 1; a <= ell - 1; ++a) {
    double2 sum = make_double2(0., 0.);
    for (int b = a + 1; b <= ell - 1; ++b) {
      const double2 t = zmul(S->tau[a][b], S->g[b + 1]);
      sum.x += t.x;
      sum.y += t.y;
    }
    S->gpp[a] = make_double2(S->g[a + 1].x + sum.x, S->g[a + 1].y + sum.y);
  }
}

// One scalar step (one thread).
__device__ void bicgl_body(BicglScalars* S, int op, int i, int j, double tol, int max_iter,
                           double* hist, int hist_cap) {
  if (op == kBlInit) {
    S->res = sqrt(S->dot[2].x);
    S->rho0 = make_double2(1., 0.);
    S->alpha = make_double2(0., 0.);
    S->omega = make_double2(1., 0.);
    S->broke = 0;
    S->it = 0;
    S->cycles = 0;
    S->stop = !(S->res > tol) || max_iter <= 0;
    if (hist_cap > 0) hist[0] = S->res;
    if (!S->stop) bicgl_start(S);
    return;
  }
  if (S->stop) return;
  switch (op) {
    case kBlBeta: {
      const double2 rho1 = S->dot[0];
      if (zzero(S->rho0)) {
        S->broke = 1;
        S->stop = 1;
        break;
      }
      S->beta = zdiv(zmul(S->alpha, rho1), S->rho0);
      S->rho0 = rho1;
      break;
    }
    case kBlAlpha: {
      const double2 gam = S->dot[0];
      if (zzero(gam)) {
        S->broke = 1;
        S->stop = 1;
        break;
      }
      S->alpha = zdiv(S->rho0, gam);
      break;
    }
    case kBlTau: {
      const double2 d = S->dot[0];
      const double sg = S->sigma[i];
      S->tau[i][j] = make_double2(d.x / sg, d.y / sg);
      break;
    }
    case kBlSigma: {
      const double sg = S->dot[0].x;
      S->sigma[j] = sg;
      if (!(sg > 0)) {
        S->broke = 1;
        S->stop = 1;
        break;
      }
      const double2 d0 = S->dot[1];
      S->gp[j] = make_double2(d0.x / sg, d0.y / sg);
      if (j == i) bicgl_solve(S, i);
      break;
    }
    case kBlEnd: {
      S->res = sqrt(S->dot[2].x);
      S->it += i;
      S->cycles += 1;
      if (S->cycles < hist_cap) hist[S->cycles] = S->res;
      if (zzero(S->omega)) S->broke = 1;
      S->stop = S->broke || !(S->res > tol) || S->it >= max_iter;
      if (!S->stop) bicgl_start(S);
      break;
    }
    default:
      break;
  }
}

// Partials of two dot products in one pass: conj(a) . b and conj(c) . d,
// 4 doubles per block.
__global__ void __launch_bounds__(256) dot2_partial(const double* __restrict__ a,
                                                    const double* __restrict__ b,
                                                    const double* __restrict__ c,
                                                    const double* __restrict__ d, size_t n,
                                                    double* __restrict__ partial) {
  __shared__ double sh[4][256];
  double acc[4] = {0, 0, 0, 0};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 p = reinterpret_cast<const double2*>(a)[i];
    const double2 q = reinterpret_cast<const double2*>(b)[i];
    const double2 u = reinterpret_cast<const double2*>(c)[i];
    const double2 v = reinterpret_cast<const double2*>(d)[i];
    acc[0] += p.x * q.x + p.y * q.y;
    acc[1] += p.x * q.y - p.y * q.x;
    acc[2] += u.x * v.x + u.y * v.y;
    acc[3] += u.x * v.y - u.y * v.x;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
#pragma unroll
      for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) partial[4 * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// Fixed-order sum of nb records of dot2_partial (one block of 256 threads)
// into S->dot[k] and S->dot[k + 1], then the scalar step op (thread 0).
__global__ void __launch_bounds__(256) bicgl_dot2_step(const double* __restrict__ partial, int nb,
                                                       BicglScalars* S, int k, int op, int i,
                                                       int j, double tol, int max_iter,
                                                       double* hist, int hist_cap) {
  __shared__ double sh[4][256];
  double acc[4] = {0, 0, 0, 0};
  for (int r = threadIdx.x; r < nb; r += blockDim.x)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[c] += partial[4 * r + c];
#pragma unroll
  for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
#pragma unroll
      for (int c = 0; c < 4; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    S->dot[k] = make_double2(sh[0][0], sh[1][0]);
    S->dot[k + 1] = make_double2(sh[2][0], sh[3][0]);
    bicgl_body(S, op, i, j, tol, max_iter, hist, hist_cap);
  }
}

// dot_final into S->dot[k], then the scalar step op (thread 0).
__global__ void __launch_bounds__(256) bicgl_dot_step(const double* __restrict__ partial, int nb,
                                                      BicglScalars* S, int k, int op, int i,
                                                      int j, double tol, int max_iter,
                                                      double* hist, int hist_cap) {
  double re, im;
  dot_reduce(partial, nb, &re, &im);
  if (threadIdx.x == 0) {
    S->dot[k] = make_double2(re, im);
    bicgl_body(S, op, i, j, tol, max_iter, hist, hist_cap);
  }
}

__global__ void __launch_bounds__(256) zaxpy_dev(double* y, const double* a, const double* b,
                                                 size_t n, const double2* c, double sign,
                                                 const int* stop) {
  if (*stop) return;  // uniform: every block returns
  const double2 cc = make_double2(sign * c->x, sign * c->y);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 av = reinterpret_cast<const double2*>(a)[i];
    const double2 bv = reinterpret_cast<const double2*>(b)[i];
    reinterpret_cast<double2*>(y)[i] =
        make_double2(av.x + (cc.x * bv.x - cc.y * bv.y), av.y + (cc.x * bv.y + cc.y * bv.x));
  }
}

__global__ void __launch_bounds__(256) zaxpy_pairs(ZPairs pr, size_t n, const double2* c,
                                                   const int* stop) {
  if (*stop) return;  // uniform: every block returns
  const double2 cc = *c;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int p = 0; p < kBicglMaxEll + 1; ++p) {
      if (p < pr.count) {
        const double2 av = reinterpret_cast<const double2*>(pr.a[p])[i];
        const double2 bv = reinterpret_cast<const double2*>(pr.b[p])[i];
        const double2 k = make_double2(pr.sign[p] * cc.x, pr.sign[p] * cc.y);
        reinterpret_cast<double2*>(pr.y[p])[i] =
            make_double2(av.x + (k.x * bv.x - k.y * bv.y), av.y + (k.x * bv.y + k.y * bv.x));
      }
    }
  }
}

__global__ void __launch_bounds__(256) zcomb_dev(double* y, const double* a, ZTerms t, size_t n,
                                                 const int* stop) {
  if (*stop) return;  // uniform: every block returns
  double2 k[kBicglMaxEll];
#pragma unroll
  for (int q = 0; q < kBicglMaxEll; ++q)
    k[q] = q < t.count ? make_double2(t.sign[q] * t.c[q]->x, t.sign[q] * t.c[q]->y)
                       : make_double2(0., 0.);
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double2 acc = reinterpret_cast<const double2*>(a)[i];
#pragma unroll
    for (int q = 0; q < kBicglMaxEll; ++q) {
      if (q < t.count) {
        const double2 bv = reinterpret_cast<const double2*>(t.b[q])[i];
        acc.x += k[q].x * bv.x - k[q].y * bv.y;
        acc.y += k[q].x * bv.y + k[q].y * bv.x;
      }
    }
    reinterpret_cast<double2*>(y)[i] = acc;
  }
}

}  // namespace dev

static int ZBlocks(size_t n) {
  int blocks = (int)((n + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  return blocks < 1 ? 1 : blocks;
}

hipError_t LaunchZaxpyPairs(const ZPairs& pr, size_t n, const double2* c, const int* stop,
                            hipStream_t s) {
  if (pr.count < 1 || pr.count > kBicglMaxEll + 1) return hipErrorInvalidValue;
  // unused slots repeat slot 0 (never stored to), so every pointer is valid
  ZPairs q = pr;
  for (int p = q.count; p < kBicglMaxEll + 1; ++p) {
    q.y[p] = q.y[0];
    q.a[p] = q.a[0];
    q.b[p] = q.b[0];
    q.sign[p] = 0;
  }
  hipLaunchKernelGGL(dev::zaxpy_pairs, dim3(ZBlocks(n)), dim3(256), 0, s, q, n, c, stop);
  return hipGetLastError();
}

hipError_t LaunchZcombDev(double* y, const double* a, const ZTerms& t, size_t n, const int* stop,
                          hipStream_t s) {
  if (t.count < 1 || t.count > kBicglMaxEll) return hipErrorInvalidValue;
  // unused slots repeat slot 0 with weight 0, so every pointer is valid
  ZTerms q = t;
  for (int k = q.count; k < kBicglMaxEll; ++k) {
    q.b[k] = q.b[0];
    q.c[k] = q.c[0];
    q.sign[k] = 0;
  }
  hipLaunchKernelGGL(dev::zcomb_dev, dim3(ZBlocks(n)), dim3(256), 0, s, y, a, q, n, stop);
  return hipGetLastError();
}

hipError_t LaunchBicglDotStep(const double* u, const double* v, size_t n, double* partial,
                             int nblocks, BicglScalars* S, int k, int op, int i, int j, double tol,
                             int max_iter, double* hist, int hist_cap, hipStream_t s) {
  if (k < 0 || k > 3) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dev::dot_partial, dim3(nblocks), dim3(256), 0, s, u, v, n, 0, partial);
  hipLaunchKernelGGL(dev::bicgl_dot_step, dim3(1), dim3(256), 0, s, partial, nblocks, S, k, op,
                     i, j, tol, max_iter, hist, hist_cap);
  return hipGetLastError();
}

hipError_t LaunchBicglReduceStep(const double* partial, int nb, BicglScalars* S, int k, int op,
                                int i, int j, double tol, int max_iter, double* hist,
                                int hist_cap, hipStream_t s) {
  if (k < 0 || k > 3) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dev::bicgl_dot_step, dim3(1), dim3(256), 0, s, partial, nb, S, k, op, i, j,
                     tol, max_iter, hist, hist_cap);
  return hipGetLastError();
}

hipError_t LaunchBicglDot2Step(const double* a, const double* b, const double* c,
                              const double* d, size_t n, double* partial, int nblocks,
                              BicglScalars* S, int k, int op, int i, int j, double tol,
                              int max_iter, double* hist, int hist_cap, hipStream_t s) {
  if (k < 0 || k > 2) return hipErrorInvalidValue;
  hipLaunchKernelGGL(dev::dot2_partial, dim3(nblocks), dim3(256), 0, s, a, b, c, d, n, partial);
  hipLaunchKernelGGL(dev::bicgl_dot2_step, dim3(1), dim3(256), 0, s, partial, nblocks, S, k, op,
                     i, j, tol, max_iter, hist, hist_cap);
  return hipGetLastError();
}

hipError_t LaunchZaxpyDev(double* y, const double* a, const double* b, size_t n,
                          const double2* c, double sign, const int* stop, hipStream_t s) {
  int blocks = (int)((n + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(dev::zaxpy_dev, dim3(blocks), dim3(256), 0, s, y, a, b, n, c, sign, stop);
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
