// HIP kernels of the MI355X HEOM integrator (gfx950): the general RHS, the
// small utility kernels, the reductions of the steady-state solver and the
// stage-kernel dispatch.  The stage kernels themselves are templates in
// heom_device.h, instantiated per method and precision in heom_rk4_f64.hip,
// heom_rk4_f32.hip, heom_rkf45_f64.hip and heom_rkf45_f32.hip.
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
__global__ void __launch_bounds__(256) rhs_general(RhsArgs A, const double* __restrict__ x,
                                                   double* __restrict__ y, int trace_row,
                                                   const unsigned char* __restrict__ act) {
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
          if (nx >= 0 && (!act || act[nx])) {
            const double s = A.next_s[j * L + k];
            const double2 u = reinterpret_cast<const double2*>(x)[(size_t)nx * NE + r];
            // column: + i s u ; row: - i s u
            const double sg = side == 0 ? s : -s;
            ox -= sg * u.y;
            oy += sg * u.x;
          }
          const int pv = A.prev[(size_t)n * D + j];
          if (pv >= 0 && (!act || act[pv])) {
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

// As snapshot_node0 for an ensemble: the mean of every member's node 0,
// summed in member order (deterministic).  One block.
__global__ void snapshot_mean(const double* __restrict__ src, const int* __restrict__ moff,
                              int members, double* __restrict__ hist, int* __restrict__ counter,
                              int ne, int capacity) {
  const int slot = *counter;
  if (slot < capacity) {
    const double2* s = reinterpret_cast<const double2*>(src);
    const double w = 1.0 / members;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
      double x = 0, y = 0;
      for (int m = 0; m < members; ++m) {
        const double2 v = s[(size_t)moff[m] * ne + e];
        x += v.x;
        y += v.y;
      }
      reinterpret_cast<double2*>(hist)[(size_t)slot * ne + e] = make_double2(x * w, y * w);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) *counter = slot + 1;
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

__global__ void __launch_bounds__(256) bicg_update(int which, double* __restrict__ y, size_t n,
                                                   const double* x0, const double* x1,
                                                   const double* x2, const BicgScalars* st) {
  double2 c[3];
  int nv;
  if (which == 0) {
    const double2 beta = zmul(zdiv(st->rho_new, st->rho_old), zdiv(st->alpha, st->omega));
    const double2 bo = zmul(beta, st->omega);
    c[0] = make_double2(1., 0.);
    c[1] = beta;
    c[2] = make_double2(-bo.x, -bo.y);
    nv = 3;
  } else {
    const double2 al = zdiv(st->rho_new, st->rv);
    c[0] = make_double2(1., 0.);
    c[1] = make_double2(-al.x, -al.y);
    c[2] = make_double2(0., 0.);
    nv = 2;
  }
  const double* xs[3] = {x0, x1, x2};
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double ar = 0, ai = 0;
    for (int t = 0; t < nv; ++t) {
      const double2 v = reinterpret_cast<const double2*>(xs[t])[i];
      ar += c[t].x * v.x - c[t].y * v.y;
      ai += c[t].x * v.y + c[t].y * v.x;
    }
    reinterpret_cast<double2*>(y)[i] = make_double2(ar, ai);
  }
}

__global__ void __launch_bounds__(256) bicg_xr(double* __restrict__ x, double* __restrict__ r,
                                               const double* __restrict__ p,
                                               const double* __restrict__ sv,
                                               const double* __restrict__ t, size_t n,
                                               const BicgScalars* st,
                                               double* __restrict__ partial) {
  __shared__ double sr[256];
  if (st->done) return;  // uniform: every block returns
  const double2 al = zdiv(st->rho_new, st->rv), om = bicg_omega(st);
  double acc = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 xv = reinterpret_cast<const double2*>(x)[i];
    const double2 pv = reinterpret_cast<const double2*>(p)[i];
    const double2 s = reinterpret_cast<const double2*>(sv)[i];
    const double2 tv = reinterpret_cast<const double2*>(t)[i];
    // x = 1 x + alpha p + omega s ; r = 1 s - omega t (left to right)
    double xr = xv.x, xi = xv.y;
    xr += al.x * pv.x - al.y * pv.y;
    xi += al.x * pv.y + al.y * pv.x;
    xr += om.x * s.x - om.y * s.y;
    xi += om.x * s.y + om.y * s.x;
    reinterpret_cast<double2*>(x)[i] = make_double2(xr, xi);
    double rr = s.x, ri = s.y;
    rr += -om.x * tv.x + om.y * tv.y;
    ri += -om.x * tv.y - om.y * tv.x;
    reinterpret_cast<double2*>(r)[i] = make_double2(rr, ri);
    acc += rr * rr + ri * ri;
  }
  sr[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sr[threadIdx.x] += sr[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr[0];
    partial[2 * blockIdx.x + 1] = 0;
  }
}

__global__ void bicg_finish(BicgScalars* st, double* hist, double tol, int max_iter) {
  if (threadIdx.x != 0 || st->done) return;
  const double2 al = zdiv(st->rho_new, st->rv), om = bicg_omega(st);
  st->alpha = al;
  st->omega = om;
  st->rho_old = st->rho_new;
  const double res = sqrt(st->rn2.x);
  st->res = res;
  const int it = st->it + 1;
  st->it = it;
  hist[it] = res;
  if (!(res > tol) || (om.x == 0 && om.y == 0) || it >= max_iter) st->done = 1;
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

hipError_t LaunchRk4F64(const RhsArgs& a, int stage, bool horner, const void* in, double* rho,
                        double* acc, void* yout, double dt, hipStream_t s);
hipError_t LaunchRk4F32(const RhsArgs& a, int stage, bool horner, const void* in, double* rho,
                        double* acc, void* yout, double dt, hipStream_t s);
hipError_t LaunchRkf45F64(const RhsArgs& a, int stage, const void* in, const double* rho,
                         const Rkf45Bufs<double>& b, const Rkf45Coefs& c, hipStream_t s);
hipError_t LaunchRkf45F32(const RhsArgs& a, int stage, const void* in, const double* rho,
                         const Rkf45Bufs<float>& b, const Rkf45Coefs& c, hipStream_t s);

hipError_t LaunchRk4Stage(const RhsArgs& a, bool fp32, int stage, bool horner, const void* in,
                          double* rho, double* acc, void* yout, double dt, hipStream_t s) {
  return fp32 ? LaunchRk4F32(a, stage, horner, in, rho, acc, yout, dt, s)
              : LaunchRk4F64(a, stage, horner, in, rho, acc, yout, dt, s);
}

hipError_t LaunchRkf45Stage(const RhsArgs& a, bool fp32, int stage, const void* in,
                            const double* rho, const Rkf45Bufs<float>& bf,
                            const Rkf45Bufs<double>& bd, const Rkf45Coefs& c, hipStream_t s) {
  return fp32 ? LaunchRkf45F32(a, stage, in, rho, bf, c, s)
              : LaunchRkf45F64(a, stage, in, rho, bd, c, s);
}

hipError_t LaunchRhsGeneral(const RhsArgs& a, const double* x, double* y, int trace_row,
                            hipStream_t s, const unsigned char* active) {
  const size_t tot = (size_t)a.num_nodes * a.NE;
  int blocks = (int)((tot + 255) / 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(dev::rhs_general, dim3(blocks), dim3(256), 0, s, a, x, y, trace_row,
                     active);
  return hipGetLastError();
}

hipError_t LaunchSnapshot(const double* src, double* hist, int* counter, int ne, int capacity,
                          hipStream_t s) {
  hipLaunchKernelGGL(dev::snapshot_node0, dim3(1), dim3(256), 0, s, src, hist, counter, ne, capacity);
  return hipGetLastError();
}

hipError_t LaunchSnapshotMean(const double* src, const int* member_offsets, int members,
                              double* hist, int* counter, int ne, int capacity, hipStream_t s) {
  hipLaunchKernelGGL(dev::snapshot_mean, dim3(1), dim3(256), 0, s, src, member_offsets, members,
                     hist, counter, ne, capacity);
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

hipError_t LaunchBicgUpdate(int which, double* y, size_t n, const double* x0, const double* x1,
                            const double* x2, const BicgScalars* st, hipStream_t s) {
  int blocks = (int)((n + 255) / 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(dev::bicg_update, dim3(blocks), dim3(256), 0, s, which, y, n, x0, x1, x2,
                     st);
  return hipGetLastError();
}

hipError_t LaunchBicgXR(double* x, double* r, const double* p, const double* sv,
                        const double* t, size_t n, BicgScalars* st, double* partial, int nblocks,
                        hipStream_t s) {
  hipLaunchKernelGGL(dev::bicg_xr, dim3(nblocks), dim3(256), 0, s, x, r, p, sv, t, n, st,
                     partial);
  hipLaunchKernelGGL(dev::dot_final, dim3(1), dim3(256), 0, s, partial, nblocks,
                     reinterpret_cast<double*>(&st->rn2));
  return hipGetLastError();
}

hipError_t LaunchBicgFinish(BicgScalars* st, double* hist, double tol, int max_iter,
                            hipStream_t s) {
  hipLaunchKernelGGL(dev::bicg_finish, dim3(1), dim3(64), 0, s, st, hist, tol, max_iter);
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
