// RHS of the diagonal and full coupling models (formulas in bath_kernels.h).
// One block per node (grid-stride over nodes); the node's ADO and one
// commutator live in LDS, each thread keeps up to four output elements in
// registers.  Neighbour ADOs and the operators are read from global memory
// (they stay in L2).  FP64 throughout.
#include "bath_kernels.h"

#include "heom_device.h"

namespace phi {
namespace dev {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) {
  return make_double2(s * a.x, s * a.y);
}
// -i z
__device__ __forceinline__ double2 mul_negi(double2 z) { return make_double2(z.y, -z.x); }

constexpr int kMaxPerThread = 4;  // N <= 32 with 256 threads

__global__ void __launch_bounds__(256) rhs_bath(RhsArgs A, BathArgs B,
                                                const double2* __restrict__ x,
                                                double2* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = A.N, NE = A.NE, D = A.D, L = A.L;
  double2* xs = reinterpret_cast<double2*>(smem);
  double2* cs = xs + NE;
  for (int n = blockIdx.x; n < A.num_nodes; n += gridDim.x) {
    __syncthreads();
    for (int e = threadIdx.x; e < NE; e += blockDim.x) xs[e] = x[(size_t)n * NE + e];
    __syncthreads();
    double2 out[kMaxPerThread];
#pragma unroll
    for (int i = 0; i < kMaxPerThread; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      out[i] = make_double2(0., 0.);
      if (e >= NE) continue;
      const int a = e % N, b = e / N;  // element (a, b)
      // -(i/hbar)(H x - x H^dagger)
      double2 u = make_double2(0., 0.);
      for (int c = 0; c < N; ++c) {
        u = cadd(u, cmul(A.Hraw[a + c * N], xs[c + b * N]));
        u = csub(u, cmul(xs[a + c * N], A.Hadj[c + b * N]));
      }
      double2 o = cscale(A.inv_hbar, mul_negi(u));
      if (!A.vacuum) {
        o = csub(o, cscale(A.mu[n], xs[e]));
        for (int j = 0; j < D; ++j) {
          const int m = j / B.K;
          const int k = A.occ[(size_t)n * D + j];
          const int nx = A.next[(size_t)n * D + j];
          const int pv = A.prev[(size_t)n * D + j];
          if (B.model == 1) {
            const double2 ba = B.ops[m * N + a], bb = B.ops[m * N + b];
            if (nx >= 0)
              o = cadd(o, cmul(mul_negi(csub(ba, bb)), x[(size_t)nx * NE + e]));
            if (pv >= 0) {
              const double2 c = B.corr[j];
              const double2 op = mul_negi(csub(cmul(ba, c), cmul(bb, cconj(c))));
              o = cadd(o, cscale((double)k, cmul(op, x[(size_t)pv * NE + e])));
            }
          } else {
            const double2* F = B.ops + (size_t)m * NE;
            if (nx >= 0) {
              const double2* xp = x + (size_t)nx * NE;
              double2 comm = make_double2(0., 0.);
              for (int c = 0; c < N; ++c) {
                comm = cadd(comm, cmul(F[a + c * N], xp[c + b * N]));
                comm = csub(comm, cmul(xp[a + c * N], F[c + b * N]));
              }
              o = cadd(o, cscale(A.next_s[j * L + k], mul_negi(comm)));
            }
            if (pv >= 0) {
              const double2* xp = x + (size_t)pv * NE;
              double2 fx = make_double2(0., 0.), xf = make_double2(0., 0.);
              for (int c = 0; c < N; ++c) {
                fx = cadd(fx, cmul(F[a + c * N], xp[c + b * N]));
                xf = cadd(xf, cmul(xp[a + c * N], F[c + b * N]));
              }
              o = csub(o, cadd(cmul(A.prev_row[j * L + k], fx), cmul(A.prev_col[j * L + k], xf)));
            }
          }
        }
        if (B.model == 1) {
          for (int m = 0; m < B.M; ++m) {
            const double2 d = csub(B.ops[m * N + a], B.ops[m * N + b]);
            o = csub(o, cscale(B.alpha[m], cmul(cmul(d, d), xs[e])));
          }
        }
      }
      out[i] = o;
    }
    // Full operators: - alpha_m [F_m, [F_m, x]] through C = F x - x F in LDS.
    if (B.model == 2 && !A.vacuum) {
      for (int m = 0; m < B.M; ++m) {
        const double2* F = B.ops + (size_t)m * NE;
        __syncthreads();
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
          const int a = e % N, b = e / N;
          double2 c2 = make_double2(0., 0.);
          for (int c = 0; c < N; ++c) {
            c2 = cadd(c2, cmul(F[a + c * N], xs[c + b * N]));
            c2 = csub(c2, cmul(xs[a + c * N], F[c + b * N]));
          }
          cs[e] = c2;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kMaxPerThread; ++i) {
          const int e = threadIdx.x + i * blockDim.x;
          if (e >= NE) continue;
          const int a = e % N, b = e / N;
          double2 d2 = make_double2(0., 0.);
          for (int c = 0; c < N; ++c) {
            d2 = cadd(d2, cmul(F[a + c * N], cs[c + b * N]));
            d2 = csub(d2, cmul(cs[a + c * N], F[c + b * N]));
          }
          out[i] = csub(out[i], cscale(B.alpha[m], d2));
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kMaxPerThread; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      if (e < NE) y[(size_t)n * NE + e] = out[i];
    }
  }
}

// One block per node of the range; x_n, Q_m, X_m (and F_m) in LDS.
__global__ void __launch_bounds__(256) timelocal_term(int N, TimeLocalArgs T,
                                                      const double2* __restrict__ x,
                                                      double2* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int NE = N * N;
  double2* xs = reinterpret_cast<double2*>(smem);
  double2* qs = xs + NE;
  double2* Xs = qs + NE;
  double2* fs = Xs + NE;
  for (int b = blockIdx.x; b < T.count; b += gridDim.x) {
    const size_t n = (size_t)T.first + b;
    __syncthreads();
    for (int e = threadIdx.x; e < NE; e += blockDim.x) xs[e] = x[n * NE + e];
    double2 out[kMaxPerThread];
#pragma unroll
    for (int i = 0; i < kMaxPerThread; ++i) out[i] = make_double2(0, 0);
    for (int m = 0; m < T.M; ++m) {
      __syncthreads();
      for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        qs[e] = T.Q[(size_t)m * NE + e];
        if (T.model == 2) fs[e] = T.ops[(size_t)m * NE + e];
      }
      __syncthreads();
      for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        const int r = e % N, c = e / N;
        double2 s = make_double2(0, 0);
        for (int k = 0; k < N; ++k) {
          s = cadd(s, cmul(qs[r + k * N], xs[k + c * N]));
          s = csub(s, cmul(xs[r + k * N], cconj(qs[c + k * N])));
        }
        Xs[e] = s;
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < kMaxPerThread; ++i) {
        const int e = threadIdx.x + i * blockDim.x;
        if (e >= NE) continue;
        const int r = e % N, c = e / N;
        if (T.model == 0) {
          const int st = T.site[m];
          if (c == st) out[i] = cadd(out[i], Xs[e]);
          if (r == st) out[i] = csub(out[i], Xs[e]);
        } else if (T.model == 1) {
          const double2 d = csub(T.ops[(size_t)m * N + r], T.ops[(size_t)m * N + c]);
          out[i] = csub(out[i], cmul(d, Xs[e]));
        } else {
          double2 s = make_double2(0, 0);
          for (int k = 0; k < N; ++k) {
            s = cadd(s, cmul(fs[r + k * N], Xs[k + c * N]));
            s = csub(s, cmul(Xs[r + k * N], fs[k + c * N]));
          }
          out[i] = csub(out[i], s);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < kMaxPerThread; ++i) {
      const int e = threadIdx.x + i * blockDim.x;
      if (e < NE) y[n * NE + e] = cadd(y[n * NE + e], out[i]);
    }
  }
}

}  // namespace dev

hipError_t LaunchTimeLocal(int N, const TimeLocalArgs& t, const double* x, double* y,
                           hipStream_t s) {
  if (N < 1 || N * N > 256 * dev::kMaxPerThread || t.model < 0 || t.model > 2 || t.M < 1 ||
      !t.Q || (t.model != 0 && !t.ops) || (t.model == 0 && !t.site))
    return hipErrorInvalidValue;
  if (t.count <= 0) return hipSuccess;
  const size_t sh = 4 * (size_t)N * N * sizeof(double2);
  const int grid = t.count < 65535 ? t.count : 65535;
  hipLaunchKernelGGL(dev::timelocal_term, dim3(grid), dim3(256), sh, s, N, t, (const double2*)x,
                     (double2*)y);
  return hipGetLastError();
}

hipError_t LaunchRhsBath(const RhsArgs& a, const BathArgs& b, const double* x, double* y,
                         hipStream_t s) {
  if (a.N < 1 || a.NE > 256 * dev::kMaxPerThread || (b.model != 1 && b.model != 2))
    return hipErrorInvalidValue;
  const size_t sh = 2 * (size_t)a.NE * sizeof(double2);
  const int grid = a.num_nodes < 65535 ? a.num_nodes : 65535;
  hipLaunchKernelGGL(dev::rhs_bath, dim3(grid), dim3(256), sh, s, a, b, (const double2*)x,
                     (double2*)y);
  return hipGetLastError();
}

}  // namespace phi
