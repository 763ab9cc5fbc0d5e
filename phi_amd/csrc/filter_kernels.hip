// Combination, final-update and flag kernels of the explicit-k path; the
// semantics are stated in filter_kernels.h.  All arithmetic is FP64.
#include "filter_kernels.h"

#include "heom_device.h"

namespace phi {
namespace dev {

struct KList {
  const double2* k[6];
  double c[6];
};

__global__ void __launch_bounds__(256) filter_combine(const double2* __restrict__ rho, KList L,
                                                      int nk, double2* __restrict__ out,
                                                      size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double2 s = rho[i];
    for (int t = 0; t < nk; ++t) {
      const double2 v = L.k[t][i];
      s.x += v.x * L.c[t];
      s.y += v.y * L.c[t];
    }
    out[i] = s;
  }
}

__global__ void __launch_bounds__(256) filter_rk4_final(double2* __restrict__ rho, KList L,
                                                        size_t n) {
  const double c = L.c[0];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 a = L.k[0][i], b = L.k[1][i], d = L.k[2][i], e = L.k[3][i];
    const double sx = a.x + 2. * b.x + 2. * d.x + e.x;
    const double sy = a.y + 2. * b.y + 2. * d.y + e.y;
    const double2 r = rho[i];
    rho[i] = make_double2(r.x + c * sx, r.y + c * sy);
  }
}

// L.k = (k0, k2, k3, k4, k5), L.c = c4 (4 weights) then c5 in c5w.
__global__ void __launch_bounds__(256) filter_rkf45_final(const double2* __restrict__ rho, KList L,
                                                          double c50, double c52, double c53,
                                                          double c54, double c55,
                                                          double2* __restrict__ out,
                                                          unsigned long long* err, size_t n) {
  __shared__ double red[4];
  double m = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 r = rho[i];
    const double2 k0 = L.k[0][i], k2 = L.k[1][i], k3 = L.k[2][i], k4 = L.k[3][i], k5 = L.k[4][i];
    const double y4x = r.x + k0.x * L.c[0] + k2.x * L.c[1] + k3.x * L.c[2] + k4.x * L.c[3];
    const double y4y = r.y + k0.y * L.c[0] + k2.y * L.c[1] + k3.y * L.c[2] + k4.y * L.c[3];
    const double y5x = r.x + k0.x * c50 + k2.x * c52 + k3.x * c53 + k4.x * c54 + k5.x * c55;
    const double y5y = r.y + k0.y * c50 + k2.y * c52 + k3.y * c53 + k4.y * c54 + k5.y * c55;
    out[i] = make_double2(y5x, y5y);
    const double d = hypot(y4x - y5x, y4y - y5y);
    m = d > m ? d : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(m, o, 64);
    m = v > m ? v : m;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) b = red[w] > b ? red[w] : b;
    atomic_max_nonneg(err, b);
  }
}

// One wavefront per node.
__global__ void __launch_bounds__(256) update_active(const double2* __restrict__ rho,
                                                     unsigned char* __restrict__ act, int ne,
                                                     int nodes, double tol) {
  const int node = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (node >= nodes) return;
  const double2* r = rho + (size_t)node * ne;
  double m = 0;
  for (int e = lane; e < ne; e += 64) {
    const double a = hypot(r[e].x, r[e].y);
    m = a > m ? a : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(m, o, 64);
    m = v > m ? v : m;
  }
  if (lane == 0) {
    const unsigned char a = act[node];
    if (m < tol && a) act[node] = 0;
    else if (m > tol && !a) act[node] = 1;
  }
}

}  // namespace dev

namespace {
int GridFor(size_t n) {
  size_t b = (n + 255) / 256;
  return (int)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}
}  // namespace

hipError_t LaunchFilterCombine(const double* rho, const double* const* k, const double* c, int nk,
                               double* out, size_t n, hipStream_t s) {
  if (nk < 0 || nk > 6) return hipErrorInvalidValue;
  dev::KList L{};
  for (int t = 0; t < nk; ++t) {
    L.k[t] = (const double2*)k[t];
    L.c[t] = c[t];
  }
  hipLaunchKernelGGL(dev::filter_combine, dim3(GridFor(n)), dim3(256), 0, s, (const double2*)rho,
                     L, nk, (double2*)out, n);
  return hipGetLastError();
}

hipError_t LaunchFilterRk4Final(double* rho, const double* const* k, double c, size_t n,
                                hipStream_t s) {
  dev::KList L{};
  for (int t = 0; t < 4; ++t) L.k[t] = (const double2*)k[t];
  L.c[0] = c;
  hipLaunchKernelGGL(dev::filter_rk4_final, dim3(GridFor(n)), dim3(256), 0, s, (double2*)rho, L,
                     n);
  return hipGetLastError();
}

hipError_t LaunchFilterRkf45Final(const double* rho, const double* const* k, const double* c4,
                                  const double* c5, double* out, unsigned long long* err,
                                  size_t n, hipStream_t s) {
  dev::KList L{};
  const int idx[5] = {0, 2, 3, 4, 5};
  for (int t = 0; t < 5; ++t) L.k[t] = (const double2*)k[idx[t]];
  for (int t = 0; t < 4; ++t) L.c[t] = c4[t];
  hipLaunchKernelGGL(dev::filter_rkf45_final, dim3(GridFor(n)), dim3(256), 0, s,
                     (const double2*)rho, L, c5[0], c5[1], c5[2], c5[3], c5[4], (double2*)out,
                     err, n);
  return hipGetLastError();
}

hipError_t LaunchUpdateActive(const double* rho, unsigned char* active, int ne, int nodes,
                              double tol, hipStream_t s) {
  const int blocks = (nodes + 3) / 4;
  hipLaunchKernelGGL(dev::update_active, dim3(blocks), dim3(256), 0, s, (const double2*)rho,
                     active, ne, nodes, tol);
  return hipGetLastError();
}

}  // namespace phi
