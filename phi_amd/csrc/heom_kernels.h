// Device argument blocks and launchers of the HEOM kernels (heom_kernels.hip).
#ifndef PHI_AMD_HEOM_KERNELS_H_
#define PHI_AMD_HEOM_KERNELS_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace phi {

// Everything the RHS of one hierarchy needs, as device pointers.
struct RhsArgs {
  int N, NE, D, L, num_nodes, tile, vacuum;
  double inv_hbar;
  const int* next;          // [num_nodes * D]
  const int* prev;          // [num_nodes * D]
  const uint8_t* occ;       // [num_nodes * D]
  const double* mu;         // [num_nodes]
  const int* site_ptr;      // [N + 1]
  const int* site_modes;    // [D]
  const double* next_s;     // [D * L]
  const double2* prev_col;  // [D * L]
  const double2* prev_row;  // [D * L]
  const double* delta;      // [N]
  const double2* H;         // N*N, shifted by mean(diag H) (Hermitian path)
  const double2* Hraw;      // N*N, as given (general path)
  const double2* Hadj;      // N*N, conjugate transpose of Hraw
};

template <typename TB>
struct Rkf45Bufs {
  TB* k[5];                  // k0..k4 (bulk precision)
  TB* tmp_out;               // next stage input
  double* rho_new;           // y5 (FP64)
  unsigned long long* err;   // max |y4 - y5| as the bits of a double
};

// RKF45 weights already multiplied by the current dt (host keeps PHI's
// in-place rescaling).  b[s][i]: stage-(s+1) input weights; c4: weights of
// k0,k2,k3,k4 in y4; c5: weights of k0,k2,k3,k4,k5 in y5.
struct Rkf45Coefs {
  double b[5][5];
  double c4[4];
  double c5[5];
};

size_t RhsSharedBytes(const RhsArgs& a, bool fp32);

hipError_t LaunchRk4Stage(const RhsArgs& a, bool fp32, int stage, const void* in, double* rho,
                          double* acc, void* yout, double dt, hipStream_t s);
hipError_t LaunchRkf45Stage(const RhsArgs& a, bool fp32, int stage, const void* in,
                            const double* rho, const Rkf45Bufs<float>& bf,
                            const Rkf45Bufs<double>& bd, const Rkf45Coefs& c, hipStream_t s);
hipError_t LaunchRhsGeneral(const RhsArgs& a, const double* x, double* y, int trace_row,
                            hipStream_t s);
hipError_t LaunchSnapshot(const double* src, double* hist, int* counter, int ne, int capacity,
                          hipStream_t s);
hipError_t LaunchFillU64(unsigned long long* p, unsigned long long v, int n, hipStream_t s);
hipError_t LaunchCastToFloat(const double* src, float* dst, size_t n, hipStream_t s);
hipError_t LaunchDot(const double* u, const double* v, size_t n, int mode, double* partial,
                     int nblocks, double* out, hipStream_t s);
hipError_t LaunchLincomb(double* y, size_t n, int nv, const double* const* x, const double2* c,
                         hipStream_t s);

}  // namespace phi

#endif  // PHI_AMD_HEOM_KERNELS_H_
