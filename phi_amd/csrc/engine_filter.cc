// Adaptive hierarchy truncation (PHI's FilterTolerance > 0): the RK4 and
// RKF45 steps of the filtered operator and the flag update.
//
// PHI (IndependentBathAdaptiveTruncation, hierarchy_updater.h:217-245,
// hierarchy_updater.cc:256-304; flag loops hierarchy_integrator.cc:2358-2385
// for RKF45 and 2828-2851 for RK4).  Stage s of a step, with the flags fixed:
//   kraw = L(input with the ADOs of inactive nodes zeroed)   (plain operator)
//   k_s  = kraw on active nodes, k_s(previous contents) + kraw on inactive
//          ones (PHI's per-node k buffers persist across steps and attempts)
// then PHI's classic combinations, all in FP64.  Flags change after each
// accepted step.  PHI's RK4 updates a thread's flags right after its own
// final update while other threads may still evaluate k3 (no barrier
// between them, hierarchy_integrator.cc:2810-2851); with one thread the
// update follows the whole step, which is the order used here.
#include <cstring>
#include <vector>

#include "engine.h"
#include "filter_kernels.h"

namespace phi {

int Engine::FilterAlloc(int nk) {
  const size_t bytes = state_elems() * 2 * sizeof(double);
  int rc = kOk;
  if (!d_ym_) rc = Check(hipMalloc((void**)&d_ym_, bytes), "alloc filter input");
  if (!rc && !d_kraw_) rc = Check(hipMalloc((void**)&d_kraw_, bytes), "alloc filter rhs");
  for (int s = 0; s < nk && !rc; ++s) {
    if (d_kf_[s]) continue;
    rc = Check(hipMalloc((void**)&d_kf_[s], bytes), "alloc filter k");
    if (!rc) rc = Check(hipMemsetAsync(d_kf_[s], 0, bytes, stream_), "zero filter k");
  }
  return rc;
}

int Engine::FilterRhs(const double* in, int s) {
  const int ne = N() * N(), nodes = num_nodes();
  int rc = Check(LaunchMaskNodes(in, d_active_, d_ym_, ne, nodes, stream_), "mask nodes");
  if (!rc) rc = Check(LaunchRhsGeneral(args_, d_ym_, d_kraw_, 0, stream_), "rhs_general");
  if (!rc && s >= 0) rc = Check(LaunchFilterK(d_kraw_, d_kf_[s], d_active_, ne, nodes, stream_),
                                "filter k");
  return rc;
}

int Engine::FilterRk4Step(double dt, bool snapshot) {
  if (fp32_) {
    err_ = "FilterTolerance (adaptive truncation) runs in FP64 only (BulkPrecision=fp64)";
    return kErrUnsupported;
  }
  int rc = FilterAlloc(4);
  if (rc) return rc;
  const size_t n = state_elems();
  double* y = (double*)d_y_[0];
  const double half = dt / 2;
  const double c[3] = {half, half, dt};
  const double* in = d_rho_;
  for (int s = 0; s < 3 && !rc; ++s) {
    rc = FilterRhs(in, s);
    const double* k[1] = {d_kf_[s]};
    if (!rc) rc = Check(LaunchFilterCombine(d_rho_, k, &c[s], 1, y, n, stream_), "filter combine");
    in = y;
  }
  if (!rc) rc = FilterRhs(y, 3);
  const double* k4[4] = {d_kf_[0], d_kf_[1], d_kf_[2], d_kf_[3]};
  if (!rc) rc = Check(LaunchFilterRk4Final(d_rho_, k4, dt / 6, n, stream_), "filter rk4 update");
  if (!rc) rc = Check(LaunchUpdateActive(d_rho_, d_active_, N() * N(), num_nodes(), filter_,
                                         stream_), "flag update");
  if (!rc && snapshot)
    rc = Check(LaunchSnapshot(d_rho_, d_hist_, d_hist_count_, N() * N(), hist_capacity_, stream_),
               "snapshot");
  return rc;
}

int Engine::FilterRkf45Attempt(const Rkf45Coefs& c, double* err) {
  if (fp32_) {
    err_ = "FilterTolerance (adaptive truncation) runs in FP64 only (BulkPrecision=fp64)";
    return kErrUnsupported;
  }
  int rc = FilterAlloc(6);
  if (rc) return rc;
  const size_t n = state_elems();
  double* y = (double*)d_y_[0];
  rc = Check(LaunchFillU64(d_err_, 0ull, 1, stream_), "zero err");
  const double* in = d_rho_;
  for (int s = 0; s < 5 && !rc; ++s) {
    rc = FilterRhs(in, s);
    const double* k[5];
    double w[5];
    for (int j = 0; j <= s; ++j) {
      k[j] = d_kf_[j];
      w[j] = c.b[s][j];
    }
    if (!rc) rc = Check(LaunchFilterCombine(d_rho_, k, w, s + 1, y, n, stream_), "filter combine");
    in = y;
  }
  if (!rc) rc = FilterRhs(y, 5);
  const double* k6[6] = {d_kf_[0], d_kf_[1], d_kf_[2], d_kf_[3], d_kf_[4], d_kf_[5]};
  if (!rc) rc = Check(LaunchFilterRkf45Final(d_rho_, k6, c.c4, c.c5, d_aux_, d_err_, n, stream_),
                      "filter rkf45 final");
  unsigned long long bits = 0;
  if (!rc) rc = Check(hipMemcpyAsync(&bits, d_err_, sizeof(bits), hipMemcpyDeviceToHost, stream_),
                      "err");
  if (!rc) rc = Synchronize();
  if (rc) return rc;
  std::memcpy(err, &bits, sizeof(double));
  return kOk;
}

int Engine::ActiveFlags(int* out) {
  const int nodes = num_nodes();
  if (!filter_on()) {
    for (int i = 0; i < nodes; ++i) out[i] = 1;
    return nodes;
  }
  std::vector<unsigned char> f(nodes);
  int rc = Check(hipMemcpyAsync(f.data(), d_active_, nodes, hipMemcpyDeviceToHost, stream_),
                 "flags");
  if (!rc) rc = Synchronize();
  if (rc) return rc;
  int count = 0;
  for (int i = 0; i < nodes; ++i) {
    out[i] = f[i] ? 1 : 0;
    count += out[i];
  }
  return count;
}

}  // namespace phi
