// Time-local truncation in the engine (PHI's TimeLocal option).
//
// The host advances the matrices (timelocal.h) on PHI's schedule and
// uploads Q after every advance; the device adds the term to the RHS of
// the deepest tier (bath_kernels.h, LaunchTimeLocal) after the operator
// RHS.  Runs take the explicit-k path (engine_filter.cc), FP64.
//   RK4 (hierarchy_integrator.cc:2721, 2784): Q advances by dt/2 after the
//   k0 combination (with H(t)) and after the k2 combination (with
//   H(t + dt/2)), each before that point's H update.
//   RKF45 (:2028-2247, 2386-2389, 2476-2480): before the H update of stage
//   s = 1..4 Q advances by a_s - a_(s-1) (a_0 = 0); before stage 5 it is
//   restored to its value at t and advanced by a5; an accepted step
//   advances it by dt - a5 and saves it; a rejected one restores it (here:
//   every attempt starts from the saved matrices, which is the same).
#include <cmath>
#include <cstring>

#include "bath_kernels.h"
#include "engine.h"
#include "timelocal.h"

namespace phi {

namespace {
constexpr int kTlSlots = 8;  // Q uploads in flight between synchronisations
constexpr double kPi = 3.14159265358979323846;
}  // namespace

int Engine::InitTimeLocal(const PhiParameters& p) {
  const int N = h_.N, M = h_.M, K = h_.K;
  if (bath_model_ == 0 && M > N) {
    err_ = "time-local truncation with more site-projector baths than states is not supported";
    return kErrUnsupported;
  }
  if (filter_on()) {
    err_ = "time-local truncation with FilterTolerance is not supported";
    return kErrUnsupported;
  }
  const double nu = 2 * kPi * p.boltzmann_constant * p.temperature / p.hbar;
  std::vector<double> rate((size_t)M * K);
  for (int m = 0; m < M; ++m)
    for (int k = 0; k < K; ++k) rate[(size_t)m * K + k] = k == 0 ? p.gamma[m] : nu * k;
  tl_ = new TimeLocal;
  tl_->Init(N, M, K, bath_model_, h_.corr.data(), rate.data());
  const size_t qn = (size_t)M * N * N;
  int rc = Check(hipMalloc((void**)&d_tlq_, qn * sizeof(double2)), "alloc time-local Q");
  if (!rc) rc = Check(hipMemset(d_tlq_, 0, qn * sizeof(double2)), "zero time-local Q");
  if (!rc)
    rc = Check(hipHostMalloc((void**)&h_tlstage_, kTlSlots * qn * sizeof(double2)),
               "pinned time-local staging");
  if (!rc && bath_model_ == 0) {
    std::vector<int> site(M);
    for (int m = 0; m < M; ++m) site[m] = h_.site_of_mode[(size_t)m * K];
    rc = Check(hipMalloc((void**)&d_tlsite_, M * sizeof(int)), "alloc time-local sites");
    if (!rc)
      rc = Check(hipMemcpy(d_tlsite_, site.data(), M * sizeof(int), hipMemcpyHostToDevice),
                 "upload time-local sites");
  }
  tl_slot_ = 0;
  return rc;
}

int Engine::TlUpload() {
  const size_t qn = (size_t)h_.M * N() * N();
  if (tl_slot_ >= kTlSlots) {
    // Every staging slot may still be read by a queued copy.
    int rc = Synchronize();
    if (rc) return rc;
    tl_slot_ = 0;
  }
  double2* slot = reinterpret_cast<double2*>(h_tlstage_) + (size_t)tl_slot_ * qn;
  const std::complex<double>* Q = tl_->Q();
  for (size_t i = 0; i < qn; ++i) slot[i] = make_double2(Q[i].real(), Q[i].imag());
  ++tl_slot_;
  return Check(hipMemcpyAsync(d_tlq_, slot, qn * sizeof(double2), hipMemcpyHostToDevice, stream_),
               "upload time-local Q");
}

int Engine::TlStep(double ts) {
  if (bath_model_ != 0 && h_opscur_.empty()) {
    err_ = "diagonal/full coupling operators were not given (bath_coupling_op)";
    return kErrInvalidParameters;
  }
  tl_->Step(h_hcur_.data(), ts, h_opscur_.empty() ? nullptr : h_opscur_.data());
  return TlUpload();
}

int Engine::TlTerm(const double* x, double* y) {
  const int first = h_.tier_start[h_.Lh - 1];
  TimeLocalArgs t{bath_model_, h_.M, first, num_nodes() - first, d_tlq_, d_ops_, d_tlsite_};
  return Check(LaunchTimeLocal(N(), t, x, y, stream_), "time-local term");
}

}  // namespace phi
