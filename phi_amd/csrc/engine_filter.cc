// Explicit-k RK4 and RKF45: the runs whose RHS the fused stage kernels do
// not cover.  The RHS of a stage is evaluated into its k buffer, then PHI's
// classic combinations run as separate FP64 kernels.  Two cases:
//
// * Adaptive truncation (FilterTolerance > 0; IndependentBathAdaptiveTruncation's
//   DoNext / DoPrevious, hierarchy_updater.cc:256-304, called from the base
//   Get, hierarchy_updater.h:36-52 -- Get is not virtual and the integrator
//   holds a HierarchyUpdater*, hierarchy_integrator.cc:1798, 2011, so the
//   class's own Get is never called; flag loops hierarchy_integrator.cc:
//   2358-2385 for RKF45 and 2828-2851 for RK4).  With the flags fixed during
//   a step, k_s = L(input) with the terms of inactive neighbours skipped;
//   every node keeps its own terms.  Flags change after each accepted step.
//   PHI's RK4 updates a thread's
//   flags right after its own final update while other threads may still
//   evaluate k3 (no barrier, hierarchy_integrator.cc:2810-2851); with one
//   thread the update follows the whole step, the order used here.
// * Diagonal and full coupling operators (DiagonalBath / FullBath,
//   bath_kernels.h): k_s = L(input) with the general-coupling RHS kernel.
//   With a time-dependent Hamiltonian the H updates, the basis rotations of
//   the live buffers (rho, the stage input and the earlier k's, as
//   hierarchy_integrator.cc:2684-2826 for RK4 and 1984-2270 for RKF45) and
//   the rotated coupling operators of a truncated system are applied
//   between stages.
#include <cstring>
#include <vector>

#include "bath_kernels.h"
#include "engine.h"
#include "filter_kernels.h"

namespace phi {

int Engine::ExplicitAlloc(int nk) {
  const size_t bytes = state_elems() * 2 * sizeof(double);
  int rc = kOk;
  for (int s = 0; s < nk && !rc; ++s) {
    if (d_kf_[s]) continue;
    rc = Check(hipMalloc((void**)&d_kf_[s], bytes), "alloc stage k");
    if (!rc) rc = Check(hipMemsetAsync(d_kf_[s], 0, bytes, stream_), "zero stage k");
  }
  return rc;
}

int Engine::OperatorRhs(const double* x, double* y) {
  int rc;
  if (bath_model_ != 0 && !d_ops_) {
    err_ = "diagonal/full coupling operators were not given (bath_coupling_op)";
    return kErrInvalidParameters;
  }
  if (bath_model_ != 0) {
    BathArgs b{bath_model_, h_.M, h_.K, d_ops_, d_alpha_, d_corr_};
    rc = Check(LaunchRhsBath(args_, b, x, y, stream_), "rhs_bath");
  } else {
    rc = Check(LaunchRhsGeneral(args_, x, y, 0, stream_, filter_on() ? d_active_ : nullptr),
               "rhs_general");
  }
  if (!rc && tl_) rc = TlTerm(x, y);
  return rc;
}

int Engine::ExplicitRhs(const double* in, int s) { return OperatorRhs(in, d_kf_[s]); }

int Engine::ExplicitUpdate(HamiltonianSource* src, double t, int nbuf, void* const* bufs) {
  if (!src || !src->Update(t)) return kOk;
  int rc = SetHamiltonian(src->H());
  if (!rc && src->BathOps()) rc = SetBathOps(src->BathOps());
  if (!rc && src->Rotation()) {
    bool f64[8];
    for (int i = 0; i < nbuf; ++i) f64[i] = false;
    rc = Rotate(src->Rotation(), nbuf, bufs, f64);
  }
  ++stage_slot_;
  return rc;
}

int Engine::ExplicitRk4Step(double t, double dt, bool snapshot, HamiltonianSource* src) {
  if (fp32_) {
    err_ = "adaptive truncation, diagonal/full coupling operators and time-local "
           "truncation run in FP64 only (BulkPrecision=fp64)";
    return kErrUnsupported;
  }
  int rc = ExplicitAlloc(4);
  if (rc) return rc;
  const size_t n = state_elems();
  double* y = (double*)d_y_[0];
  const double half = dt / 2;
  // H at t (rho live), at t + dt/2 (rho, Y, k0), at t + dt (rho, Y, k0..k2).
  {
    void* b[1] = {d_rho_};
    rc = ExplicitUpdate(src, t, 1, b);
  }
  if (!rc) rc = ExplicitRhs(d_rho_, 0);
  const double* k0[1] = {d_kf_[0]};
  if (!rc) rc = Check(LaunchFilterCombine(d_rho_, k0, &half, 1, y, n, stream_), "combine");
  if (!rc && tl_) rc = TlStep(dt / 2);  // Q(t + dt/2), with H(t)
  if (!rc) {
    void* b[3] = {d_rho_, y, d_kf_[0]};
    rc = ExplicitUpdate(src, t + dt / 2, 3, b);
  }
  if (!rc) rc = ExplicitRhs(y, 1);
  const double* k1[1] = {d_kf_[1]};
  if (!rc) rc = Check(LaunchFilterCombine(d_rho_, k1, &half, 1, y, n, stream_), "combine");
  if (!rc) rc = ExplicitRhs(y, 2);
  const double* k2[1] = {d_kf_[2]};
  if (!rc) rc = Check(LaunchFilterCombine(d_rho_, k2, &dt, 1, y, n, stream_), "combine");
  if (!rc && tl_) rc = TlStep(dt / 2);  // Q(t + dt), with H(t + dt/2)
  if (!rc) {
    void* b[5] = {d_rho_, y, d_kf_[0], d_kf_[1], d_kf_[2]};
    rc = ExplicitUpdate(src, t + dt, 5, b);
  }
  if (!rc) rc = ExplicitRhs(y, 3);
  const double* k4[4] = {d_kf_[0], d_kf_[1], d_kf_[2], d_kf_[3]};
  if (!rc) rc = Check(LaunchFilterRk4Final(d_rho_, k4, dt / 6, n, stream_), "rk4 update");
  if (!rc && filter_on())
    rc = Check(LaunchUpdateActive(d_rho_, d_active_, N() * N(), num_nodes(), filter_, stream_),
               "flag update");
  if (!rc && snapshot)
    rc = Check(LaunchSnapshot(d_rho_, d_hist_, d_hist_count_, N() * N(), hist_capacity_, stream_),
               "snapshot");
  if (!rc && (src || tl_)) {
    // The step's staging slots are reused by the next step.
    rc = Synchronize();
    stage_slot_ = 0;
    tl_slot_ = 0;
  }
  return rc;
}

int Engine::ExplicitRkf45Attempt(const Rkf45Coefs& c, double* err, double t, const double* a,
                                 HamiltonianSource* src) {
  if (fp32_) {
    err_ = "adaptive truncation, diagonal/full coupling operators and time-local "
           "truncation run in FP64 only (BulkPrecision=fp64)";
    return kErrUnsupported;
  }
  int rc = ExplicitAlloc(6);
  if (rc) return rc;
  const size_t n = state_elems();
  double* y = (double*)d_y_[0];
  rc = Check(LaunchFillU64(d_err_, 0ull, 1, stream_), "zero err");
  if (tl_) {
    if (!a) {
      err_ = "time-local truncation needs the RKF45 stage times";
      return kErrInvalidArgument;
    }
    for (int i = 0; i < 5; ++i) tl_a_[i] = a[i];
    tl_->Restore();
    if (!rc) rc = TlUpload();
  }
  // Time-local matrices before the H update of stage s: advanced by the
  // stage-time increment (s = 1..4), restored and advanced by a5 (s = 5).
  auto tl_advance = [&](int s) -> int {
    if (!tl_ || s == 0) return kOk;
    if (s == 5) {
      tl_->Restore();
      return TlStep(a[4]);
    }
    return TlStep(s == 1 ? a[0] : a[s - 1] - a[s - 2]);
  };
  // H at t (rho live) and at t + a[s-1] before stage s (rho, Y and
  // k0..k(s-1) live, k1 excepted before stage 5).
  auto update = [&](int s) -> int {
    int e = tl_advance(s);
    if (e || !src) return e;
    void* b[6];
    int nb = 0;
    b[nb++] = d_rho_;
    if (s > 0) {
      b[nb++] = y;
      for (int j = 0; j < s; ++j)
        if (!(s == 5 && j == 1)) b[nb++] = d_kf_[j];
    }
    return ExplicitUpdate(src, s == 0 ? t : t + a[s - 1], nb, b);
  };
  const double* in = d_rho_;
  for (int s = 0; s < 5 && !rc; ++s) {
    rc = update(s);
    if (!rc) rc = ExplicitRhs(in, s);
    const double* k[5];
    double w[5];
    for (int j = 0; j <= s; ++j) {
      k[j] = d_kf_[j];
      w[j] = c.b[s][j];
    }
    if (!rc) rc = Check(LaunchFilterCombine(d_rho_, k, w, s + 1, y, n, stream_), "combine");
    in = y;
  }
  if (!rc) rc = update(5);
  if (!rc) rc = ExplicitRhs(y, 5);
  const double* k6[6] = {d_kf_[0], d_kf_[1], d_kf_[2], d_kf_[3], d_kf_[4], d_kf_[5]};
  if (!rc) rc = Check(LaunchFilterRkf45Final(d_rho_, k6, c.c4, c.c5, d_aux_, d_err_, n, stream_),
                      "rkf45 final");
  unsigned long long bits = 0;
  if (!rc) rc = Check(hipMemcpyAsync(&bits, d_err_, sizeof(bits), hipMemcpyDeviceToHost, stream_),
                      "err");
  if (!rc) rc = Synchronize();
  // The attempt's staging slots are free again.
  stage_slot_ = 0;
  tl_slot_ = 0;
  if (rc) return rc;
  std::memcpy(err, &bits, sizeof(double));
  return kOk;
}

int Engine::SetBathOps(const std::complex<double>* ops) {
  if (bath_model_ == 0) return kOk;
  const size_t n = (size_t)h_.M * (bath_model_ == 1 ? N() : N() * N());
  h_opscur_.assign(ops, ops + n);
  int rc = kOk;
  if (!h_ops_) {
    rc = Check(hipHostMalloc((void**)&h_ops_, 3 * n * sizeof(double2)), "pinned operators");
    if (rc) return rc;
  }
  if (!d_ops_) rc = Check(hipMalloc((void**)&d_ops_, n * sizeof(double2)), "alloc operators");
  if (rc) return rc;
  // A slot per update of a step (the step synchronises before reuse).
  double2* slot = reinterpret_cast<double2*>(h_ops_) + (size_t)(stage_slot_ % 3) * n;
  for (size_t i = 0; i < n; ++i) slot[i] = make_double2(ops[i].real(), ops[i].imag());
  return Check(hipMemcpyAsync(d_ops_, slot, n * sizeof(double2), hipMemcpyHostToDevice, stream_),
               "upload operators");
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
