// RK4 with a time-dependent system Hamiltonian (PHI's annealing runs,
// optionally Hilbert-space truncated).
//
// PHI's RK4 loop (hierarchy_integrator.cc:2684-2826) calls
// UpdateTimeDependentHamiltonian at t (before k0), at t + dt/2 (before k1)
// and at t + dt (before k3); after an update of a truncated system it
// rotates, into the new eigenbasis, every ADO buffer that is still live:
// rho at t; rho, rho_tmp and k0 at t + dt/2; rho, rho_tmp, k0, k1 and k2 at
// t + dt.  The stage kernels here run PHI's classic order with an
// accumulator acc = rho + dt/6 k0 (+ dt/3 k1 + dt/3 k2); the rotation is
// linear, so rotating rho, the stage input Y and acc is the same change.
// The nested form is not used: the RHS is not time independent.
#include <cstring>
#include <vector>

#include "engine.h"

namespace phi {

namespace {
void HermitianParts(int N, const std::complex<double>* H, double2* Hs, double2* Hraw,
                    double2* Hadj) {
  double mean = 0;
  for (int a = 0; a < N; ++a) mean += H[(size_t)a * (N + 1)].real();
  mean /= N;
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) {
      const std::complex<double> v = H[(size_t)a + (size_t)b * N];
      Hraw[a + b * N] = make_double2(v.real(), v.imag());
      Hs[a + b * N] = make_double2(v.real() - (a == b ? mean : 0.0), v.imag());
      Hadj[b + a * N] = make_double2(v.real(), -v.imag());
    }
}
}  // namespace

// Staging slots of one step: up to three updates, each H (three forms) and
// R: 4 N^2 complex per update.
static constexpr int kSlotsPerStep = 3;

int Engine::SetHamiltonian(const std::complex<double>* H) {
  const int N = this->N(), NE = N * N;
  h_hcur_.assign(H, H + NE);
  int rc = kOk;
  if (!h_stage_) {
    rc = Check(hipHostMalloc((void**)&h_stage_, (size_t)kSlotsPerStep * 4 * NE * sizeof(double2)),
               "pinned staging");
    if (rc) return rc;
    stage_slot_ = 0;
  }
  if (stage_slot_ >= kSlotsPerStep) {
    // Every slot is in flight: wait for the stream before reusing them.
    rc = Synchronize();
    if (rc) return rc;
    stage_slot_ = 0;
  }
  double2* base = reinterpret_cast<double2*>(h_stage_) + (size_t)stage_slot_ * 4 * NE;
  HermitianParts(N, H, base, base + NE, base + 2 * NE);
  bool real_h = true;
  for (int i = 0; i < NE; ++i)
    if (H[i].imag() != 0) real_h = false;
  const size_t bytes = NE * sizeof(double2);
  rc = Check(hipMemcpyAsync(d_H_, base, bytes, hipMemcpyHostToDevice, stream_), "upload H");
  if (!rc) rc = Check(hipMemcpyAsync(d_Hraw_, base + NE, bytes, hipMemcpyHostToDevice, stream_),
                      "upload Hraw");
  if (!rc) rc = Check(hipMemcpyAsync(d_Hadj_, base + 2 * NE, bytes, hipMemcpyHostToDevice,
                                     stream_), "upload Hadj");
  // The kernels skip the imaginary part of H when it is zero everywhere.
  args_.real_h = real_h ? 1 : 0;
  return rc;
}

int Engine::Rotate(const std::complex<double>* R, int nbuf, void* const* bufs, const bool* fp32) {
  const int N = this->N(), NE = N * N;
  int rc = kOk;
  if (!d_rot_) rc = Check(hipMalloc((void**)&d_rot_, NE * sizeof(double2)), "alloc rotation");
  if (rc) return rc;
  // The slot of the H just staged holds R in its fourth part.
  double2* base = reinterpret_cast<double2*>(h_stage_) + (size_t)stage_slot_ * 4 * NE + 3 * NE;
  for (int i = 0; i < NE; ++i) base[i] = make_double2(R[i].real(), R[i].imag());
  rc = Check(hipMemcpyAsync(d_rot_, base, NE * sizeof(double2), hipMemcpyHostToDevice, stream_),
             "upload rotation");
  for (int b = 0; b < nbuf && !rc; ++b)
    rc = Check(LaunchRotateNodes(bufs[b], fp32[b], d_rot_, N, num_nodes(), stream_), "rotate");
  return rc;
}

int Engine::Rk4StepsTD(long long nsteps, const double* t_before, double dt, long long step_offset,
                       int snap_every, HamiltonianSource* src) {
  if (members_) return Unsupported("a time-dependent Hamiltonian");
  if (filter_on()) {
    err_ = "FilterTolerance with a time-dependent Hamiltonian is not supported";
    return kErrUnsupported;
  }
  if (closure_) {
    err_ = "a closure with a time-dependent Hamiltonian is not supported";
    return kErrUnsupported;
  }
  int rc = AllocBulk();
  if (rc) return rc;
  if (explicit_path()) {
    // Diagonal / full coupling operators, time-local truncation: the
    // explicit-k step applies the updates itself (engine_filter.cc).
    for (long long s = 0; s < nsteps && !rc; ++s)
      rc = ExplicitRk4Step(t_before[s], dt, snap_every > 0 && (step_offset + s) % snap_every == 0,
                           src);
    return rc;
  }
  const void* in[4] = {d_rho_, d_y_[0], d_y_[1], d_y_[0]};
  void* out[4] = {d_y_[0], d_y_[1], d_y_[0], nullptr};
  for (long long s = 0; s < nsteps && !rc; ++s) {
    const double t = t_before[s];
    const double when[4] = {t, t + dt / 2, -1, t + dt};
    for (int st = 0; st < 4 && !rc; ++st) {
      if (st != 2 && src->Update(when[st])) {
        rc = SetHamiltonian(src->H());
        const std::complex<double>* R = src->Rotation();
        if (!rc && R) {
          // Live buffers: rho at t; rho, Y and acc later in the step.
          void* bufs[3] = {d_rho_, d_y_[0], d_aux_};
          const bool f32[3] = {false, fp32_, false};
          rc = Rotate(R, st == 0 ? 1 : 3, bufs, f32);
        }
        ++stage_slot_;
      }
      if (!rc)
        rc = Check(LaunchRk4Stage(StageArgs(st), fp32_, st, kRk4Classic, in[st], d_rho_, d_aux_,
                                  out[st], dt, stream_), "rk4 stage");
    }
    if (!rc && snap_every > 0 && (step_offset + s) % snap_every == 0)
      rc = Check(LaunchSnapshot(d_rho_, d_hist_, d_hist_count_, N() * N(), hist_capacity_, stream_),
                 "snapshot");
    // The step's staging slots are reused by the next step.
    if (!rc) rc = Synchronize();
    stage_slot_ = 0;
  }
  return rc;
}

}  // namespace phi
