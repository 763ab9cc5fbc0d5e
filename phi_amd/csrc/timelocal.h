// Time-local truncation of the hierarchy (PHI's TimeLocal option): the
// deepest tier's missing deeper ADOs are replaced by a time-local term
//   -sum_m [V_m, Q_m rho_n - rho_n Q_m^dagger]
// (hierarchy_updater.cc:28-43, 94-107, 170-213), with matrices Q_m that the
// host advances between stages (hierarchy.cc:264-341,
// hierarchy_integrator.cc:1675-1731):
//   T_mk <- exp(-nu_k ts) P^dagger (T_mk + ts c_mk V'_m) P,  P = 1 - i H ts,
//   Q_m = sum_k T_mk,
// H the current system Hamiltonian as stored (PHI does not divide it by
// hbar here), nu_0 = gamma_m, nu_k = 2 pi k kB T / hbar, c_mk the
// correlation constants.  V'_m (the seed) is the unit matrix element (m, m)
// for site-projector baths (AddIndependentBathOperator uses the bath's own
// index), diag(b_m) for diagonal and F_m for full coupling operators.
// N and M are small (N x N matrices, M K of them), so this runs on the host
// in FP64 and the device receives Q after each advance.
#pragma once

#include <complex>
#include <vector>

namespace phi {

class TimeLocal {
 public:
  typedef std::complex<double> Z;
  // model: 0 site projectors, 1 diagonal, 2 full.  corr: M K constants
  // (bath-major); rate: M K decay rates.
  void Init(int N, int M, int K, int model, const Z* corr, const double* rate);
  // Advances every T_mk by ts with Hamiltonian H (N x N column-major) and
  // the current coupling operators (M N diagonal entries or M N^2 full
  // matrices; unused for site projectors).  Steps below 1e-9 are skipped.
  void Step(const Z* H, double ts, const Z* ops);
  void Store() { saved_ = T_; }
  void Restore() { T_ = saved_; }
  // Q_m (M x N x N, column-major), summed over k in order.
  const Z* Q();
  int M() const { return M_; }
  int N() const { return N_; }
  int K() const { return K_; }
  // All T_mk (M x K x N x N) and their saved copy, for restart files.
  std::vector<Z>& terms() { return T_; }
  std::vector<Z>& saved() { return saved_; }

 private:
  int N_ = 0, M_ = 0, K_ = 0, model_ = 0;
  std::vector<Z> corr_;
  std::vector<double> rate_;
  std::vector<Z> T_, saved_, Q_, X_, Y_;
};

}  // namespace phi
