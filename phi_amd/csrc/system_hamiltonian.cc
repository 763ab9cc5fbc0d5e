// Host side of time-dependent and truncated system Hamiltonians.
#include "system_hamiltonian.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

namespace phi {

// --- Hermitian eigen-decomposition (cyclic Jacobi) -----------------------------
// Each rotation acts on the pair (p, q) with the unitary
//   W = diag(1, e*) [[c, s], [-s, c]],  e = a_pq / |a_pq|,
// which first makes a_pq real and then zeroes it with the real Jacobi angle
// t = tan(theta) = sgn(tau) / (|tau| + sqrt(1 + tau^2)),
// tau = (a_qq - a_pp) / (2 |a_pq|) (Golub and Van Loan, Matrix Computations,
// section 8.5); A <- W^dagger A W, V <- V W.
void HermitianEigen(int n, const cplx* A, int k, double* w, cplx* V) {
  std::vector<cplx> a((size_t)n * n), v((size_t)n * n, 0.0);
  for (int c = 0; c < n; ++c)
    for (int r = 0; r < n; ++r) {
      // Lower triangle as given, upper triangle its conjugate.
      a[r + (size_t)c * n] = r >= c ? A[r + (size_t)c * n] : std::conj(A[c + (size_t)r * n]);
    }
  for (int i = 0; i < n; ++i) v[i + (size_t)i * n] = 1.0;
  double scale = 0;
  for (const cplx& x : a) scale = std::max(scale, std::abs(x));
  const double eps = std::numeric_limits<double>::epsilon();
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0;
    for (int q = 1; q < n; ++q)
      for (int p = 0; p < q; ++p) off = std::max(off, std::abs(a[p + (size_t)q * n]));
    if (off <= eps * eps * std::max(scale, std::numeric_limits<double>::min())) break;
    for (int q = 1; q < n; ++q)
      for (int p = 0; p < q; ++p) {
        const cplx apq = a[p + (size_t)q * n];
        const double g = std::abs(apq);
        if (g == 0) continue;
        const cplx e = apq / g;
        const double app = a[p + (size_t)p * n].real(), aqq = a[q + (size_t)q * n].real();
        const double tau = (aqq - app) / (2 * g);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::abs(tau) + std::sqrt(1 + tau * tau));
        const double c = 1 / std::sqrt(1 + t * t), s = t * c;
        const cplx ec = std::conj(e);
        // columns: B = A W
        for (int r = 0; r < n; ++r) {
          const cplx xp = a[r + (size_t)p * n], xq = a[r + (size_t)q * n];
          a[r + (size_t)p * n] = c * xp - s * ec * xq;
          a[r + (size_t)q * n] = s * xp + c * ec * xq;
          const cplx vp = v[r + (size_t)p * n], vq = v[r + (size_t)q * n];
          v[r + (size_t)p * n] = c * vp - s * ec * vq;
          v[r + (size_t)q * n] = s * vp + c * ec * vq;
        }
        // rows: A = W^dagger B
        for (int col = 0; col < n; ++col) {
          const cplx xp = a[p + (size_t)col * n], xq = a[q + (size_t)col * n];
          a[p + (size_t)col * n] = c * xp - s * e * xq;
          a[q + (size_t)col * n] = s * xp + c * e * xq;
        }
        a[p + (size_t)q * n] = 0;
        a[q + (size_t)p * n] = 0;
        a[p + (size_t)p * n] = a[p + (size_t)p * n].real();
        a[q + (size_t)q * n] = a[q + (size_t)q * n].real();
      }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return a[x + (size_t)x * n].real() < a[y + (size_t)y * n].real();
  });
  for (int j = 0; j < k; ++j) {
    const int src = order[j];
    w[j] = a[src + (size_t)src * n].real();
    int big = 0;
    double mag = -1;
    for (int r = 0; r < n; ++r) {
      const double m = std::abs(v[r + (size_t)src * n]);
      if (m > mag * (1 + 1e-12)) {
        mag = m;
        big = r;
      }
    }
    const cplx ph = std::conj(v[big + (size_t)src * n]) / std::abs(v[big + (size_t)src * n]);
    for (int r = 0; r < n; ++r) V[r + (size_t)j * n] = v[r + (size_t)src * n] * ph;
  }
}

// --- CubicSpline ----------------------------------------------------------------
CubicSpline::CubicSpline(const std::vector<double>& y) {
  const size_t m = y.size();
  x_.resize(m);
  for (size_t i = 0; i < m; ++i) x_[i] = i * (1. / (m - 1));
  a_ = y;
  b_.assign(m, 0);
  c_.assign(m, 0);
  d_.assign(m, 0);
  b_[0] = 3. * (y[1] - y[0]);
  b_[m - 1] = 3. * (y[m - 1] - y[m - 2]);
  for (size_t i = 1; i + 1 < m; ++i) b_[i] = 3. * (y[i + 1] - y[i]);
  // Symmetric positive definite tridiagonal solve, L D L^T (the
  // factorisation dpttrf computes; dpttrs then solves).
  std::vector<double> dg(m, 4.), e(m > 1 ? m - 1 : 0, 1.);
  dg[0] = dg[m - 1] = 2.;
  for (size_t i = 0; i + 1 < m; ++i) {
    const double l = e[i] / dg[i];
    dg[i + 1] -= l * e[i];
    e[i] = l;
  }
  for (size_t i = 1; i < m; ++i) b_[i] -= e[i - 1] * b_[i - 1];
  b_[m - 1] /= dg[m - 1];
  for (size_t i = m - 1; i-- > 0;) b_[i] = b_[i] / dg[i] - e[i] * b_[i + 1];
  for (size_t i = 0; i + 1 < m; ++i) {
    c_[i] = 3. * (a_[i + 1] - a_[i]) - 2. * b_[i] - b_[i + 1];
    d_[i] = 2. * (a_[i] - a_[i + 1]) + b_[i] + b_[i + 1];
  }
}

double CubicSpline::Get(double x) {
  const size_t m = x_.size();
  if (x > x_[m - 1]) return a_[m - 1];
  if (x < x_[0]) return a_[0];
  size_t i = last_;
  if (!(x > x_[i] && x < x_[i + 1])) {
    while (i + 2 < m && x > x_[i + 1]) ++i;
    while (i > 0 && x < x_[i]) --i;
    last_ = i;
  }
  const double t = (x - x_[i]) / (x_[i + 1] - x_[i]);
  return a_[i] + b_[i] * t + c_[i] * t * t + d_[i] * t * t * t;
}

// --- SystemHamiltonian ----------------------------------------------------------
SystemHamiltonian::~SystemHamiltonian() {
  delete fi_;
  delete ff_;
}

std::string SystemHamiltonian::Init(const PhiParameters& p) {
  n_ = p.num_states;
  full_ = std::max(p.hilbert_space_size, n_);
  T_ = p.integration_time;
  min_change_ = p.minimum_field_change;
  const size_t FE = (size_t)full_ * full_;
  if (!p.time_indepenent_hamiltonian.empty()) {
    if (p.time_indepenent_hamiltonian.size() != FE)
      return "the Hamiltonian must be HilbertSpaceSize x HilbertSpaceSize";
    Hstatic_.assign(p.time_indepenent_hamiltonian.begin(), p.time_indepenent_hamiltonian.end());
    td_ = false;
  } else {
    if (p.initial_annealing_hamiltonian.size() != FE || p.final_annealing_hamiltonian.size() != FE)
      return "a Hamiltonian is required (static, or both annealing Hamiltonians)";
    if (p.initial_hamiltonian_field.size() < 2 || p.final_hamiltonian_field.size() < 2)
      return "annealing fields need at least two values";
    Hi_.assign(p.initial_annealing_hamiltonian.begin(), p.initial_annealing_hamiltonian.end());
    Hf_.assign(p.final_annealing_hamiltonian.begin(), p.final_annealing_hamiltonian.end());
    fi_ = new CubicSpline(std::vector<double>(p.initial_hamiltonian_field.begin(),
                                              p.initial_hamiltonian_field.end()));
    ff_ = new CubicSpline(std::vector<double>(p.final_hamiltonian_field.begin(),
                                              p.final_hamiltonian_field.end()));
    td_ = true;
  }
  // Start: the static H, or the annealing H at alpha = 0
  // (hierarchy_integrator.cc:924-946); the multipliers' last values start
  // at the largest double so the first update always happens (:518-523).
  old_i_ = old_f_ = std::numeric_limits<double>::max();
  std::vector<cplx> Hfull;
  if (td_) FullAt(0.0, &Hfull);
  else Hfull = Hstatic_;
  SetFrom(Hfull, false);
  if (truncated() && p.is_diagonal_bath_coupling)
    return "diagonal coupling operators of a truncated system (HilbertSpaceSize > NumStates) "
           "are not supported";
  if (truncated() && p.is_full_bath_coupling) {
    M_ = p.num_bath_couplings;
    if (p.bath_coupling_op.size() != (size_t)M_ * FE) return "BathCouplingOp has the wrong size";
    ops_.assign((size_t)M_ * n_ * n_, 0.0);
    std::vector<cplx> F(FE);
    for (int m = 0; m < M_; ++m) {
      for (size_t e = 0; e < FE; ++e) F[e] = p.bath_coupling_op[(size_t)m * FE + e];
      Truncate(F.data(), ops_.data() + (size_t)m * n_ * n_);
    }
  }
  return "";
}

void SystemHamiltonian::FullAt(double alpha, std::vector<cplx>* Hfull) {
  const cplx ai = fi_->Get(alpha), af = ff_->Get(alpha);
  Hfull->resize(Hi_.size());
  for (size_t i = 0; i < Hi_.size(); ++i) (*Hfull)[i] = ai * Hi_[i] + af * Hf_[i];
}

void SystemHamiltonian::SetFrom(const std::vector<cplx>& Hfull, bool rotate) {
  if (!truncated()) {
    H_ = Hfull;
    return;
  }
  std::vector<cplx> Vold = V_;
  V_.assign((size_t)full_ * n_, 0.0);
  w_.assign(n_, 0.0);
  HermitianEigen(full_, Hfull.data(), n_, w_.data(), V_.data());
  H_.assign((size_t)n_ * n_, 0.0);
  Truncate(Hfull.data(), H_.data());
  // R(i, j) = <new_i | old_j>: an ADO in the old basis a maps to R a R^dagger.
  R_.assign((size_t)n_ * n_, 0.0);
  if (rotate && !Vold.empty()) {
    for (int i = 0; i < n_; ++i)
      for (int j = 0; j < n_; ++j) {
        cplx s = 0;
        for (int r = 0; r < full_; ++r)
          s += std::conj(V_[r + (size_t)i * full_]) * Vold[r + (size_t)j * full_];
        R_[i + (size_t)j * n_] = s;
      }
  } else {
    for (int i = 0; i < n_; ++i) R_[i + (size_t)i * n_] = 1.0;
  }
}

bool SystemHamiltonian::Update(double t) {
  if (!td_) return false;
  const double alpha = t / T_;
  const double vi = fi_->Get(alpha), vf = ff_->Get(alpha);
  const double di = std::fabs(vi / old_i_ - 1.0), df = std::fabs(vf / old_f_ - 1.0);
  if (di < min_change_ && df < min_change_) return false;
  old_i_ = vi;
  old_f_ = vf;
  std::vector<cplx> Hfull;
  FullAt(alpha, &Hfull);
  SetFrom(Hfull, true);
  // Coupling operators follow the basis: F <- R F R^dagger.
  const size_t NE = (size_t)n_ * n_;
  std::vector<cplx> T(NE);
  for (int m = 0; m < M_; ++m) {
    cplx* F = ops_.data() + (size_t)m * NE;
    for (int c = 0; c < n_; ++c)
      for (int r = 0; r < n_; ++r) {
        cplx s = 0;
        for (int k = 0; k < n_; ++k) s += F[r + (size_t)k * n_] * std::conj(R_[c + (size_t)k * n_]);
        T[r + (size_t)c * n_] = s;  // F R^dagger
      }
    for (int c = 0; c < n_; ++c)
      for (int r = 0; r < n_; ++r) {
        cplx s = 0;
        for (int k = 0; k < n_; ++k) s += R_[r + (size_t)k * n_] * T[k + (size_t)c * n_];
        F[r + (size_t)c * n_] = s;
      }
  }
  return true;
}

void SystemHamiltonian::Truncate(const cplx* A, cplx* out) const {
  if (!truncated()) {
    std::copy(A, A + (size_t)n_ * n_, out);
    return;
  }
  // out(j, i) = <v_j | A | v_i> (hilbert_space_truncator.cc:34-61).
  std::vector<cplx> Av(full_);
  for (int i = 0; i < n_; ++i) {
    for (int r = 0; r < full_; ++r) {
      cplx s = 0;
      for (int c = 0; c < full_; ++c) s += A[r + (size_t)c * full_] * V_[c + (size_t)i * full_];
      Av[r] = s;
    }
    for (int j = 0; j < n_; ++j) {
      cplx s = 0;
      for (int r = 0; r < full_; ++r) s += std::conj(V_[r + (size_t)j * full_]) * Av[r];
      out[j + (size_t)i * n_] = s;
    }
  }
}

void SystemHamiltonian::UnTruncate(const cplx* a, cplx* full) const {
  if (!truncated()) {
    std::copy(a, a + (size_t)n_ * n_, full);
    return;
  }
  // full = V a V^dagger (hilbert_space_truncator.cc:84-95).
  for (int c = 0; c < full_; ++c)
    for (int r = 0; r < full_; ++r) {
      cplx s = 0;
      for (int i = 0; i < n_; ++i)
        for (int j = 0; j < n_; ++j)
          s += V_[r + (size_t)j * full_] * a[j + (size_t)i * n_] *
               std::conj(V_[c + (size_t)i * full_]);
      full[r + (size_t)c * full_] = s;
    }
}

}  // namespace phi
