#include "timelocal.h"

#include <cmath>

namespace phi {

void TimeLocal::Init(int N, int M, int K, int model, const Z* corr, const double* rate) {
  N_ = N;
  M_ = M;
  K_ = K;
  model_ = model;
  corr_.assign(corr, corr + (size_t)M * K);
  rate_.assign(rate, rate + (size_t)M * K);
  const size_t NE = (size_t)N * N;
  T_.assign((size_t)M * K * NE, Z(0, 0));
  saved_ = T_;
  Q_.assign((size_t)M * NE, Z(0, 0));
  X_.assign(NE, Z(0, 0));
  Y_.assign(NE, Z(0, 0));
}

void TimeLocal::Step(const Z* H, double ts, const Z* ops) {
  if (ts < 1e-9) return;
  const int N = N_;
  const size_t NE = (size_t)N * N;
  // P = 1 - i H ts; P^dagger(r, c) = conj(P(c, r)).
  auto P = [&](int r, int c) { return Z(r == c ? 1.0 : 0.0, 0) - Z(0, ts) * H[r + (size_t)c * N]; };
  for (int m = 0; m < M_; ++m)
    for (int k = 0; k < K_; ++k) {
      Z* T = T_.data() + ((size_t)m * K_ + k) * NE;
      const Z f = ts * corr_[(size_t)m * K_ + k];
      // X = T + ts c V'
      for (size_t e = 0; e < NE; ++e) X_[e] = T[e];
      if (model_ == 0) {
        X_[(size_t)m * (N + 1)] += f;
      } else if (model_ == 1) {
        for (int a = 0; a < N; ++a) X_[(size_t)a * (N + 1)] += f * ops[(size_t)m * N + a];
      } else {
        for (size_t e = 0; e < NE; ++e) X_[e] += f * ops[(size_t)m * NE + e];
      }
      // Y = P^dagger X, T = damp Y P
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
          Z s(0, 0);
          for (int n = 0; n < N; ++n) s += std::conj(P(n, r)) * X_[n + (size_t)c * N];
          Y_[r + (size_t)c * N] = s;
        }
      const double damp = std::exp(-rate_[(size_t)m * K_ + k] * ts);
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
          Z s(0, 0);
          for (int n = 0; n < N; ++n) s += Y_[r + (size_t)n * N] * P(n, c);
          T[r + (size_t)c * N] = damp * s;
        }
    }
}

const TimeLocal::Z* TimeLocal::Q() {
  const size_t NE = (size_t)N_ * N_;
  for (int m = 0; m < M_; ++m) {
    Z* q = Q_.data() + (size_t)m * NE;
    for (size_t e = 0; e < NE; ++e) q[e] = Z(0, 0);
    for (int k = 0; k < K_; ++k) {
      const Z* T = T_.data() + ((size_t)m * K_ + k) * NE;
      for (size_t e = 0; e < NE; ++e) q[e] += T[e];
    }
  }
  return Q_.data();
}

}  // namespace phi
