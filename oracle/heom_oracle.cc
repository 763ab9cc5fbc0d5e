// CPU oracle for the HEOM hot path of PHI (strj/phi).
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain FP64 restatement of the
// reference's algorithm, used by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py as the checker.  Nothing in phi_amd/ links,
// loads or calls it.
//
// What it restates (file:line in the reference):
//   * hierarchy index graph, PHI id order  ... hierarchy.cc:45-155
//   * +/- neighbour linking                ... hierarchy_integrator.cc:1204-1338
//   * correlation-function constants       ... hierarchy_integrator.cc:427-473
//   * per-node prefactors                  ... hierarchy_integrator.cc:1343-1466
//   * initial ADO scaling                  ... hierarchy_integrator.cc:1075-1091
//   * IndependentBath RHS (Get)            ... hierarchy_updater.h:36-75,
//                                              hierarchy_updater.cc:129-144,
//                                              216-254 (NOBLAS Gemm order,
//                                              blas_wrapper.cc:11-56)
//   * Shi filter RHS (adaptive truncation) ... hierarchy_updater.h:36-52,
//                                              hierarchy_updater.cc:256-304:
//     HierarchyUpdater::Get is not virtual and the integrator calls it
//     through a HierarchyUpdater* (hierarchy_integrator.cc:1798, 2011), so
//     the base Get runs: every node's own terms (DoUnitary overwrites drho),
//     and the adaptive DoNext / DoPrevious, which skip inactive neighbours.
//     IndependentBathAdaptiveTruncation::Get (hierarchy_updater.h:227-239)
//     is never called.
//   * Vacuum updater (Lh == 1)             ... hierarchy_updater.h:251-265
//   * RK4 loop (step count, update order)  ... hierarchy_integrator.cc:2578-2939
//   * RKF45 loop and dt controller         ... hierarchy_integrator.cc:1797-2573
//
// Steady state: PHI's solver is unusable as shipped (PrepForSteadyState reads
// an unassigned hamiltonian_, hierarchy_integrator.cc:1496), so the steady
// state operator is defined here, not restated: A x = L x with the (0,0)
// element of node 0 replaced by the trace of node 0; b = e_(0,(0,0)).
//
// Threading: nodes are split into contiguous blocks, one per pthread; each
// node's work is independent, so results do not depend on the thread count
// (as in PHI, hierarchy_integrator.cc:113-175).

#include <pthread.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

typedef std::complex<double> cplx;

namespace {

const double kPi = 3.14159265358979323846264338327950288419716939937510;

struct Node {
  std::vector<int> index;        // multi-index n (D entries)
  int tier = 0;                  // |n|
  bool active = true;            // Shi filter flag
  double dephasing = 0;          // mu_n
  std::vector<int> next_node, next_mode, next_site;
  std::vector<cplx> next_pref;
  std::vector<int> prev_node, prev_mode, prev_site;
  std::vector<cplx> prev_row, prev_col;
};

struct Oracle {
  int N = 0, M = 0, K = 0, D = 0, Lh = 0;
  double hbar = 0, kb = 0, temperature = 0;
  std::vector<double> gamma, lambda;
  std::vector<int> bath_site;
  bool normalized = true;
  double filter = 0;
  bool vacuum = false;
  std::vector<cplx> H, Hadj;      // column-major N x N
  // Coupling model (hierarchy_integrator.cc:1772-1790): 0 IndependentBath
  // (site projectors), 1 DiagonalBath, 2 FullBath.  bath_ops: M operators,
  // N entries each (diagonal) or N x N column-major (full).
  int bath_model = 0;
  std::vector<cplx> bath_ops;
  // Time-local truncation: the summed matrices Q_m (M x N x N, column-major),
  // applied to the nodes of the deepest tier (hierarchy_integrator.cc:
  // 1318-1333) in place of their (empty) DoNext.
  bool tl_on = false;
  std::vector<cplx> tl_Q;
  std::vector<cplx> corr;         // c_{jk}, D entries
  std::vector<double> abs_corr;
  std::vector<cplx> matsubara;    // Delta_j, M entries
  std::vector<Node> nodes;
  std::vector<cplx> rho;          // nodes * N * N, PHI id order
  // Per-stage k buffers (PHI's k0..k5; Get overwrites them every stage).
  std::vector<cplx> kbuf[6];
  int threads = 1;
};

// hierarchy.cc:45-155.  Tier 0 is the zero index, tier 1 is e_0..e_{D-1} in
// order, and tier l >= 2 is filled by visiting, for k = 0..D-1, every node n
// of tier l-1 in id order and incrementing entry (n + k) % D; new indices get
// the next id.
void ConstructIndex(Oracle* o) {
  const int D = o->D, Lh = o->Lh;
  std::map<std::vector<int>, int> lookup;
  std::vector<std::vector<int>> idx;
  idx.push_back(std::vector<int>(D, 0));
  lookup[idx[0]] = 0;
  std::vector<int> level_count(Lh > 0 ? Lh : 1, 0);
  level_count[0] = 1;
  if (Lh >= 2) {
    for (int i = 0; i < D; ++i) {
      std::vector<int> v(D, 0);
      v[i] = 1;
      lookup[v] = (int)idx.size();
      idx.push_back(v);
    }
    level_count[1] = D;
  }
  int count = (int)idx.size();
  for (int level = 2; level < Lh; ++level) {
    int added = 0;
    const int prev_start = count - level_count[level - 1];
    for (int k = 0; k < D; ++k) {
      for (int n = 0; n < level_count[level - 1]; ++n) {
        std::vector<int> v = idx[prev_start + n];
        v[(n + k) % D] += 1;
        if (lookup.count(v) == 0) {
          lookup[v] = count + added;
          idx.push_back(v);
          ++added;
        }
      }
    }
    level_count[level] = added;
    count += added;
  }
  o->nodes.resize(idx.size());
  for (size_t i = 0; i < idx.size(); ++i) {
    Node& nd = o->nodes[i];
    nd.index = idx[i];
    nd.tier = 0;
    for (int v : idx[i]) nd.tier += v;
  }
  // Linking, hierarchy_integrator.cc:1240-1320.  prev: modes with n_j > 0 in
  // increasing j.  next: only when |n| < Lh - 1, then every mode in order.
  for (size_t i = 0; i < idx.size(); ++i) {
    Node& nd = o->nodes[i];
    for (int j = 0; j < D; ++j) {
      if (nd.index[j] > 0) {
        std::vector<int> v = nd.index;
        v[j] -= 1;
        auto it = lookup.find(v);
        if (it != lookup.end()) {
          nd.prev_node.push_back(it->second);
          nd.prev_mode.push_back(j);
          nd.prev_site.push_back(o->bath_site[j / o->K]);
        }
      }
    }
    if (nd.tier < Lh - 1) {
      for (int j = 0; j < D; ++j) {
        std::vector<int> v = nd.index;
        v[j] += 1;
        auto it = lookup.find(v);
        if (it != lookup.end()) {
          nd.next_node.push_back(it->second);
          nd.next_mode.push_back(j);
          nd.next_site.push_back(o->bath_site[j / o->K]);
        }
      }
    }
  }
}

// hierarchy_integrator.cc:427-473 (constants) and 1343-1466 (prefactors).
void InitializeConstants(Oracle* o) {
  const double hbar = o->hbar;
  const double thermal = o->kb * o->temperature;
  const double nu = 2 * kPi * thermal / hbar;
  const int K = o->K;
  o->corr.assign(o->D, cplx(0, 0));
  o->abs_corr.assign(o->D, 0.0);
  for (int i = 0; i < o->M; ++i) {
    const double g = o->gamma[i], l = o->lambda[i];
    if (g > 0) {
      o->corr[i * K] = cplx(g * l / hbar) *
                       cplx(1 / std::tan(g * hbar / (2 * thermal)), -1);
    }
    o->abs_corr[i * K] = std::abs(o->corr[i * K]);
    for (int k = 1; k < K; ++k) {
      if (g > 0) {
        o->corr[i * K + k] = cplx(4 * l * g * thermal / hbar / hbar * (nu * k) /
                                  (std::pow(nu * k, 2) - std::pow(g, 2)));
      }
      o->abs_corr[i * K + k] = std::abs(o->corr[i * K + k]);
    }
  }
  o->matsubara.assign(o->M, cplx(0, 0));
  for (int j = 0; j < o->M; ++j) {
    const double g = o->gamma[j], l = o->lambda[j];
    if (g > 0) {
      o->matsubara[j] = cplx(2 * l * thermal / hbar / hbar / g -
                                 l / hbar / std::tan(hbar * g / (2 * thermal)),
                             0);
      for (int k = 1; k < K; ++k) {
        o->matsubara[j] -= cplx(4 * l * thermal / hbar / hbar * g /
                                    (std::pow(nu * k, 2) - std::pow(g, 2)),
                                0);
      }
    }
  }
  for (Node& nd : o->nodes) {
    cplx deph(0, 0);
    for (int m = 0; m < o->D; ++m) {
      const bool drude = m % K == 0;
      const int k = drude ? m / K : m % K;
      const double nu_ak = drude ? o->gamma[k] : nu * k;
      deph += cplx(nd.index[m] * nu_ak);
    }
    nd.dephasing = deph.real();
    nd.next_pref.resize(nd.next_node.size());
    for (size_t j = 0; j < nd.next_node.size(); ++j) {
      cplx p = cplx(0, 1);
      if (o->normalized) {
        const int nm = nd.next_mode[j];
        p *= cplx(std::sqrt((nd.index[nm] + 1) * o->abs_corr[nm]));
      }
      nd.next_pref[j] = p;
    }
    nd.prev_row.resize(nd.prev_node.size());
    nd.prev_col.resize(nd.prev_node.size());
    for (size_t j = 0; j < nd.prev_node.size(); ++j) {
      const int nm = nd.prev_mode[j];
      cplx row = cplx(0, nd.index[nm]) * o->corr[nm];
      cplx col = cplx(0, -nd.index[nm]) * std::conj(o->corr[nm]);
      if (o->normalized) {
        const cplx norm =
            cplx(1, 0) / cplx(std::sqrt(nd.index[nm] * o->abs_corr[nm]));
        row *= norm;
        col *= norm;
      }
      nd.prev_row[j] = row;
      nd.prev_col[j] = col;
    }
  }
}

// c = alpha * A B + beta * c for N x N column-major matrices (NOBLAS Gemm
// order, blas_wrapper.cc:11-31: c(a,b) = beta c(a,b) + sum_n alpha (A(a,n)
// B(n,b))).
void Gemm(int N, cplx alpha, const cplx* A, const cplx* B, cplx beta, cplx* C) {
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) {
      cplx c = C[a + b * N] * beta;
      for (int n = 0; n < N; ++n) c += alpha * (A[a + n * N] * B[n + b * N]);
      C[a + b * N] = c;
    }
}

// The coupling terms of DiagonalBath and FullBath (hierarchy_updater.cc:
// 10-127) with the operators of Hierarchy::FillDiagonalBathOperators /
// FillFullBathOperators (hierarchy.cc:157-225):
//   diagonal: same_m(r,c) = alpha_m (b_r - b_c)^2, next_m(r,c) = -i (b_r - b_c),
//             prev_{mk}(r,c) = -i (b_r c_mk - b_c conj(c_mk));
//             DoNext adds next_m .* rho_{n+e} (no prefactor), DoPrevious adds
//             n_mk prev_mk .* rho_{n-e} (PHI's DiagonalBath ignores the ADO
//             normalisation);
//   full:     same_m = sqrt(alpha_m) F_m, next_m = -i F_m, prev_m = -F_m;
//             DoMatsubara: drho -= [S, [S, rho]];
//             DoNext: drho += (-i p) next_m rho' + (i p) rho' next_m;
//             DoPrevious: drho += row prev_m rho' + col rho' prev_m.
// Called after DoUnitary and DoDephasing.
void NodeRhsBath(const Oracle* o, const cplx* x, const Node& nd, const cplx* rho, cplx* drho) {
  const int N = o->N, NE = N * N, K = o->K;
  const cplx I(0, 1);
  if (o->bath_model == 1) {
    auto b = [&](int m, int r) { return o->bath_ops[(size_t)m * N + r]; };
    for (int m = 0; m < o->M; ++m)
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
          const cplx same = o->matsubara[m] * std::pow(b(m, r) - b(m, c), 2);
          drho[r + c * N] -= same * rho[r + c * N];
        }
    for (size_t j = 0; j < nd.next_node.size(); ++j) {
      const int m = nd.next_mode[j] / K;
      const cplx* r1 = x + (size_t)nd.next_node[j] * NE;
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r)
          drho[r + c * N] += -I * (b(m, r) - b(m, c)) * r1[r + c * N];
    }
    for (size_t j = 0; j < nd.prev_node.size(); ++j) {
      const int mode = nd.prev_mode[j], m = mode / K;
      const cplx cc = o->corr[mode];
      const cplx pf(nd.index[mode], 0);
      const cplx* r1 = x + (size_t)nd.prev_node[j] * NE;
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r)
          drho[r + c * N] +=
              pf * (-I * (b(m, r) * cc - b(m, c) * std::conj(cc))) * r1[r + c * N];
    }
    return;
  }
  // Full coupling operators.
  std::vector<cplx> S(NE), T(NE), F(NE);
  for (int m = 0; m < o->M; ++m) {
    const cplx sq = std::sqrt(o->matsubara[m]);
    for (int e = 0; e < NE; ++e) S[e] = sq * o->bath_ops[(size_t)m * NE + e];
    Gemm(N, cplx(1, 0), S.data(), rho, cplx(0, 0), T.data());     // T = S rho
    Gemm(N, cplx(-1, 0), rho, S.data(), cplx(1, 0), T.data());    // T -= rho S
    Gemm(N, cplx(-1, 0), S.data(), T.data(), cplx(1, 0), drho);   // drho -= S T
    Gemm(N, cplx(1, 0), T.data(), S.data(), cplx(1, 0), drho);    // drho += T S
  }
  for (size_t j = 0; j < nd.next_node.size(); ++j) {
    const int m = nd.next_mode[j] / K;
    for (int e = 0; e < NE; ++e) F[e] = -I * o->bath_ops[(size_t)m * NE + e];
    const cplx* r1 = x + (size_t)nd.next_node[j] * NE;
    Gemm(N, -I * nd.next_pref[j], F.data(), r1, cplx(1, 0), drho);
    Gemm(N, I * nd.next_pref[j], r1, F.data(), cplx(1, 0), drho);
  }
  for (size_t j = 0; j < nd.prev_node.size(); ++j) {
    const int m = nd.prev_mode[j] / K;
    for (int e = 0; e < NE; ++e) F[e] = -o->bath_ops[(size_t)m * NE + e];
    const cplx* r1 = x + (size_t)nd.prev_node[j] * NE;
    Gemm(N, nd.prev_row[j], F.data(), r1, cplx(1, 0), drho);
    Gemm(N, nd.prev_col[j], r1, F.data(), cplx(1, 0), drho);
  }
}

// DoTimelocal of the three coupling models (hierarchy_updater.cc:28-43,
// 94-107, 170-213): with X_m = Q_m rho - rho Q_m^dagger,
//   IndependentBath: column s += X_m(:, s), row s -= X_m(s, :)  (s = site of
//                    bath m), i.e. drho -= [P_s, X_m];
//   DiagonalBath:    drho(r, c) -= (b_r - b_c) X_m(r, c);
//   FullBath:        drho -= F_m X_m - X_m F_m.
void NodeTimeLocal(const Oracle* o, const cplx* rho, cplx* drho) {
  const int N = o->N, NE = N * N;
  std::vector<cplx> X(NE), Qa(NE);
  for (int m = 0; m < o->M; ++m) {
    const cplx* Q = o->tl_Q.data() + (size_t)m * NE;
    for (int r = 0; r < N; ++r)
      for (int c = 0; c < N; ++c) Qa[r + c * N] = std::conj(Q[c + r * N]);
    Gemm(N, cplx(1, 0), Q, rho, cplx(0, 0), X.data());       // X = Q rho
    Gemm(N, cplx(-1, 0), rho, Qa.data(), cplx(1, 0), X.data());  // X -= rho Q^dagger
    if (o->bath_model == 0) {
      const int s = o->bath_site[m];
      for (int r = 0; r < N; ++r) drho[r + s * N] += X[r + s * N];
      for (int c = 0; c < N; ++c) drho[s + c * N] -= X[s + c * N];
    } else if (o->bath_model == 1) {
      const cplx* b = o->bath_ops.data() + (size_t)m * N;
      for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) drho[r + c * N] -= (b[r] - b[c]) * X[r + c * N];
    } else {
      const cplx* F = o->bath_ops.data() + (size_t)m * NE;
      Gemm(N, cplx(-1, 0), F, X.data(), cplx(1, 0), drho);  // drho -= F X
      Gemm(N, cplx(1, 0), X.data(), F, cplx(1, 0), drho);   // drho += X F
    }
  }
}

// hierarchy_updater.h:36-75 and hierarchy_updater.cc:129-144, 216-304.
// drho for node i from the ADO array x (PHI id order, column-major blocks).
// Arithmetic order follows the NOBLAS Gemm (blas_wrapper.cc:11-31).  With
// time-local truncation and one tier the base Get applies DoUnitary,
// DoDephasing (mu_0 = 0), DoMatsubara and DoTimelocal, which is also what
// the (never called) NoHierarchy* Get would (hierarchy_updater.h:266-345).
// Under the Shi filter every node keeps its own terms and inactive
// neighbours are skipped.
void NodeRhs(const Oracle* o, const cplx* x, int i, cplx* drho) {
  const int N = o->N;
  const int NE = N * N;
  const Node& nd = o->nodes[i];
  const cplx* rho = x + (size_t)i * NE;
  const bool adaptive = o->filter > 0 && !o->vacuum;
  {
    // DoUnitary: Hemm(Right, -i/hbar, rho, H) -> Gemm(H, rho);
    //            Hemm(Left, +i/hbar, rho, H^dagger, beta=1) -> Gemm(rho, H^dag).
    const cplx alpha(0, -1. / o->hbar), beta(0, 1. / o->hbar);
    for (int a = 0; a < N; ++a) {
      for (int b = 0; b < N; ++b) {
        cplx c = drho[a + b * N] * cplx(0, 0);
        for (int n = 0; n < N; ++n) {
          c += alpha * (o->H[a + n * N] * rho[b * N + n]);
        }
        drho[a + b * N] = c;
      }
    }
    for (int a = 0; a < N; ++a) {
      for (int b = 0; b < N; ++b) {
        cplx c = drho[a + b * N] * cplx(1, 0);
        for (int n = 0; n < N; ++n) {
          c += beta * (rho[a + n * N] * o->Hadj[b * N + n]);
        }
        drho[a + b * N] = c;
      }
    }
    if (o->vacuum && !o->tl_on) return;
    const bool tl = o->tl_on && nd.tier == o->Lh - 1;
    // DoDephasing.
    const cplx mu(nd.dephasing, 0);
    for (int n = 0; n < NE; ++n) drho[n] -= mu * rho[n];
    if (o->bath_model != 0) {
      NodeRhsBath(o, x, nd, rho, drho);
      if (tl) NodeTimeLocal(o, rho, drho);
      return;
    }
    // IndependentBath::DoMatsubara (single bath per state).
    for (int m = 0; m < N; ++m) {
      for (int n = 0; n < N; ++n) {
        drho[m * N + n] -=
            rho[m * N + n] * (o->matsubara[n] + o->matsubara[m]);
      }
      const int d = (N + 1) * m;
      drho[d] += cplx(2, 0) * o->matsubara[m] * rho[d];
    }
    if (tl) NodeTimeLocal(o, rho, drho);
  }
  // DoNext: column m += p rho_next(:, m); row m -= p rho_next(m, :).
  const size_t nn = nd.next_node.size();
  for (size_t j = 0; j < nn; ++j) {
    if (adaptive && !o->nodes[nd.next_node[j]].active) continue;
    const int m = nd.next_site[j];
    const cplx* r = x + (size_t)nd.next_node[j] * NE;
    for (int n = 0; n < N; ++n) drho[N * m + n] += nd.next_pref[j] * r[N * m + n];
  }
  for (int n = 0; n < N; ++n) {
    for (size_t j = 0; j < nn; ++j) {
      if (adaptive && !o->nodes[nd.next_node[j]].active) continue;
      const int m = nd.next_site[j];
      const cplx* r = x + (size_t)nd.next_node[j] * NE;
      drho[N * n + m] -= nd.next_pref[j] * r[N * n + m];
    }
  }
  // DoPrevious: column m -= col rho_prev(:, m); row m -= row rho_prev(m, :).
  const size_t np = nd.prev_node.size();
  for (size_t j = 0; j < np; ++j) {
    if (adaptive && !o->nodes[nd.prev_node[j]].active) continue;
    const int m = nd.prev_site[j];
    const cplx* r = x + (size_t)nd.prev_node[j] * NE;
    for (int n = 0; n < N; ++n) drho[N * m + n] -= nd.prev_col[j] * r[N * m + n];
  }
  for (int n = 0; n < N; ++n) {
    for (size_t j = 0; j < np; ++j) {
      if (adaptive && !o->nodes[nd.prev_node[j]].active) continue;
      const int m = nd.prev_site[j];
      const cplx* r = x + (size_t)nd.prev_node[j] * NE;
      drho[N * n + m] -= nd.prev_row[j] * r[N * n + m];
    }
  }
}

// --- pthread fan-out over contiguous node blocks ---------------------------
struct Job {
  const Oracle* o;
  int lo, hi;
  void (*fn)(const Job*, int);
  const cplx* in;
  cplx* out;
  const cplx* v[8];
  cplx coef[12];
  int nv;
  cplx* out2;
  double red;    // reduction result (max or sum)
  cplx red_c;
};

void* JobEntry(void* p) {
  Job* j = (Job*)p;
  for (int i = j->lo; i < j->hi; ++i) j->fn(j, i);
  return nullptr;
}

void RunParallel(const Oracle* o, int count, Job proto, std::vector<Job>* jobs) {
  const int T = o->threads < 1 ? 1 : o->threads;
  jobs->assign(T, proto);
  std::vector<pthread_t> th(T);
  for (int t = 0; t < T; ++t) {
    (*jobs)[t].lo = (int)((long long)count * t / T);
    (*jobs)[t].hi = (int)((long long)count * (t + 1) / T);
    (*jobs)[t].red = 0;
    (*jobs)[t].red_c = cplx(0, 0);
  }
  if (T == 1) {
    JobEntry(&(*jobs)[0]);
    return;
  }
  for (int t = 0; t < T; ++t) pthread_create(&th[t], nullptr, JobEntry, &(*jobs)[t]);
  for (int t = 0; t < T; ++t) pthread_join(th[t], nullptr);
}

void RhsNodeJob(const Job* j, int i) {
  const size_t NE = (size_t)j->o->N * j->o->N;
  NodeRhs(j->o, j->in, i, j->out + i * NE);
}

void AllRhs(const Oracle* o, const cplx* in, cplx* out) {
  Job proto{};
  proto.o = o;
  proto.fn = RhsNodeJob;
  proto.in = in;
  proto.out = out;
  std::vector<Job> jobs;
  RunParallel(o, (int)o->nodes.size(), proto, &jobs);
}

// out = in0 + sum_t v_t * coef_t, elementwise, left-to-right like PHI's
// non-BLAS update loops (hierarchy_integrator.cc:2019-2024 etc.).
void CombineJob(const Job* j, int i) {
  const size_t NE = (size_t)j->o->N * j->o->N;
  for (size_t e = i * NE; e < (i + 1) * NE; ++e) {
    cplx s = j->in[e];
    for (int t = 0; t < j->nv; ++t) s = s + j->v[t][e] * j->coef[t];
    j->out[e] = s;
  }
}

void Combine(const Oracle* o, const cplx* base, cplx* out, int nv,
             const cplx* const* v, const cplx* coef) {
  Job proto{};
  proto.o = o;
  proto.fn = CombineJob;
  proto.in = base;
  proto.out = out;
  proto.nv = nv;
  for (int t = 0; t < nv; ++t) {
    proto.v[t] = v[t];
    proto.coef[t] = coef[t];
  }
  std::vector<Job> jobs;
  RunParallel(o, (int)o->nodes.size(), proto, &jobs);
}

// RK4 final update: rho += dt/6 * (k0 + 2 k1 + 2 k2 + k3)
// (hierarchy_integrator.cc:2822-2825).
void Rk4UpdateJob(const Job* j, int i) {
  const size_t NE = (size_t)j->o->N * j->o->N;
  const cplx two(2, 0);
  for (size_t e = i * NE; e < (i + 1) * NE; ++e) {
    j->out[e] += j->coef[0] *
                 (j->v[0][e] + two * j->v[1][e] + two * j->v[2][e] + j->v[3][e]);
  }
}

// Filter: max |rho_ij| per node against the tolerance
// (hierarchy_integrator.cc:2828-2851).
void ApplyFilter(Oracle* o) {
  if (!(o->filter > 0)) return;
  const size_t NE = (size_t)o->N * o->N;
  for (size_t i = 0; i < o->nodes.size(); ++i) {
    double mx = 0;
    for (size_t e = i * NE; e < (i + 1) * NE; ++e) {
      const double a = std::abs(o->rho[e]);
      if (a > mx) mx = a;
    }
    Node& nd = o->nodes[i];
    if (mx < o->filter && nd.active) {
      nd.active = false;
    } else if (mx > o->filter && !nd.active) {
      nd.active = true;
    }
  }
}

// RKF45 combination + error (hierarchy_integrator.cc:2317-2328).
void Rkf45ErrJob(const Job* j, int i) {
  const size_t NE = (size_t)j->o->N * j->o->N;
  const cplx* rho = j->in;
  Job* jm = const_cast<Job*>(j);
  for (size_t e = i * NE; e < (i + 1) * NE; ++e) {
    const cplx y4 = rho[e] + j->v[0][e] * j->coef[0] + j->v[2][e] * j->coef[2] +
                    j->v[3][e] * j->coef[3] + j->v[4][e] * j->coef[4];
    const cplx y5 = rho[e] + j->v[0][e] * j->coef[5] + j->v[2][e] * j->coef[6] +
                    j->v[3][e] * j->coef[7] + j->v[4][e] * j->coef[8] +
                    j->v[5][e] * j->coef[1];
    j->out[e] = y5;
    const double d = std::abs(y4 - y5);
    if (d > jm->red) jm->red = d;
  }
}

}  // namespace

extern "C" {

// Plain-data description of one IndependentBath HEOM problem.
struct oracle_params {
  int num_states;          // N
  int num_baths;           // M (must equal N for IndependentBath)
  int num_terms;           // PHI's num_matsubara_terms (MatsubaraTerms + 1)
  int hierarchy_truncation;  // PHI's Lh (tiers 0 .. Lh-1)
  double hbar, kb, temperature;
  const double* gamma;     // M
  const double* lambda;    // M
  const int* bath_site;    // M, site coupled to each bath (default j)
  const double* hamiltonian;  // N*N complex interleaved, column-major
  const double* rho0;      // N*N complex interleaved, column-major
  int normalized;          // rho_normalized
  double filter;           // FilterTolerance (0 = off)
  int threads;
};

void* oracle_create(const oracle_params* p) {
  Oracle* o = new Oracle;
  o->N = p->num_states;
  o->M = p->num_baths;
  o->K = p->num_terms;
  o->D = o->M * o->K;
  o->Lh = p->hierarchy_truncation;
  o->hbar = p->hbar;
  o->kb = p->kb;
  o->temperature = p->temperature;
  o->gamma.assign(p->gamma, p->gamma + o->M);
  o->lambda.assign(p->lambda, p->lambda + o->M);
  o->bath_site.resize(o->M);
  for (int m = 0; m < o->M; ++m) o->bath_site[m] = p->bath_site ? p->bath_site[m] : m;
  o->normalized = p->normalized != 0;
  o->filter = p->filter;
  o->threads = p->threads;
  // GetUpdater, hierarchy_integrator.cc:1741-1744: Lh == 1 without
  // time-local truncation is the Vacuum updater.
  o->vacuum = o->Lh == 1;
  const int N = o->N;
  o->H.resize(N * N);
  o->Hadj.resize(N * N);
  for (int e = 0; e < N * N; ++e) o->H[e] = cplx(p->hamiltonian[2 * e], p->hamiltonian[2 * e + 1]);
  // ConjugateTranspose (complex_matrix.cc:36-44).
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) o->Hadj[i * N + j] = std::conj(o->H[j * N + i]);
  ConstructIndex(o);
  InitializeConstants(o);
  const size_t NE = (size_t)N * N;
  o->rho.assign(o->nodes.size() * NE, cplx(0, 0));
  // InitializeDensityMatrix (hierarchy_integrator.cc:1070-1091): node 0
  // gets rho0 / scale, scale = prod sqrt(|c|^n n!) = 1 at n = 0.
  for (size_t e = 0; e < NE; ++e) o->rho[e] = cplx(p->rho0[2 * e], p->rho0[2 * e + 1]);
  for (int s = 0; s < 6; ++s) o->kbuf[s].assign(o->rho.size(), cplx(0, 0));
  return o;
}

void oracle_destroy(void* h) { delete (Oracle*)h; }
int oracle_num_nodes(void* h) { return (int)((Oracle*)h)->nodes.size(); }
int oracle_num_modes(void* h) { return ((Oracle*)h)->D; }
void oracle_set_threads(void* h, int t) { ((Oracle*)h)->threads = t; }

void oracle_get_index(void* h, int* out) {
  Oracle* o = (Oracle*)h;
  for (size_t i = 0; i < o->nodes.size(); ++i)
    for (int j = 0; j < o->D; ++j) out[i * o->D + j] = o->nodes[i].index[j];
}

// Correlation constants c (D complex), |c| (D) and Delta (M).
void oracle_get_constants(void* h, double* corr, double* abs_corr, double* matsubara) {
  Oracle* o = (Oracle*)h;
  for (int j = 0; j < o->D; ++j) {
    corr[2 * j] = o->corr[j].real();
    corr[2 * j + 1] = o->corr[j].imag();
    abs_corr[j] = o->abs_corr[j];
  }
  for (int j = 0; j < o->M; ++j) matsubara[j] = o->matsubara[j].real();
}

void oracle_get_state(void* h, double* out) {
  Oracle* o = (Oracle*)h;
  std::memcpy(out, (const void*)o->rho.data(), o->rho.size() * sizeof(cplx));
}

void oracle_set_state(void* h, const double* in) {
  Oracle* o = (Oracle*)h;
  std::memcpy((void*)o->rho.data(), in, o->rho.size() * sizeof(cplx));
}

// Coupling model and operators (interleaved complex): model 1 (diagonal)
// M x N entries, model 2 (full) M x N x N column-major.  The adaptive filter
// does not apply to these models (phi_parameters.cc:847-851).
void oracle_set_bath(void* h, int model, const double* ops) {
  Oracle* o = (Oracle*)h;
  o->bath_model = model;
  const size_t n = model == 1 ? (size_t)o->M * o->N : (size_t)o->M * o->N * o->N;
  o->bath_ops.assign(n, cplx(0, 0));
  if (model != 0)
    for (size_t e = 0; e < n; ++e) o->bath_ops[e] = cplx(ops[2 * e], ops[2 * e + 1]);
  if (model != 0) o->filter = 0;
}

// Replaces the system Hamiltonian (N x N, column-major interleaved complex)
// and its adjoint, as UpdateTimeDependentHamiltonian does between stages
// (hierarchy_integrator.cc:1617-1667).
void oracle_set_hamiltonian(void* h, const double* H) {
  Oracle* o = (Oracle*)h;
  const int N = o->N;
  std::memcpy((void*)o->H.data(), H, (size_t)N * N * sizeof(cplx));
  for (int i = 0; i < N; ++i)
    for (int j = 0; j < N; ++j) o->Hadj[i * N + j] = std::conj(o->H[j * N + i]);
}

// Time-local truncation matrices Q_m (M x N x N column-major interleaved
// complex), as UpdateTimeLocalTruncation leaves them between stages
// (hierarchy_integrator.cc:1675-1731); null switches the term off.
void oracle_set_timelocal(void* h, const double* Q) {
  Oracle* o = (Oracle*)h;
  o->tl_on = Q != nullptr;
  const size_t n = (size_t)o->M * o->N * o->N;
  o->tl_Q.assign(n, cplx(0, 0));
  if (Q)
    for (size_t e = 0; e < n; ++e) o->tl_Q[e] = cplx(Q[2 * e], Q[2 * e + 1]);
}

void oracle_get_active(void* h, int* out) {
  Oracle* o = (Oracle*)h;
  for (size_t i = 0; i < o->nodes.size(); ++i) out[i] = o->nodes[i].active ? 1 : 0;
}

// One HierarchyUpdater::Get over every node: out = L(in) (under the Shi
// filter with the current flags).
void oracle_rhs(void* h, const double* in, double* out) {
  Oracle* o = (Oracle*)h;
  AllRhs(o, (const cplx*)in, (cplx*)out);
}

// Fixed-step RK4, hierarchy_integrator.cc:2578-2939.  Integrates from
// t_start while time_now < t_end with time_now += dt (accumulated in double).
// Records (time_now, rho_0) after step i when i % output_factor == 0, as the
// writer thread does, into hist (capacity max_hist).  Returns steps taken.
long long oracle_rk4(void* h, double t_start, double t_end, double dt,
                     double out_min, double* hist_t, double* hist_rho0,
                     int max_hist, int* n_hist) {
  Oracle* o = (Oracle*)h;
  const size_t NE = (size_t)o->N * o->N;
  std::vector<cplx> tmp(o->rho.size());
  const cplx dt2(dt / 2), dt1(dt), dt6(dt / 6);
  const int output_factor = (int)std::ceil(out_min / dt);
  double time_now = t_start;
  long long i = -1;
  int nh = 0;
  while (time_now < t_end) {
    i += 1;
    const cplx* v[4];
    cplx c[4];
    AllRhs(o, o->rho.data(), o->kbuf[0].data());
    v[0] = o->kbuf[0].data(); c[0] = dt2;
    Combine(o, o->rho.data(), tmp.data(), 1, v, c);
    AllRhs(o, tmp.data(), o->kbuf[1].data());
    v[0] = o->kbuf[1].data(); c[0] = dt2;
    Combine(o, o->rho.data(), tmp.data(), 1, v, c);
    AllRhs(o, tmp.data(), o->kbuf[2].data());
    v[0] = o->kbuf[2].data(); c[0] = dt1;
    Combine(o, o->rho.data(), tmp.data(), 1, v, c);
    AllRhs(o, tmp.data(), o->kbuf[3].data());
    {
      Job proto{};
      proto.o = o;
      proto.fn = Rk4UpdateJob;
      proto.out = o->rho.data();
      for (int s = 0; s < 4; ++s) proto.v[s] = o->kbuf[s].data();
      proto.coef[0] = dt6;
      std::vector<Job> jobs;
      RunParallel(o, (int)o->nodes.size(), proto, &jobs);
    }
    ApplyFilter(o);
    time_now += dt;
    if ((i % output_factor) == 0 && nh < max_hist) {
      hist_t[nh] = time_now;
      std::memcpy(hist_rho0 + 2 * NE * nh, o->rho.data(), NE * sizeof(cplx));
      ++nh;
    }
  }
  if (n_hist) *n_hist = nh;
  return i + 1;
}

// RKF45, hierarchy_integrator.cc:1797-2573, with PHI's coefficient algebra:
// the Butcher weights are pre-multiplied by the initial dt and then rescaled
// in place by the same factors as dt (so their rounding follows PHI's), the
// end-clamp factor replaces the growth factor for the rest of the run
// (hierarchy_integrator.cc:2395-2400), and a step is accepted when
// err < tol or dt * 0.5 < mindt.
//
// Output: PHI's RKF45 writer cadence is undefined behaviour (it compares the
// uninitialised wall-clock `now` with simulated time, lines 2500-2509).  This
// oracle records rho_0 after every accepted step whose time_now has reached
// next_write_time, then sets next_write_time = time_now + out_min.
//
// stats[0] = attempted steps, stats[1] = accepted steps, stats[2] = final dt,
// stats[3] = final time.
void oracle_rkf45(void* h, double t_start, double t_end, double dt0, double tol,
                  double mindt, double out_min, double* hist_t,
                  double* hist_rho0, int max_hist, int* n_hist,
                  double* stats) {
  Oracle* o = (Oracle*)h;
  const size_t NE = (size_t)o->N * o->N;
  std::vector<cplx> tmp(o->rho.size());
  double mydt = dt0;
  double dt_f = dt0;
  cplx dtup(1.1), dtdn(0.5);
  const cplx dtc(dt0);
  double a1 = 2. / 9, a2 = 1. / 3, a3 = 3. / 4, a4 = 1.0, a5 = 5. / 6;
  cplx b10(2. / 9.);
  cplx b20(1. / 12.), b21(1. / 4.);
  cplx b30(69. / 128.), b31(-243. / 128.), b32(135. / 64.);
  cplx b40(-17. / 12.), b41(27. / 4.), b42(-27. / 5), b43(16. / 15);
  cplx b50(65. / 432.), b51(-5. / 16), b52(13. / 16.), b53(4. / 27.), b54(5. / 144.);
  cplx c50(1. / 9), c52(9. / 20), c53(16. / 45), c54(1. / 12);
  cplx c60(47. / 450), c62(12. / 25), c63(32. / 225), c64(1. / 30), c65(6. / 25);
  cplx* co[] = {&b10, &b20, &b21, &b30, &b31, &b32, &b40, &b41, &b42, &b43,
                &b50, &b51, &b52, &b53, &b54, &c50, &c52, &c53, &c54,
                &c60, &c62, &c63, &c64, &c65};
  a1 *= mydt; a2 *= mydt; a3 *= mydt; a4 *= mydt; a5 *= mydt;
  for (cplx* c : co) *c *= dtc;
  double time_now = t_start;
  double next_write_time = t_start;
  long long i = -1;
  long long accepted = 0;
  int nh = 0;
  cplx* k[6];
  for (int s = 0; s < 6; ++s) k[s] = o->kbuf[s].data();
  while (time_now < t_end) {
    i += 1;
    const cplx* v[5];
    cplx c[5];
    AllRhs(o, o->rho.data(), k[0]);
    v[0] = k[0]; c[0] = b10;
    Combine(o, o->rho.data(), tmp.data(), 1, v, c);
    AllRhs(o, tmp.data(), k[1]);
    v[0] = k[0]; v[1] = k[1]; c[0] = b20; c[1] = b21;
    Combine(o, o->rho.data(), tmp.data(), 2, v, c);
    AllRhs(o, tmp.data(), k[2]);
    v[2] = k[2]; c[0] = b30; c[1] = b31; c[2] = b32;
    Combine(o, o->rho.data(), tmp.data(), 3, v, c);
    AllRhs(o, tmp.data(), k[3]);
    v[3] = k[3]; c[0] = b40; c[1] = b41; c[2] = b42; c[3] = b43;
    Combine(o, o->rho.data(), tmp.data(), 4, v, c);
    AllRhs(o, tmp.data(), k[4]);
    v[4] = k[4]; c[0] = b50; c[1] = b51; c[2] = b52; c[3] = b53; c[4] = b54;
    Combine(o, o->rho.data(), tmp.data(), 5, v, c);
    AllRhs(o, tmp.data(), k[5]);
    double max_diff = 0;
    {
      Job proto{};
      proto.o = o;
      proto.fn = Rkf45ErrJob;
      proto.in = o->rho.data();
      proto.out = tmp.data();
      for (int s = 0; s < 6; ++s) proto.v[s] = k[s];
      proto.coef[0] = c50; proto.coef[2] = c52; proto.coef[3] = c53; proto.coef[4] = c54;
      proto.coef[5] = c60; proto.coef[6] = c62; proto.coef[7] = c63; proto.coef[8] = c64;
      proto.coef[1] = c65;
      std::vector<Job> jobs;
      RunParallel(o, (int)o->nodes.size(), proto, &jobs);
      for (const Job& j : jobs) max_diff = j.red > max_diff ? j.red : max_diff;
    }
    if (max_diff < tol || dt_f * dtdn.real() < mindt) {
      o->rho.swap(tmp);
      ApplyFilter(o);
      time_now += mydt;
      accepted += 1;
      bool update_dt = max_diff < tol / 10;
      if (time_now + mydt > t_end) {
        update_dt = true;
        dtup = cplx((t_end - time_now) / mydt);
      }
      if (update_dt) {
        mydt *= dtup.real();
        a1 *= dtup.real(); a2 *= dtup.real(); a3 *= dtup.real();
        a4 *= dtup.real(); a5 *= dtup.real();
        for (cplx* cc : co) *cc *= dtup;
      }
      dt_f = mydt;
      if (time_now >= next_write_time && nh < max_hist) {
        hist_t[nh] = time_now;
        std::memcpy(hist_rho0 + 2 * NE * nh, o->rho.data(), NE * sizeof(cplx));
        ++nh;
        next_write_time = time_now + out_min;
      }
    } else {
      mydt *= dtdn.real();
      a1 *= dtdn.real(); a2 *= dtdn.real(); a3 *= dtdn.real();
      a4 *= dtdn.real(); a5 *= dtdn.real();
      for (cplx* cc : co) *cc *= dtdn;
      dt_f = mydt;
    }
  }
  (void)a1; (void)a2; (void)a3; (void)a4; (void)a5;
  if (n_hist) *n_hist = nh;
  stats[0] = (double)(i + 1);
  stats[1] = (double)accepted;
  stats[2] = dt_f;
  stats[3] = time_now;
}

// Steady-state operator (defined here; see header): y = A x where A is the
// hierarchy Liouvillian L with row (node 0, element (0,0)) replaced by
// trace(x_0).
void oracle_steady_matvec(void* h, const double* in, double* out) {
  Oracle* o = (Oracle*)h;
  AllRhs(o, (const cplx*)in, (cplx*)out);
  cplx tr(0, 0);
  const cplx* x = (const cplx*)in;
  for (int a = 0; a < o->N; ++a) tr += x[a * (o->N + 1)];
  ((cplx*)out)[0] = tr;
}

}  // extern "C"
