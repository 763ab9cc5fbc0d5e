// Hierarchy construction for the MI355X HEOM integrator.  See hierarchy.h.
#include "hierarchy.h"

#include <cmath>
#include <unordered_map>

namespace phi {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288419716939937510;

// Multi-indices are keyed by their occupation bytes.
struct KeyHash {
  size_t operator()(const std::string& s) const { return std::hash<std::string>()(s); }
};

}  // namespace

double HierarchyTables::MeanNeighbours() const {
  if (num_nodes == 0) return 0;
  size_t edges = 0;
  for (size_t i = 0; i < next.size(); ++i) edges += (next[i] >= 0) + (prev[i] >= 0);
  return double(edges) / num_nodes;
}

std::string BuildHierarchy(const PhiParameters& p, HierarchyTables* h) {
  const int N = p.num_states;
  const int M = p.num_bath_couplings;
  const int K = p.num_matsubara_terms;
  if (N <= 0) return "NumStates must be positive";
  if (M <= 0 || K <= 0) return "at least one bath coupling term is required";
  if ((int)p.gamma.size() < M || (int)p.lambda.size() < M)
    return "Gamma and Lambda must be given for every bath";
  if (p.hierarchy_truncation_level < 1) return "HierarchyTruncation must be positive";
  h->N = N;
  h->M = M;
  h->K = K;
  h->D = M * K;
  h->Lh = p.hierarchy_truncation_level;
  h->L = h->Lh;
  // One tier is PHI's Vacuum updater, unless time-local truncation turns it
  // into the NoHierarchy* updaters (DoUnitary, DoMatsubara, DoTimelocal;
  // hierarchy_integrator.cc:1741-1758).
  h->vacuum = h->Lh == 1 && !p.use_time_local_truncation;
  const int D = h->D;
  const int Lh = h->Lh;
  if (!h->vacuum && M < N && !p.is_diagonal_bath_coupling && !p.is_full_bath_coupling)
    return "IndependentBath needs one bath per state (NumCouplingTerms >= NumStates)";
  if (Lh > 255) return "HierarchyTruncation above 255 is not supported";
  // The mode -> site map reads one index per bath.
  if (!p.multiple_independent_bath_indices.empty() &&
      (int)p.multiple_independent_bath_indices.size() < M)
    return "MultipleIndependentBaths needs one index per bath (NumCouplingTerms entries)";
  // Size bound before anything is enumerated: C(D + Lh - 1, D) nodes (every
  // multi-index of D modes with at most Lh - 1 quanta), each of N^2 complex
  // elements.  Node ids are 32-bit and every state buffer holds at most
  // kMaxStateElems elements.
  {
    double count = 1;
    for (int i = 1; i < Lh && count <= 1e18; ++i) count = count * (D + i) / i;
    if (count > (double)kMaxNodes || count * (double)N * N > (double)kMaxStateElems)
      return "hierarchy too large (" + std::to_string(count) + " nodes of " +
             std::to_string(N) + "x" + std::to_string(N) + "; at most " +
             std::to_string(kMaxNodes) + " nodes and " + std::to_string(kMaxStateElems) +
             " complex elements per state)";
  }

  // --- index graph, PHI id order (hierarchy.cc:45-155) ----------------------
  std::vector<std::string> idx;
  std::unordered_map<std::string, int, KeyHash> lookup;
  idx.emplace_back(std::string(D, '\0'));
  lookup.emplace(idx.back(), 0);
  h->tier_start.assign(1, 0);
  if (Lh >= 2) {
    h->tier_start.push_back(1);
    for (int j = 0; j < D; ++j) {
      std::string v(D, '\0');
      v[j] = 1;
      lookup.emplace(v, (int)idx.size());
      idx.push_back(v);
    }
  }
  int prev_begin = 1, prev_end = (int)idx.size();
  for (int level = 2; level < Lh; ++level) {
    h->tier_start.push_back((int)idx.size());
    const int width = prev_end - prev_begin;
    for (int k = 0; k < D; ++k) {
      for (int n = 0; n < width; ++n) {
        std::string v = idx[prev_begin + n];
        v[(n + k) % D] += 1;
        if (lookup.find(v) == lookup.end()) {
          lookup.emplace(v, (int)idx.size());
          idx.push_back(std::move(v));
        }
      }
    }
    prev_begin = prev_end;
    prev_end = (int)idx.size();
  }
  h->tier_start.push_back((int)idx.size());
  const int nodes = (int)idx.size();
  h->num_nodes = nodes;

  // --- tables ---------------------------------------------------------------
  h->occ.assign((size_t)nodes * D, 0);
  h->next.assign((size_t)nodes * D, -1);
  h->prev.assign((size_t)nodes * D, -1);
  h->tier.assign(nodes, 0);
  for (int i = 0; i < nodes; ++i) {
    int t = 0;
    for (int j = 0; j < D; ++j) {
      h->occ[(size_t)i * D + j] = (uint8_t)idx[i][j];
      t += (uint8_t)idx[i][j];
    }
    h->tier[i] = t;
  }
  // Linking (hierarchy_integrator.cc:1240-1320): n - e_j for every n_j > 0;
  // n + e_j only below the top tier.
  for (int i = 0; i < nodes; ++i) {
    std::string v = idx[i];
    for (int j = 0; j < D; ++j) {
      if (v[j] > 0) {
        v[j] -= 1;
        auto it = lookup.find(v);
        if (it != lookup.end()) h->prev[(size_t)i * D + j] = it->second;
        v[j] += 1;
      }
      if (h->tier[i] < Lh - 1) {
        v[j] += 1;
        auto it = lookup.find(v);
        if (it != lookup.end()) h->next[(size_t)i * D + j] = it->second;
        v[j] -= 1;
      }
    }
  }

  // --- bath constants (hierarchy_integrator.cc:427-473, 1343-1466) ----------
  const double hbar = p.hbar;
  const double thermal = p.boltzmann_constant * p.temperature;
  const double nu = 2 * kPi * thermal / hbar;
  h->corr.assign(D, std::complex<double>(0, 0));
  h->abs_corr.assign(D, 0);
  for (int b = 0; b < M; ++b) {
    const double g = p.gamma[b], l = p.lambda[b];
    if (g > 0) {
      h->corr[b * K] = std::complex<double>(g * l / hbar) *
                       std::complex<double>(1 / std::tan(g * hbar / (2 * thermal)), -1);
      for (int k = 1; k < K; ++k) {
        const double nk = nu * k;
        h->corr[b * K + k] = std::complex<double>(
            4 * l * g * thermal / hbar / hbar * nk / (std::pow(nk, 2) - std::pow(g, 2)));
      }
    }
    for (int k = 0; k < K; ++k) h->abs_corr[b * K + k] = std::abs(h->corr[b * K + k]);
  }
  // Matsubara/high-temperature correction Delta, indexed by state as PHI's
  // IndependentBath::DoMatsubara does (hierarchy_updater.cc:129-144).
  std::vector<double> delta_bath(M, 0);
  for (int b = 0; b < M; ++b) {
    const double g = p.gamma[b], l = p.lambda[b];
    if (g > 0) {
      double d = 2 * l * thermal / hbar / hbar / g - l / hbar / std::tan(hbar * g / (2 * thermal));
      for (int k = 1; k < K; ++k)
        d -= 4 * l * thermal / hbar / hbar * g / (std::pow(nu * k, 2) - std::pow(g, 2));
      delta_bath[b] = d;
    }
  }
  h->delta.assign(N, 0);
  if (!h->vacuum)
    for (int a = 0; a < N && a < M; ++a) h->delta[a] = delta_bath[a];
  h->alpha = delta_bath;

  // Mode -> site map (LinkHierarchyNodes: bath_indices[j / K]).  Diagonal
  // and full coupling operators act through their matrices instead; their
  // modes map to site 0 here and the tables are unused.
  const bool projectors = !p.is_diagonal_bath_coupling && !p.is_full_bath_coupling;
  h->site_of_mode.assign(D, 0);
  for (int j = 0; j < D && projectors; ++j) {
    const int site = p.multiple_independent_bath_indices.empty()
                         ? j / K
                         : p.multiple_independent_bath_indices[j / K];
    if (site < 0 || site >= N) return "bath index outside the system";
    h->site_of_mode[j] = site;
  }
  h->site_ptr.assign(N + 1, 0);
  for (int j = 0; j < D; ++j) h->site_ptr[h->site_of_mode[j] + 1]++;
  for (int a = 0; a < N; ++a) h->site_ptr[a + 1] += h->site_ptr[a];
  h->site_modes.assign(D, 0);
  {
    std::vector<int> fill(h->site_ptr.begin(), h->site_ptr.end() - 1);
    for (int j = 0; j < D; ++j) h->site_modes[fill[h->site_of_mode[j]]++] = j;
  }

  // Dephasing rates: nu_j = gamma_b for the Drude term, 2 pi k kT / hbar for
  // Matsubara term k (hierarchy_integrator.cc:1378-1385).
  std::vector<double> nu_mode(D, 0);
  for (int j = 0; j < D; ++j) nu_mode[j] = (j % K == 0) ? p.gamma[j / K] : nu * (j % K);
  h->mu.assign(nodes, 0);
  for (int i = 0; i < nodes; ++i) {
    double s = 0;
    for (int j = 0; j < D; ++j) s += h->occ[(size_t)i * D + j] * nu_mode[j];
    h->mu[i] = h->vacuum ? 0 : s;
  }

  // Edge prefactors by (mode, occupation) (hierarchy_integrator.cc:1408-1451):
  //   next  = i sqrt((n_j + 1) |c_j|)                (normalized) or i
  //   row   = i n_j c_j / sqrt(n_j |c_j|)            (normalized) or i n_j c_j
  //   col   = -i n_j c_j^* / sqrt(n_j |c_j|)         (normalized) or -i n_j c_j^*
  const int L = h->L;
  h->next_s.assign((size_t)D * L, 0);
  h->prev_col.assign((size_t)D * L, 0);
  h->prev_row.assign((size_t)D * L, 0);
  for (int j = 0; j < D; ++j) {
    for (int k = 0; k < L; ++k) {
      h->next_s[(size_t)j * L + k] =
          p.is_rho_normalized ? std::sqrt((k + 1) * h->abs_corr[j]) : 1.0;
      if (k == 0) continue;
      std::complex<double> row = std::complex<double>(0, k) * h->corr[j];
      std::complex<double> col = std::complex<double>(0, -k) * std::conj(h->corr[j]);
      if (p.is_rho_normalized) {
        const std::complex<double> norm =
            std::complex<double>(1, 0) / std::complex<double>(std::sqrt(k * h->abs_corr[j]));
        row *= norm;
        col *= norm;
      }
      h->prev_row[(size_t)j * L + k] = row;
      h->prev_col[(size_t)j * L + k] = col;
    }
  }
  return "";
}

}  // namespace phi
