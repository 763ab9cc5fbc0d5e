// Hierarchy index graph and per-node coefficient tables of the HEOM.
//
// Node ids follow PHI's enumeration (hierarchy.cc:45-155), so rho_0 is node 0
// and restart files list nodes in the same order.  The host builds
// structure-of-arrays tables that the device kernels read:
//   next[n*D + j] / prev[n*D + j]   id of n + e_j / n - e_j, or -1
//   occ[n*D + j]                    occupation n_j (uint8)
//   mu[n]                           sum_j n_j nu_j (dephasing rate)
//   next_s[j*L + k], prev_col[...], prev_row[...]
//                                   edge prefactors for n_j = k, which depend
//                                   only on (j, n_j)
//   site_ptr/site_modes             CSR list of the modes coupled to each site
#ifndef PHI_AMD_HIERARCHY_H_
#define PHI_AMD_HIERARCHY_H_

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

#include "phi_parameters.h"

namespace phi {

// Size limits of one hierarchy (or ensemble slab): node ids are int32, and
// a state buffer holds at most 2^31 - 1 complex elements (32 GiB in FP64).
constexpr long long kMaxNodes = 1LL << 28;
constexpr long long kMaxStateElems = (1LL << 31) - 1;

struct HierarchyTables {
  int N = 0;        // system size
  int M = 0;        // baths
  int K = 0;        // exponential terms per bath (MatsubaraTerms + 1)
  int D = 0;        // modes = M * K
  int Lh = 0;       // PHI's HierarchyTruncation: tiers 0 .. Lh-1
  int L = 0;        // table stride for occupations (= max(Lh, 1))
  int num_nodes = 0;
  bool vacuum = false;  // Lh == 1: unitary part only (hierarchy_updater.h:251-265)

  std::vector<uint8_t> occ;
  std::vector<int> next, prev;
  std::vector<int> tier;
  std::vector<int> tier_start;   // first id of each tier (ids are tier-major)
  std::vector<double> mu;
  std::vector<double> next_s;
  std::vector<std::complex<double>> prev_col, prev_row;
  std::vector<double> delta;     // per state, PHI's matsubara_prefactor
  std::vector<double> alpha;     // per bath, the same prefactor (general couplings)
  std::vector<int> site_of_mode;
  std::vector<int> site_ptr, site_modes;
  std::vector<std::complex<double>> corr;
  std::vector<double> abs_corr;

  // Average number of existing +/- neighbours per node.
  double MeanNeighbours() const;
};

// Builds the tables from validated parameters.  Returns an empty string on
// success, else a message (and leaves `out` unspecified).
std::string BuildHierarchy(const PhiParameters& p, HierarchyTables* out);

}  // namespace phi

#endif  // PHI_AMD_HIERARCHY_H_
