// PHI parameter set and `.prm` reader for the MI355X HEOM integrator.
//
// Drop-in for PHI's PhiParameters (phi_parameters.h:32-122): same field names,
// same defaults, same `.prm` grammar and the same error convention (invalid
// input leaves are_parameters_valid == false and run_method == BROKENINPUT;
// nothing throws).  Matrices are stored column-major: element (i, j) of an
// N x N matrix sits at [i + j * N], as in PHI.
#ifndef PHI_AMD_PHI_PARAMETERS_H_
#define PHI_AMD_PHI_PARAMETERS_H_

#include <complex>
#include <iosfwd>
#include <string>
#include <vector>

namespace phi {

typedef double Float;
typedef std::complex<double> Complex;

// Reduced Planck constant in cm^-1 ps and Boltzmann constant in cm^-1 / K
// (phi_parameters.h:13-16).
constexpr double kHbar = 5.3088354;
constexpr double kBoltzmann = 0.6950344;

// Same enumerators and order as PHI's RunMethod (phi_parameters.h:18-30).
enum RunMethod {
  BROKENINPUT,
  NONE,
  RK4,
  RKF45,
  RK4SPECTRUM,
  RKF45SPECTRUM,
  BICGSTABL,
  BICGSTAB,
  BICGSTABU,
  PRINTHIERARCHY,
  PRINTMEMORY
};

// Precision of the bulk derivative evaluation (a key PHI does not know, so
// PHI ignores it in shared `.prm` files): FP64 everywhere, or FP32 RHS with
// FP64 state, accumulator, error norms and inner products.
enum BulkPrecision { kBulkFp64 = 0, kBulkFp32 = 1 };

struct PhiParameters {
  int verbose;
  Float hbar;
  Float boltzmann_constant;

  int num_states;
  int hilbert_space_num_elements;
  int hilbert_space_size;
  int num_bath_couplings;
  int num_matsubara_terms;  // MatsubaraTerms + 1
  int hierarchy_truncation_level;
  int matrix_count;

  std::vector<Complex> time_indepenent_hamiltonian;   // empty == NULL
  std::vector<Complex> final_annealing_hamiltonian;
  std::vector<Complex> initial_annealing_hamiltonian;
  std::vector<Float> initial_hamiltonian_field;
  std::vector<Float> final_hamiltonian_field;
  Float minimum_field_change;

  Float temperature;
  std::vector<Float> kappa;
  std::vector<Float> gamma;
  std::vector<Float> lambda;
  std::vector<Float> bath_coupling_op;
  std::vector<int> multiple_independent_bath_indices;

  std::vector<Complex> initial_density_matrix;
  std::vector<Complex> spectrum_initial_density_matrix;

  Float integration_time;
  Float timestep;
  Float output_minimum_timestep;
  Float start_time;
  Float restart_output_time;
  Float rkf45_minimum_timestep;
  Float rkf45_tolerance;
  Float bicgstab_tolerance;
  Float filter_tolerance;

  std::string output_filename;
  std::string restart_input_filename;
  std::string restart_output_filename;
  std::string restart_backup_filename;

  bool is_truncated_system;
  bool use_time_local_truncation;
  bool is_multiple_independent_baths;
  bool is_spectrum_calculation;
  bool is_diagonal_bath_coupling;
  bool is_full_bath_coupling;
  bool is_rho_normalized;
  bool is_restarted;

  int bicgstab_ell;
  int num_threads;
  bool stupid_hierarchy_partition;
  bool are_parameters_valid;
  RunMethod run_method;

  // Keys of this build (ignored by PHI, which skips unknown keys):
  //   BulkPrecision=fp64|fp32   derivative precision
  //   FullPrecisionOutput=file  rho_0 history at %.17g for parity checks
  //   BicgstabMaxIterations=n   iteration cap of the steady-state solver
  BulkPrecision bulk_precision;
  std::string full_precision_output;
  int bicgstab_max_iterations;

  PhiParameters();
  void SetDefaults();
  void ReadCommandLine(int num_args, const char* const argv[]);
  void ReadInputFile(const std::string& parameter_file);
  void FixIncompatibleCombinations();
  void Write(std::ostream* out, int num_matrices) const;
};

// Number of multi-indices with |n| < max_level over `terms` modes, as PHI
// estimates it in floating point (phi_parameters.cc:1138-1147).
int GetHierarchyNodeCountEstimate(int max_level, int terms);

// Value readers with PHI's input checks (parameter_input.cc).  Exposed for
// the C-ABI so that the boundary's validation can be tested directly.
namespace prm {
bool ValidInteger(const std::string& s, bool* ok);
bool ValidFloat(const std::string& s, bool* ok);
bool ValidBool(const std::string& s, bool* ok);
bool ValidRealList(const std::string& s, bool* ok);
bool ValidComplexList(const std::string& s, bool* ok);
}  // namespace prm

}  // namespace phi

#endif  // PHI_AMD_PHI_PARAMETERS_H_
