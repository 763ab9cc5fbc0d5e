// HierarchyIntegrator: PHI's orchestration entry point, run on one MI355X.
//
// Mirrors PHI's class of the same name (hierarchy_integrator.h:25-33):
// construct from a caller-owned PhiParameters, Launch() blocks until the run
// is done and writes PHI's output file, density_matrix() returns a new[]'d
// N*N copy of rho_0 that the caller delete[]s.  Invalid parameters make
// Launch() a no-op that prints "Done." (hierarchy_integrator.cc:168-175).
#ifndef PHI_AMD_INTEGRATOR_H_
#define PHI_AMD_INTEGRATOR_H_

#include <complex>
#include <ctime>
#include <fstream>
#include <string>
#include <vector>

#include "engine.h"
#include "phi_parameters.h"

namespace phi {

// Counters of the last Launch(), readable through the C-ABI.
struct RunInfo {
  long long steps = 0;           // RK4 steps or RKF45 attempts
  long long accepted = 0;        // RKF45 accepted steps
  double final_time = 0;
  double final_dt = 0;
  double seconds = 0;            // wall time of the integration loop
  int bicgstab_iterations = 0;
  double bicgstab_residual = 0;
  int num_nodes = 0;
  int status = 0;
};

class HierarchyIntegrator {
 public:
  explicit HierarchyIntegrator(PhiParameters* p);
  ~HierarchyIntegrator();

  void Launch();
  // rho_0 after Launch; for a truncated system (HilbertSpaceSize >
  // NumStates) in the full space, HilbertSpaceSize^2 entries
  // (hierarchy_integrator.cc:177-191).
  std::complex<double>* density_matrix();
  int density_matrix_dim() const;

  // Additions of this build.
  int status() const { return info_.status; }
  const RunInfo& info() const { return info_; }
  const std::string& error() const { return error_; }
  Engine* engine() { return engine_; }
  // Full ADO state (PHI id order) after Launch().
  int GetState(std::vector<std::complex<double>>* out);

 private:
  int CheckRun();
  int Initialize();
  void WriteHeader();
  int InitialState();
  int RunRk4();
  int RunRkf45();
  int RunSteadyState();
  void WriteRho0Line(const std::complex<double>* rho0, double t);
  void WriteFinal(const std::complex<double>* rho0, double t);
  void FullPrecisionLine(double t, const std::complex<double>* rho0);
  // "#Total number of active matrices = a/n" with the filter on and
  // verbose > 0 (hierarchy_integrator.cc:2540-2543, 2908-2911).
  void ActiveLine(std::ostream* os);
  // Restart files (PHI's text layout, restart_helper.cc:30-55): written at
  // the end of a run and, during it, on PHI's wall-clock cadence.
  int WriteRestart(double t, double dt);
  // Restart output is on with RestartOutputTimeSecs > 0 and a RestartFile
  // (PHI opens an empty name, which fails silently and writes nothing).
  bool RestartOn() const {
    return p_->restart_output_time > 0 && !p_->restart_output_filename.empty();
  }
  int MaybeWriteRestart(double t, double dt, std::time_t* last);
  int ReadRestart(double* t, double* dt);
  void Fail(int code, const std::string& msg);

  PhiParameters* p_;
  Engine* engine_ = nullptr;
  // Time-dependent or truncated system Hamiltonian (null for a plain static
  // H); the engine then runs on ep_, the parameters with the NumStates x
  // NumStates integration Hamiltonian.
  class SystemHamiltonian* sh_ = nullptr;
  PhiParameters ep_;
  std::ofstream out_;
  std::ofstream full_;
  std::vector<std::complex<double>> final_rho0_;
  RunInfo info_;
  std::string error_;
  bool have_final_ = false;
};

// PHI's "#Memory to store matrices" table (hierarchy_integrator.cc:193-311),
// in the same 32-bit float arithmetic.
std::string MemoryTable(const PhiParameters& p, int matrix_count);

}  // namespace phi

#endif  // PHI_AMD_INTEGRATOR_H_
