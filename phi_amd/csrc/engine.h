// Device engine: one hierarchy resident in HBM plus the launch sequences of
// the RK4, RKF45 and steady-state paths.  Single-threaded per instance; all
// work runs on the engine's own HIP stream (default priority).
#ifndef PHI_AMD_ENGINE_H_
#define PHI_AMD_ENGINE_H_

#include <hip/hip_runtime.h>

#include <complex>
#include <string>
#include <vector>

#include "heom_kernels.h"
#include "hierarchy.h"
#include "phi_parameters.h"

namespace phi {

// Status codes of the C-ABI and the engine (0 = success).
enum Status {
  kOk = 0,
  kErrInvalidArgument = -1,
  kErrInvalidParameters = -2,
  kErrUnsupported = -3,
  kErrHip = -4,
  kErrIo = -5,
  kErrNotConverged = -6,
  kErrNoDevice = -7,
};

class Engine {
 public:
  // Builds the hierarchy and uploads every table.  Returns kOk or an error
  // with a message in *why.
  static int Create(const PhiParameters& p, Engine** out, std::string* why);
  ~Engine();

  const HierarchyTables& tables() const { return h_; }
  int num_nodes() const { return h_.num_nodes; }
  int N() const { return h_.N; }
  size_t state_elems() const { return (size_t)h_.num_nodes * h_.N * h_.N; }
  bool fp32() const { return fp32_; }
  void set_fp32(bool v);
  // Stage-kernel shape: col_nm = kernel form (0 element, 1 matrix cores,
  // 2 persistent matrix-core kernels, -1 default); tile = nodes per block
  // (0 = default for the LDS budget); threads = threads per block (0 =
  // 256); blocks = grid of the persistent kernels (0 = all resident).
  int SetLaunch(int col_nm, int tile, int threads, int blocks);
  // RK4 stage order: true = nested (Taylor/Horner) form, false = PHI's
  // classic k0..k3 form.  Both give PHI's step to rounding.
  void set_rk4_nested(bool v) { rk4_nested_ = v; }
  bool rk4_nested() const { return rk4_nested_; }
  const RhsArgs& args() const { return args_; }
  hipStream_t stream() const { return stream_; }

  // Full state in PHI id order, column-major complex blocks.
  int SetState(const std::complex<double>* host);
  int GetState(std::complex<double>* host);
  int GetNode0(std::complex<double>* host);

  // y = L(x) with PHI's Get semantics (any complex input), host pointers.
  int Rhs(const std::complex<double>* x, std::complex<double>* y);
  // Same on device pointers (interleaved doubles).
  int RhsDevice(const double* x, double* y, int trace_row);

  // `nsteps` RK4 steps of size dt.  When snap_every > 0, node 0 is
  // appended to the device history after every step i (counted from
  // step_offset) with i % snap_every == 0.
  int Rk4Steps(long long nsteps, double dt, long long step_offset, int snap_every);
  // As Rk4Steps without snapshots, timing every stage kernel with HIP events
  // on the engine stream; ms[s] += duration of stage s summed over steps.
  int Rk4StageTimes(long long nsteps, double dt, double ms[4]);
  int HistoryCount();
  int DrainHistory(std::vector<std::complex<double>>* out);  // and resets it

  // One RKF45 attempt with the given (dt-scaled) weights; *err = max|y4-y5|.
  // The 5th-order solution is held until Rkf45Accept().
  int Rkf45Attempt(const Rkf45Coefs& c, double* err);
  void Rkf45Accept();

  // Steady state: BiCGSTAB on A x = b (A = L with the trace row), from the
  // current state.  Returns kOk or kErrNotConverged; *iters, *resid report.
  int SteadyState(double tol, int max_iter, int* iters, double* resid,
                  std::vector<double>* history);

  int Synchronize();
  std::string last_error() const { return err_; }

 private:
  Engine() = default;
  int Check(hipError_t e, const char* where);
  int AllocBulk();

  HierarchyTables h_;
  RhsArgs args_{};
  bool fp32_ = false;
  bool rk4_nested_ = true;
  hipStream_t stream_ = nullptr;
  std::string err_;

  // Tables.
  int* d_next_ = nullptr;
  int* d_prev_ = nullptr;
  uint8_t* d_occ_ = nullptr;
  double* d_mu_ = nullptr;
  int* d_site_ptr_ = nullptr;
  int* d_site_modes_ = nullptr;
  double* d_next_s_ = nullptr;
  double2* d_prev_col_ = nullptr;
  double2* d_prev_row_ = nullptr;
  double* d_delta_ = nullptr;
  double2* d_H_ = nullptr;
  double2* d_Hraw_ = nullptr;
  double2* d_Hadj_ = nullptr;
  int4* d_edge_ = nullptr;
  int mps_ = 0;

  // State buffers (interleaved complex).
  double* d_rho_ = nullptr;
  double* d_aux_ = nullptr;       // RK4 acc / RKF45 rho_new
  void* d_y_[2] = {nullptr, nullptr};  // RK4 stage inputs (bulk)
  void* d_k_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // RKF45 k's
  void* d_tmp_ = nullptr;         // RKF45 stage input (bulk)
  unsigned long long* d_err_ = nullptr;
  bool bulk_is_fp32_ = false;
  bool bulk_alloc_ = false;
  bool rkf45_alloc_ = false;

  // Node-0 history.
  double* d_hist_ = nullptr;
  int* d_hist_count_ = nullptr;
  int hist_capacity_ = 0;

  // Reduction scratch.
  double* d_partial_ = nullptr;
  double* d_dot_ = nullptr;
  int red_blocks_ = 0;
};

}  // namespace phi

#endif  // PHI_AMD_ENGINE_H_
