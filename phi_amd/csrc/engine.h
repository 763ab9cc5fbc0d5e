// Device engine: one hierarchy resident in HBM plus the launch sequences of
// the RK4, RKF45 and steady-state paths.  Single-threaded per instance; all
// work runs on the engine's own HIP stream (default priority).
#ifndef PHI_AMD_ENGINE_H_
#define PHI_AMD_ENGINE_H_

#include <hip/hip_runtime.h>

#include <array>
#include <complex>
#include <string>
#include <vector>

#include "closure_kernels.h"
#include "heom_kernels.h"
#include "hierarchy.h"
#include "phi_parameters.h"
#include "timelocal.h"

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
  kErrNoMemory = -8,
  kErrInternal = -9,
};

// A time-dependent system Hamiltonian seen by the RK4 step: Update(t) is
// true when H changed at time t (PHI's UpdateTimeDependentHamiltonian);
// H() is then the new N x N Hamiltonian (column-major) and Rotation(), for a
// truncated system, the basis change R of the ADOs (a -> R a R^dagger), else
// null.
class HamiltonianSource {
 public:
  virtual ~HamiltonianSource() {}
  virtual bool Update(double t) = 0;
  virtual const std::complex<double>* H() const = 0;
  virtual const std::complex<double>* Rotation() const = 0;
  // Full coupling operators in the current basis (M x N x N), when they
  // follow the basis of a truncated system; else null.
  virtual const std::complex<double>* BathOps() const { return nullptr; }
};

class Engine {
 public:
  // Builds the hierarchy and uploads every table.  Returns kOk or an error
  // with a message in *why.
  static int Create(const PhiParameters& p, Engine** out, std::string* why);
  ~Engine();

  // An ensemble of hierarchies of one bath model: member m has Hamiltonian
  // H[m*N*N .. ] (column-major) and depth depths[m] (PHI's
  // HierarchyTruncation); the members' nodes are the first nodes of the
  // deepest member's enumeration, stored one member after the other in one
  // slab.  Every member starts from base.initial_density_matrix.  RK4 only.
  static int CreateEnsemble(const PhiParameters& base, int members,
                            const std::complex<double>* H, const int* depths, Engine** out,
                            std::string* why);

  const HierarchyTables& tables() const { return h_; }
  // Nodes of the slab (all members of an ensemble).
  int num_nodes() const { return slab_nodes_; }
  int N() const { return h_.N; }
  size_t state_elems() const { return (size_t)slab_nodes_ * h_.N * h_.N; }
  int members() const { return members_; }
  // Nodes of every member (integrated nodes; without closure slots).
  const std::vector<int>& member_counts() const { return mreal_; }
  const std::vector<int>& member_offsets() const { return moff_; }
  // Node 0 of every member (members * N * N complex).
  int GetMemberRho0(std::complex<double>* host);

  // --- Learned truncation closure (engine_closure.cc) ---------------------------
  // Training samples of the closure network for hierarchies of PHI depth
  // shallow_lh, read from this ensemble's current state (every member at
  // least one tier deeper): for every member and every node n of tier
  // shallow_lh - 1, the network input (ADO of n scaled by its largest
  // magnitude s, and n / (shallow_lh - 1)) and the target (the ADOs of its
  // D next-tier neighbours, scaled by the same s).  Host arrays of
  // *rows = members * (nodes of that tier) rows.
  int ClosureSamples(int shallow_lh, float* X, float* T, long long cap_rows, long long* rows);
  // Couples a trained network to this ensemble (every member of the same
  // depth): each node of the last tier gains D virtual next-tier neighbours
  // whose ADOs the network predicts from the node's stage input before every
  // RK4 stage.  Resets the state to the initial density matrix; RK4 runs in
  // the classic stage order (the closed RHS is not linear).  The network is
  // borrowed and must outlive the engine's use of it.
  int AttachClosure(class Mlp* mlp);
  bool has_closure() const { return closure_ != nullptr; }
  // Closure predictions by the fused pass (default) or by the unfused
  // kernels (inputs, the network's GEMMs, scatter); same arithmetic.
  void set_closure_fused(bool v) {
    closure_fused_ = v;
    DropGraph();
  }
  // Fallback of the closure to plain truncation (ClosureGuard): outputs of
  // magnitude above ymax (or not finite) zero their row's slots.  ymax <= 0
  // leaves the threshold; *fallen = rows that fell back since the attach.
  int ClosureFallback(double ymax, long long* fallen);
  bool fp32() const { return fp32_; }
  void set_fp32(bool v);
  // Stage-kernel shape: col_nm = kernel form (0 element, 1 matrix cores,
  // 2 persistent matrix-core kernels, -1 default); tile = nodes per block
  // (0 = default for the LDS budget); threads = threads per block (0 =
  // 256); blocks = grid of the persistent kernels (0 = all resident).
  int SetLaunch(int col_nm, int tile, int threads, int blocks);
  // Threads and grid of RK4 stage `stage` (0..3) of the persistent kernels
  // alone, in both precisions, with the tile of the common shape (threads 0
  // = one element per thread).  SetLaunch resets every stage to the common
  // shape.
  int SetStageLaunch(int stage, int threads, int blocks);
  // Threads and grid of the four RK4 stages in the current precision.
  void StageLaunch(int threads[4], int blocks[4]) const;
  // RK4 stage order (kRk4Classic, kRk4Nested, kRk4Increment).  The
  // increment form needs the persistent kernels; with the one-shot kernels
  // a step runs nested instead, and the closure always runs classic.
  void set_rk4_form(int f) { rk4_form_ = f; }
  int rk4_form() const { return rk4_form_; }
  const RhsArgs& args() const { return args_; }
  hipStream_t stream() const { return stream_; }

  // Full state in PHI id order, column-major complex blocks.
  int SetState(const std::complex<double>* host);
  int GetState(std::complex<double>* host);
  int GetNode0(std::complex<double>* host);

  // y = L(x) with PHI's Get semantics (any complex input), host pointers.
  int Rhs(const std::complex<double>* x, std::complex<double>* y);
  // Same through caller-provided device staging buffers dx, dy (each
  // state_elems() complex doubles on this engine's device); allocates
  // nothing.
  int RhsStaged(const std::complex<double>* x, std::complex<double>* y, double* dx, double* dy);
  // PHI's per-node Get on the live hierarchy: the RHS of nodes[0..count-1]
  // evaluated on the current device state, count N x N blocks into y.
  int RhsNodes(const int* nodes, int count, std::complex<double>* y);
  // The device this engine's buffers and stream live on.
  int device() const { return device_; }
  // Same on device pointers (interleaved doubles).
  int RhsDevice(const double* x, double* y, int trace_row);

  // `nsteps` RK4 steps of size dt.  When snap_every > 0, node 0 is
  // appended to the device history after every step i (counted from
  // step_offset) with i % snap_every == 0.
  int Rk4Steps(long long nsteps, double dt, long long step_offset, int snap_every);
  // RK4 with a time-dependent Hamiltonian (PHI's classic stage order,
  // hierarchy_integrator.cc:2684-2826): H is updated at t before stage 1,
  // at t + dt/2 before stage 2 and at t + dt before stage 4, each time
  // rotating the live ADO buffers for a truncated system.  t_before[s] is
  // the time at the start of step s (PHI's accumulated time).  Snapshots as
  // Rk4Steps.
  int Rk4StepsTD(long long nsteps, const double* t_before, double dt, long long step_offset,
                 int snap_every, HamiltonianSource* src);
  // Replaces the system Hamiltonian (N x N, column-major) on the device, in
  // stream order.
  int SetHamiltonian(const std::complex<double>* H);
  // Stage-kernel timing: stage s launched nsteps times back to back between
  // two HIP events; ms[s] = their total.  The state is restored.
  int Rk4StageTimes(long long nsteps, double dt, double ms[4]);
  // Rk4Steps between two HIP events on the engine's stream; *ms = elapsed.
  int Rk4StepsTimed(long long nsteps, double dt, double* ms, int snap_every = 0);
  int HistoryCount();
  int history_capacity() const { return hist_capacity_; }
  int DrainHistory(std::vector<std::complex<double>>* out);  // and resets it

  // One RKF45 attempt with the given (dt-scaled) weights; *err = max|y4-y5|.
  // The 5th-order solution is held until Rkf45Accept().  With a
  // time-dependent Hamiltonian (src non-null) H is updated at t before
  // stage 0 and at t + a[s-1] before stage s = 1..5 (a: the dt-scaled
  // stage times), each update rotating the live ADO buffers of a
  // truncated system: rho; then rho, the stage input and k0..k(s-1), k1
  // excepted before stage 5 (hierarchy_integrator.cc:1984-2270).
  int Rkf45Attempt(const Rkf45Coefs& c, double* err, double t = 0.0, const double* a = nullptr,
                   HamiltonianSource* src = nullptr);
  int Rkf45Accept();

  // Adaptive hierarchy truncation (PHI's FilterTolerance > 0 with a
  // hierarchy): RK4 and RKF45 run the filtered operator (the general RHS
  // with inactive neighbours' terms skipped, every node's own terms kept)
  // in FP64 with PHI's classic stage order, and the node flags are updated
  // after every accepted step (filter_kernels.h).  The per-call RHS applies
  // the current flags; the steady state uses the plain operator.
  bool filter_on() const { return filter_ > 0; }
  // Current flags (num_nodes entries, 1 = active); returns the active count.
  int ActiveFlags(int* out);

  // Diagonal / full coupling operators (PHI's DiagonalBath / FullBath,
  // bath_kernels.h): 0 = site projectors (IndependentBath), 1 diagonal, 2
  // full.  RK4 and RKF45 run the explicit-k path in FP64; the operators
  // come from the parameters (real) or from SetBathOps (complex, e.g. a
  // truncated system's operators in its eigenbasis).
  int bath_model() const { return bath_model_; }
  int SetBathOps(const std::complex<double>* ops);
  // Runs that take the explicit-k RK4/RKF45 path (engine_filter.cc).
  bool explicit_path() const { return filter_on() || bath_model_ != 0 || tl_ != nullptr; }

  // Time-local truncation (PHI's TimeLocal; engine_timelocal.cc): the
  // deepest tier gets the time-local term, the matrices advance on PHI's
  // stage schedule inside the explicit-k RK4 step and RKF45 attempt.
  // Null without the option.  The saved matrices are what restart files
  // hold.
  TimeLocal* timelocal() { return tl_; }
  int TlUpload();

  // Steady state: BiCGSTAB on A x = b (A = L with the trace row), from the
  // current state.  Returns kOk or kErrNotConverged; *iters, *resid report.
  int SteadyState(double tol, int max_iter, int* iters, double* resid,
                  std::vector<double>* history);
  // The same system by BiCGSTAB(ell) (ell = 1..8): ell BiCG steps, then a
  // minimal-residual polynomial of degree ell; *iters counts BiCG steps
  // (two matvecs each), history holds |r| after every ell of them.  The
  // scalars stay on the device (BicglScalars); batches of cycles are
  // replayed from a hipGraph when graphs are on.
  int SteadyStateL(int ell, double tol, int max_iter, int* iters, double* resid,
                   std::vector<double>* history);

  int Synchronize();
  std::string last_error() const { return err_; }

  // RK4 blocks replayed from a captured hipGraph (default on; the autotuner
  // picks the faster of graphs and eager launches unless this was called).
  void set_graphs(bool v) {
    use_graphs_ = v;
    graphs_pinned_ = true;
    if (!v) DropGraph();
  }
  bool graphs() const { return use_graphs_; }
  // Steps of the last RK4 call that replayed a captured graph / ran eagerly.
  long long last_graph_steps() const { return last_graph_steps_; }
  long long last_eager_steps() const { return last_eager_steps_; }
  // Autotuner: times `steps` RK4 steps (after one warm-up step) for each
  // candidate shape of the persistent kernels (tile x threads x grid) in
  // the current precision and RK4 form and keeps the fastest; then, with
  // that tile, gives each stage in turn the threads and grid that make
  // whole steps fastest; then (graphs not pinned) times graph blocks
  // against eager launches, and restores the state.
  // Results are cached per (device, problem shape, precision, form) for the
  // process.  *best_ms = milliseconds per step of the winner; shape_out =
  // {form, tile, threads, blocks} of the common shape (StageLaunch gives
  // the stages' own).
  int Autotune(int steps, double dt, double* best_ms, int shape_out[4]);

 private:
  Engine() = default;
  int Check(hipError_t e, const char* where);
  int AllocBulk();
  int Rk4Step(double dt, bool snapshot);
  // The stage order a step runs now, and each stage's input, output and
  // auxiliary buffer for it (inputs alternate so that a stage never
  // overwrites what its own neighbour gathers read).
  int StepForm() const;
  void StageBuffers(int form, const void* in[4], void* out[4], void** aux) const;
  int GraphBlock(int steps, int snap_every, double dt);
  void DropGraph();
  // args_ with RK4 stage s's threads and grid.
  RhsArgs StageArgs(int s) const;
  // Stage shapes {threads, grid} of the current precision, unchecked (the
  // autotuner's own, found valid for the current tile).
  void SetStageShapes(const int threads[4], const int blocks[4]);
  // Whether `threads` (0 = one element per thread) fit the persistent
  // kernels with tile t of precision f; *th = the resolved count.
  int StreamThreads(int t, bool fp32, int threads, int* th);

  struct GraphKey {
    int steps, snap_every;
    double dt;
    bool fp32;
    int form, col_nm, tile0, tile1, th0, th1, grid0, grid1;
    const void* y0;
    const void* rho;
    const void* tiles;
    std::array<int, 16> stage;  // stage_threads_ and stage_grid_
    bool operator==(const GraphKey& o) const {
      return steps == o.steps && snap_every == o.snap_every && dt == o.dt && fp32 == o.fp32 &&
             form == o.form && col_nm == o.col_nm && tile0 == o.tile0 && tile1 == o.tile1 &&
             th0 == o.th0 && th1 == o.th1 && grid0 == o.grid0 && grid1 == o.grid1 &&
             y0 == o.y0 && rho == o.rho && tiles == o.tiles && stage == o.stage;
    }
  };
  static constexpr int kGraphSteps = 32;
  bool use_graphs_ = true;
  bool graphs_pinned_ = false;
  long long last_graph_steps_ = 0, last_eager_steps_ = 0;
  hipGraphExec_t graph_exec_ = nullptr;
  GraphKey graph_key_{};

  HierarchyTables h_;
  RhsArgs args_{};
  // Threads and grid of each RK4 stage of the persistent kernels, per
  // precision ([0] FP64, [1] FP32); the tile is args_.col_tile.
  int stage_threads_[2][4] = {};
  int stage_grid_[2][4] = {};
  bool fp32_ = false;
  int rk4_form_ = kRk4Increment;
  int device_ = 0;
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

  // Slab and ensemble members (members_ == 0: a single hierarchy).  mcount_
  // is each member's slab span (integrated nodes + closure slots), mreal_
  // its integrated nodes.
  int slab_nodes_ = 0;
  int members_ = 0;
  std::vector<int> mcount_, mreal_, moff_;
  std::vector<std::complex<double>> rho0_init_;  // InitialDensityMatrix (N*N or empty)

  // Closure.
  class Mlp* closure_ = nullptr;
  bool closure_fused_ = true;
  FusedNet fused_{};
  float closure_ymax_ = 3.0e38f;
  unsigned long long* d_cfall_ = nullptr;
  int closure_first_ = 0, closure_nb_ = 0, closure_vbase_ = 0;
  float* d_cbmi_ = nullptr;
  int* d_csite_ = nullptr;   // site of each mode (the slot column the stage kernels read)
  int* d_cslot_ = nullptr;   // per packed output: float offset from a row's first slot
  float* d_cx_ = nullptr;
  float* d_cs_ = nullptr;
  float* d_cy_ = nullptr;
  void* d_cpack_ = nullptr;  // the attached network's packed parameters
  int ClosurePredict(const void* in, bool in_fp32);
  int BuildEdges(int closure_first, int closure_nb, int closure_vbase, int4** out);
  int UploadEdges(int closure_first, int closure_nb, int closure_vbase);
  int* d_moff_ = nullptr;
  double* d_mrho_ = nullptr;  // node 0 of every member (GetMemberRho0)
  double2* d_Hm_ = nullptr;
  TileDesc* d_tiles_ = nullptr;
  int tiles_cap_ = 0;
  int tiles_for_ = 0;  // tile size the descriptors were built for
  int BuildTiles();
  // Rebuilds the ensemble tile descriptors when the active tile size
  // changed (precision or launch shape); a no-op for a single hierarchy.
  int EnsureTiles() {
    return members_ && tiles_for_ != args_.col_tile[fp32_ ? 1 : 0] ? BuildTiles() : 0;
  }
  int Unsupported(const char* what);

  // Adaptive truncation.
  double filter_ = 0;
  unsigned char* d_active_ = nullptr;
  double* d_kf_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int bath_model_ = 0;
  double2* d_ops_ = nullptr;
  double* d_alpha_ = nullptr;
  double2* d_corr_ = nullptr;
  std::complex<double>* h_ops_ = nullptr;
  int ExplicitAlloc(int nk);
  int OperatorRhs(const double* x, double* y);
  int ExplicitRhs(const double* in, int s);
  int ExplicitUpdate(HamiltonianSource* src, double t, int nbuf, void* const* bufs);
  int ExplicitRk4Step(double t, double dt, bool snapshot, HamiltonianSource* src);
  int ExplicitRkf45Attempt(const Rkf45Coefs& c, double* err, double t, const double* a,
                           HamiltonianSource* src);

  // Time-local truncation: host matrices, device Q (M N^2), site of each
  // projector bath, pinned staging ring, the current H and coupling
  // operators on the host (the matrices' propagator and seeds), and the
  // stage times of the last RKF45 attempt.
  TimeLocal* tl_ = nullptr;
  double2* d_tlq_ = nullptr;
  int* d_tlsite_ = nullptr;
  std::complex<double>* h_tlstage_ = nullptr;
  int tl_slot_ = 0;
  std::vector<std::complex<double>> h_hcur_, h_opscur_;
  double tl_a_[5] = {0, 0, 0, 0, 0};
  int InitTimeLocal(const PhiParameters& p);
  int TlStep(double ts);
  int TlTerm(const double* x, double* y);

  // Time-dependent Hamiltonian: pinned staging of H (three slots per step,
  // reused after the step's synchronisation) and the device rotation.
  std::complex<double>* h_stage_ = nullptr;
  double2* d_rot_ = nullptr;
  int stage_slot_ = 0;
  int Rotate(const std::complex<double>* R, int nbuf, void* const* bufs, const bool* fp32);

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
