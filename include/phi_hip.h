/*
 * phi_hip.h - C ABI of the MI355X HEOM integrator (drop-in for PHI's hot path).
 *
 * Every entry point returns 0 on success or a negative status (PHI_HIP_E*),
 * never throws (each body runs inside a catch-all that maps exceptions to
 * PHI_HIP_ENOMEM / PHI_HIP_EINTERNAL), takes plain pointers and sizes, and
 * is single-threaded per handle.  Complex arrays are interleaved doubles
 * (re, im); matrices are column-major (element (i, j) at index i + j*N), as
 * in PHI.  ADO state
 * arrays hold every node in PHI's id order (hierarchy.cc:45-155), N*N complex
 * per node.
 *
 * Reference interfaces replaced (file:line in strj/phi):
 *   phi_hip_main                   main / PhiParameters::ReadCommandLine /
 *                                  HierarchyIntegrator::Launch
 *                                  (phi.cc:4-10, phi_parameters.cc:93-146,
 *                                  hierarchy_integrator.cc:113-175)
 *   phi_hip_params_*               struct PhiParameters and its reader
 *                                  (phi_parameters.h:32-122,
 *                                  phi_parameters.cc:152-888, 934-1118)
 *   phi_hip_integrator_*           class HierarchyIntegrator: constructor,
 *                                  Launch(), density_matrix()
 *                                  (hierarchy_integrator.h:25-33,
 *                                  hierarchy_integrator.cc:49-191)
 *   phi_hip_create / _destroy      Hierarchy + AllocateMemory +
 *                                  LinkHierarchyNodes + InitializePrefactors
 *                                  (hierarchy.cc:45-155,
 *                                  hierarchy_integrator.cc:870-1036,
 *                                  1204-1466)
 *   phi_hip_rhs                    HierarchyUpdater::Get over all nodes
 *                                  (hierarchy_updater.h:36-52)
 *   phi_hip_rhs_nodes              HierarchyUpdater::Get of given nodes on
 *                                  the live hierarchy (same lines)
 *   phi_hip_rk4_steps              the body of Rk4Integrate's loop
 *                                  (hierarchy_integrator.cc:2677-2857)
 *   phi_hip_rkf45_attempt/_accept  one attempted step of Rkf45Integrate
 *                                  (hierarchy_integrator.cc:1998-2438)
 *   phi_hip_steady_state           BicgstabSteadyState
 *   phi_hip_steady_state_l         BicgstabLSteadyState
 *                                  (hierarchy_integrator.cc:3112-3489)
 *                                  (hierarchy_integrator.cc:3495-3783)
 *   phi_hip_get_rho0               density_matrix()
 *                                  (hierarchy_integrator.cc:177-191)
 */
#ifndef PHI_HIP_H_
#define PHI_HIP_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PHI_HIP_OK 0
#define PHI_HIP_EINVAL (-1)
#define PHI_HIP_EPARAMS (-2)
#define PHI_HIP_EUNSUPPORTED (-3)
#define PHI_HIP_EHIP (-4)
#define PHI_HIP_EIO (-5)
#define PHI_HIP_ENOTCONVERGED (-6)
#define PHI_HIP_ENODEVICE (-7)
#define PHI_HIP_ENOMEM (-8)    /* host memory exhausted */
#define PHI_HIP_EINTERNAL (-9) /* an unexpected internal error (message says which) */

typedef struct phi_params phi_params;
typedef struct phi_integrator phi_integrator;
typedef struct phi_ctx phi_ctx;
typedef struct phi_mlp phi_mlp;

/* Message of the last failing call on this thread ("" if none). */
const char* phi_hip_last_error(void);
/* ABI version (major * 100 + minor). */
int phi_hip_version(void);
/* Number of visible HIP devices (0 when none or the runtime fails). */
int phi_hip_device_count(void);

/* The `phi <paramfile> <method> [threads]` program. */
int phi_hip_main(int argc, const char* const* argv);

/* ---- Parameters (PhiParameters) ------------------------------------------ */
/* Defaults of PhiParameters::SetDefaults. */
int phi_hip_params_new(phi_params** out);
/* ReadCommandLine semantics: method is the CLI word ("rk4", "rkf45",
 * "bicgstab", ...), threads <= 0 keeps the default.  Invalid input still
 * returns a handle, with are_parameters_valid = 0 and run_method BROKENINPUT
 * where PHI would set them; the call returns PHI_HIP_EPARAMS then. */
int phi_hip_params_read(const char* prm_path, const char* method, int threads,
                        phi_params** out);
void phi_hip_params_free(phi_params* p);
/* Field access by PhiParameters member name (e.g. "temperature",
 * "num_states", "gamma", "time_indepenent_hamiltonian", "run_method"). */
int phi_hip_params_get_double(const phi_params* p, const char* name, double* out);
int phi_hip_params_set_double(phi_params* p, const char* name, double v);
int phi_hip_params_get_int(const phi_params* p, const char* name, int* out);
int phi_hip_params_set_int(phi_params* p, const char* name, int v);
/* Real arrays (gamma, lambda, kappa, bath_coupling_op, fields,
 * multiple_independent_bath_indices): *n receives the length; out may be
 * NULL to query it. */
int phi_hip_params_get_array(const phi_params* p, const char* name, double* out, size_t cap,
                             size_t* n);
int phi_hip_params_set_array(phi_params* p, const char* name, const double* v, size_t n);
/* Complex matrices (time_indepenent_hamiltonian, initial_density_matrix, ...):
 * n complex values, interleaved. */
int phi_hip_params_get_matrix(const phi_params* p, const char* name, double* out, size_t cap,
                              size_t* n);
int phi_hip_params_set_matrix(phi_params* p, const char* name, const double* v, size_t n);
int phi_hip_params_get_string(const phi_params* p, const char* name, char* out, size_t cap);
int phi_hip_params_set_string(phi_params* p, const char* name, const char* v);
/* run_method by CLI word, then FixIncompatibleCombinations. */
int phi_hip_params_fix(phi_params* p, const char* method);
/* The output-file header PHI writes before the dynamics (PhiParameters::Write
 * + "#Total Number of matrices" + memory table + separator) for a hierarchy
 * of num_nodes nodes.  *len receives the full length. */
int phi_hip_params_header(const phi_params* p, int num_nodes, char* out, size_t cap, size_t* len);
/* Value checks of PHI's ParameterInput (kind: "int", "float", "bool",
 * "reals", "complexes"); *valid = 1 when accepted. */
int phi_hip_validate(const char* kind, const char* text, int* valid);

/* ---- HierarchyIntegrator ----------------------------------------------------- */
int phi_hip_integrator_create(phi_params* p, phi_integrator** out);
/* Launch(): runs params' run_method, writes the output file.  Returns the
 * run status (0, or PHI_HIP_E*). */
int phi_hip_integrator_launch(phi_integrator* it);
/* density_matrix(): rho_0 after Launch (N*N complex; for a truncated
 * system, HilbertSpaceSize > NumStates, rho_0 mapped back to the full
 * space: HilbertSpaceSize^2 complex, hierarchy_integrator.cc:177-191). */
int phi_hip_integrator_density_matrix(phi_integrator* it, double* out, size_t cap);
/* Whole ADO state after Launch (num_nodes*N*N complex). */
int phi_hip_integrator_state(phi_integrator* it, double* out, size_t cap);
/* Run counters: out[0] steps/attempts, [1] accepted, [2] final time,
 * [3] final dt, [4] wall seconds, [5] BiCGSTAB iterations, [6] residual,
 * [7] nodes. */
int phi_hip_integrator_info(phi_integrator* it, double* out, size_t cap);
void phi_hip_integrator_destroy(phi_integrator* it);

/* ---- Device hierarchy context (the hot path) ---------------------------------- */
/* precision: 0 = FP64, 1 = FP32 bulk with FP64 state. */
int phi_hip_create(const phi_params* p, int precision, phi_ctx** out);
void phi_hip_destroy(phi_ctx* c);

/* ---- Ensembles of hierarchies (static disorder) ------------------------------- */
/* No counterpart in PHI, which integrates one hierarchy per process
 * (phi.cc:4-10); replaces running PHI once per realisation.  `members`
 * hierarchies of p's bath model, member m with Hamiltonian H[m*N*N ..]
 * (complex, column-major) and depth depths[m] (HierarchyTruncation >= 2),
 * all starting from p's InitialDensityMatrix, in one device slab advanced by
 * one launch per RK4 stage (persistent kernels; N <= 32).  The context
 * takes the RK4 calls below (phi_hip_rk4_steps, phi_hip_rk4_history,
 * phi_hip_autotune, phi_hip_set_launch, ...); num_nodes and the state
 * calls refer to the whole slab. */
int phi_hip_ensemble_create(const phi_params* p, int members, const double* H, const int* depths,
                            int precision, phi_ctx** out);
/* Number of members (0 for a single hierarchy); counts[m] = nodes of member m. */
int phi_hip_ensemble_members(const phi_ctx* c, int* counts, int cap);
/* rho_0 of every member (members * N * N complex; N * N for a single hierarchy). */
int phi_hip_ensemble_rho0(phi_ctx* c, double* out, size_t n_complex);
/* nsteps RK4 steps recording, after every step i with i % every == 0, rho_0
 * (a single hierarchy) or the mean of the members' rho_0 (an ensemble):
 * *count snapshots of N * N complex into out (the output lines of
 * Rk4Integrate, hierarchy_integrator.cc:2859-2880). */
int phi_hip_rk4_history(phi_ctx* c, long long nsteps, double dt, int every, double* out,
                        size_t n_complex, int* count);
/* phi_hip_rk4_history with the steps (snapshots included) between two HIP
 * events on the context's stream; *ms = milliseconds between them.
 * PHI counterpart: PHI times its loop with the host clock (Timer,
 * timer.cc:50-75). */
int phi_hip_rk4_history_timed(phi_ctx* c, long long nsteps, double dt, int every, double* out,
                              size_t n_complex, int* count, double* ms);

/* ---- Learned truncation closure ------------------------------------------------ */
/* No counterpart in PHI, which truncates the hierarchy by dropping the tier
 * beyond HierarchyTruncation (LinkHierarchyNodes, hierarchy_integrator.cc:
 * 1204-1338, links no neighbour there).  A 3-layer perceptron
 * (n_in -> n_hidden -> n_hidden -> n_out, GELU) with bfloat16 GEMM operands
 * and float32 accumulation; parameters are one float32 array W1 (n_hidden x
 * n_in, row-major), b1, W2, b2, W3 (n_out x n_hidden), b3.  For the closure
 * of N states and D bath modes: n_in = 2 N^2 + D, n_out = 2 N^2 D. */
int phi_hip_mlp_create(int n_in, int n_hidden, int n_out, int max_batch, phi_mlp** out);
void phi_hip_mlp_destroy(phi_mlp* m);
long long phi_hip_mlp_param_count(const phi_mlp* m);
int phi_hip_mlp_set_params(phi_mlp* m, const float* flat, size_t n);
int phi_hip_mlp_get_params(phi_mlp* m, float* flat, size_t n);
/* y (batch x n_out) = network(x (batch x n_in)), row-major host arrays. */
int phi_hip_mlp_forward(phi_mlp* m, const float* x, int batch, float* y);
/* *loss = mean((network(x) - t)^2) and its gradient (parameter layout). */
int phi_hip_mlp_gradients(phi_mlp* m, const float* x, const float* t, int batch, double* loss,
                          float* grads, size_t n);
/* Adam (beta 0.9 / 0.999, eps 1e-8) on `samples` rows of X / T, minibatches
 * of `batch` from a seeded shuffle each epoch; epoch_loss[e] = mean loss.
 * The learning rate goes from lr (first epoch) to lr_final (last epoch)
 * geometrically, epoch by epoch; lr_final <= 0 keeps lr throughout. */
int phi_hip_mlp_train(phi_mlp* m, const float* X, const float* T, long long samples, int batch,
                      int epochs, float lr, float lr_final, unsigned seed, double* epoch_loss);
/* Training rows from a context's current state for hierarchies of depth
 * shallow_depth (PHI's HierarchyTruncation; the context is deeper): input
 * = [Re rho_n / s, Im rho_n / s, n / (shallow_depth - 1)], target = the D
 * next-tier ADOs of n / s (each [Re, Im], column-major), for every node n of
 * tier shallow_depth - 1 of every member; s = max |rho_n|. */
int phi_hip_closure_samples(phi_ctx* c, int shallow_depth, float* X, float* T,
                            long long cap_rows, long long* rows);
/* Couple a network to an ensemble context of uniform depth (its last tier
 * gains predicted next-tier neighbours before every RK4 stage); resets the
 * state to the initial density matrix.  Of each predicted neighbour
 * n + e_j only column site(j), the part the stage kernels read, is
 * evaluated and written (2 N D of the 2 N^2 D outputs; the rest of the
 * virtual node is never read).  The network must outlive the context's RK4
 * calls. */
int phi_hip_closure_attach(phi_ctx* c, phi_mlp* m);
/* Closure predictions by one fused launch per stage (on = 1, default) or by
 * the unfused kernels (inputs, the network's GEMMs, scatter); the network's
 * parameters are packed when the closure is attached. */
int phi_hip_closure_fused(phi_ctx* c, int on);
/* The closure's fallback to PHI's plain truncation (the dropped tier of
 * LinkHierarchyNodes): before a stage uses them, the predictions of a
 * boundary node whose evaluated outputs (next-tier entries over the node's
 * largest entry) are not all finite and at most ymax in magnitude are replaced by
 * zeros.  ymax > 0 sets the threshold (default 3e38: non-finite outputs
 * only); *fallen (if not null) = boundary-node predictions replaced since
 * the attach. */
int phi_hip_closure_fallback(phi_ctx* c, double ymax, long long* fallen);
int phi_hip_num_nodes(const phi_ctx* c);
int phi_hip_num_modes(const phi_ctx* c);
/* Multi-indices (num_nodes * D ints), PHI id order. */
int phi_hip_get_index(const phi_ctx* c, int* out, size_t cap);
int phi_hip_set_state(phi_ctx* c, const double* state, size_t n_complex);
int phi_hip_get_state(phi_ctx* c, double* state, size_t n_complex);
int phi_hip_get_rho0(phi_ctx* c, double* out, size_t n_complex);
/* out = L(in) for every node (PHI's Get, any complex input); host arrays.
 * With FilterTolerance > 0 the current node flags apply: the terms of
 * inactive neighbours are skipped and every node keeps its own terms
 * (PHI's base Get with the adaptive DoNext / DoPrevious,
 * hierarchy_updater.h:36-52, hierarchy_updater.cc:256-304).  With
 * TimeLocal the deepest tier includes the time-local term with the current
 * matrices (zero until an integrator run advances them). */
int phi_hip_rhs(phi_ctx* c, const double* in, double* out, size_t n_complex);
/* phi_hip_rhs without any device allocation.  No counterpart in PHI: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The input and output are staged in a
 * fixed per-device array of 32 MiB held until the process ends, so
 * n_complex is limited to 1048576 (PHI_HIP_EINVAL beyond; phi_hip_rhs has
 * no such limit).  Calls that overlap on one device share that array and
 * race: serialise them. */
int phi_hip_rhs_noalloc(phi_ctx* c, const double* in, double* out, size_t n_complex);
/* PHI's per-node HierarchyUpdater::Get(node, drho) (hierarchy_updater.h:36-52):
 * drho[i] = the RHS of node nodes[i] (PHI id order) evaluated on the
 * context's current state, as PHI's Get reads a node's live neighbours;
 * count N x N blocks, column-major, n_complex = count * N^2.  The current
 * flags (FilterTolerance), H and time-local matrices apply, as in
 * phi_hip_rhs.  PHI_HIP_EINVAL for an id outside 0 .. num_nodes - 1. */
int phi_hip_rhs_nodes(phi_ctx* c, const int* nodes, int count, double* drho, size_t n_complex);
/* nsteps RK4 steps of size dt on the device state. */
int phi_hip_rk4_steps(phi_ctx* c, long long nsteps, double dt);
/* phi_hip_rk4_steps bracketed by two HIP events on the context's stream;
 * *ms = milliseconds between them (returns after the second event).  No
 * PHI counterpart: PHI times its loop with the host clock (Timer,
 * timer.cc:50-75). */
int phi_hip_rk4_steps_timed(phi_ctx* c, long long nsteps, double dt, double* ms);
/* Stage-kernel timing: stage s (RK4 order, RK4's buffers) launched nsteps
 * times back to back between two HIP events on the context's stream;
 * ms[s] = total milliseconds of those nsteps launches.  The state is
 * restored afterwards. */
int phi_hip_rk4_stage_times(phi_ctx* c, long long nsteps, double dt, double* ms);
/* One RKF45 attempt.  coefs: 15 b's (b10, b20, b21, ..., b54), 4 c's of y4
 * (c50, c52, c53, c54) and 5 of y5 (c60, c62, c63, c64, c65), all already
 * multiplied by dt.  *err = max |y4 - y5|.  Accept keeps y5.  Time-local
 * truncation needs the stage times and runs through the integrator only
 * (PHI_HIP_EINVAL here). */
int phi_hip_rkf45_attempt(phi_ctx* c, const double* coefs, double* err);
int phi_hip_rkf45_accept(phi_ctx* c);
/* Adaptive hierarchy truncation (FilterTolerance > 0): the activity flag of
 * every node (HierarchyNode::is_active, hierarchy_node.h; updated after each
 * accepted step as hierarchy_integrator.cc:2358-2385, 2828-2851) into
 * flags[0 .. num_nodes-1] (1 = active), *count = active nodes.  Without the
 * filter every node is active.  n = capacity of flags. */
int phi_hip_active_flags(phi_ctx* c, int* flags, int n, int* count);
/* Steady state from the current state: BiCGSTAB on L x = 0 with the
 * (0,0) row of node 0 replaced by trace(rho_0) = 1. */
int phi_hip_steady_state(phi_ctx* c, double tol, int max_iter, int* iters, double* resid);
/* The same system by BiCGSTAB(ell), ell = 1..8 (Sleijpen and Fokkema 1993):
 * ell BiCG steps, then a minimal-residual polynomial of degree ell;
 * *iters counts BiCG steps (two matvecs each); ell = 1 is BiCGSTAB.
 * PHI_HIP_EINVAL for ell outside 1..8 or max_iter < 0 (state untouched),
 * PHI_HIP_ENOTCONVERGED when max_iter BiCG steps leave |r| above tol.
 * PHI counterpart: BicgstabLSteadyState (hierarchy_integrator.cc:
 * 3112-3489), selected by PHI's bicgstab method with ell >= 1. */
int phi_hip_steady_state_l(phi_ctx* c, int ell, double tol, int max_iter, int* iters,
                           double* resid);
int phi_hip_synchronize(phi_ctx* c);
/* Shape of the stage kernels: gemm = 0 element path, 1 matrix cores
 * (v_mfma 16x16x4, N <= 32), 2 persistent pipelined matrix-core kernels
 * (N <= 32, at most two bath modes per site; the default where they
 * apply), -1 default; tile = nodes per block (0 = default); threads =
 * threads per block, a multiple of 64 (one-shot forms: 64..256, 0 = 256;
 * persistent form: up to 1024 with one element per thread, up to 256 with
 * 2..4, 0 = one element per thread); blocks = grid of the persistent
 * kernels (0 = as many as are resident at once).  Every RK4 stage takes
 * this shape.  Results do not depend on the shape beyond rounding.  No
 * counterpart in PHI, whose partitioning is PartitionHierarchy,
 * hierarchy_integrator.cc:636-808. */
int phi_hip_set_launch(phi_ctx* c, int gemm, int tile, int threads, int blocks);
/* Threads and blocks of RK4 stage `stage` (0..3) of the persistent kernels
 * alone (same rules; the tile stays phi_hip_set_launch's), in both
 * precisions; phi_hip_set_launch resets every stage to its shape.  The
 * first and last stage move the FP64 state and can prefer another grid.
 * No counterpart in PHI. */
int phi_hip_set_stage_launch(phi_ctx* c, int stage, int threads, int blocks);
/* threads[4], blocks[4] of the four RK4 stages in the context's precision. */
int phi_hip_stage_launch(const phi_ctx* c, int* threads, int* blocks);
/* Stage order of phi_hip_rk4_steps (the RK4 loop body of Rk4Integrate,
 * hierarchy_integrator.cc:2709-2825): 0 = PHI's k0..k3 form with an
 * accumulator, 1 = nested form rho + dt L (rho + dt/2 L (rho + dt/3 L (rho +
 * dt/4 L rho))), the same polynomial for this time-independent linear RHS,
 * with half the FP64 traffic, 2 = its increment form (default: stages
 * carry K = L rho and the increments dt/3 (K + dt/4 L K), dt/2 (K + L D);
 * the FP64 state is read by the first and last stage only; with the
 * one-shot kernels of large systems a step runs the nested form).  Equal to
 * rounding. */
int phi_hip_set_rk4_form(phi_ctx* c, int form);
/* RK4 step blocks replayed from a captured hipGraph (on = 1, the default)
 * or launched kernel by kernel (0).  Same results either way.  Once called,
 * phi_hip_autotune keeps this choice; otherwise it times both. */
int phi_hip_set_graphs(phi_ctx* c, int on);
/* The current choice (1 = graph blocks, 0 = eager launches). */
int phi_hip_get_graphs(phi_ctx* c, int* on);
/* How the last RK4 call (phi_hip_rk4_steps, _timed, _history) ran: steps
 * replayed from a captured graph and steps launched eagerly.  No
 * counterpart in PHI (no device graphs). */
int phi_hip_launch_counts(const phi_ctx* c, long long* graph_steps, long long* eager_steps);
/* Autotuner of the stage-kernel shape (no counterpart in PHI, whose
 * partitioning is PartitionHierarchy, hierarchy_integrator.cc:636-808):
 * times `steps` RK4 steps of size dt for every candidate shape in the
 * context's precision and RK4 form and keeps the fastest; with its tile,
 * gives each RK4 stage in turn the threads and blocks that make whole
 * steps fastest (phi_hip_set_stage_launch); then (unless
 * phi_hip_set_graphs was called) times 129-step runs with graph blocks and
 * with eager launches and keeps the faster, and leaves the state
 * unchanged.  Cached per device and problem shape for the process.
 * *best_ms = milliseconds per step; shape[4] = {gemm, tile, threads,
 * blocks} of the common shape as in phi_hip_set_launch
 * (phi_hip_stage_launch gives the stages' own). */
int phi_hip_autotune(phi_ctx* c, int steps, double dt, double* best_ms, int* shape);

#ifdef __cplusplus
}
#endif

#endif /* PHI_HIP_H_ */
