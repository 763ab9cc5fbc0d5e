// Device argument blocks and launchers of the HEOM kernels (heom_kernels.hip).
#ifndef PHI_AMD_HEOM_KERNELS_H_
#define PHI_AMD_HEOM_KERNELS_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace phi {

// Tile of an ensemble slab: slab nodes base .. base + nt - 1 are the local
// nodes local .. local + nt - 1 of `member`, which has `count` nodes.
// 32 bytes, so a block reads one with a single scalar load.
struct __attribute__((aligned(32))) TileDesc {
  int base, local, nt, member, count, pad0, pad1, pad2;
};

// Everything the RHS of one hierarchy needs, as device pointers.
struct RhsArgs {
  int N, NE, D, L, num_nodes, tile, vacuum;
  // Form of the stage kernels: col_nm = 0 element path (H from LDS),
  // 1 matrix-core path (v_mfma 16x16x4, N <= 32), 2 persistent pipelined
  // matrix-core kernels (heom_stream.h; N <= 32, at most 2 modes per
  // site).  Nodes, threads and blocks (0 = as many as fit) per launch, per
  // precision ([0] FP64, [1] FP32).
  int col_nm;
  int col_tile[2];
  int col_threads[2];
  int col_grid[2];
  // Tile of the one-shot kernels when col_nm == 2 (their block is 256
  // threads): RKF45 stages, which have no persistent form.
  int v1_tile[2];
  // Edge records of the persistent kernels: edge[(n*N + b)*mps + q] =
  // {id of n + e_j, id of n - e_j, j*L + n_j, 0} for the q-th mode j of
  // site b ({-1, -1, -1, 0} when absent).
  const int4* edge;
  int mps;
  // Ensembles (persistent kernels only): slab of several members of one
  // bath model sharing the tables of the deepest member.  tiles[t] describes
  // tile t (TileDesc) and H holds every member's shifted H (stride N*N).
  // tiles is nullptr for a single hierarchy.
  const TileDesc* tiles;
  int ntiles;
  int real_h;                 // imag(H) == 0 everywhere
  unsigned long long mN, mNE; // ceil(2^32 / N), ceil(2^32 / N^2)
  double inv_hbar;
  const int* next;          // [num_nodes * D]
  const int* prev;          // [num_nodes * D]
  const uint8_t* occ;       // [num_nodes * D]
  const double* mu;         // [num_nodes]
  const int* site_ptr;      // [N + 1]
  const int* site_modes;    // [D]
  const double* next_s;     // [D * L]
  const double2* prev_col;  // [D * L]
  const double2* prev_row;  // [D * L]
  const double* delta;      // [N]
  const double2* H;         // N*N, shifted by mean(diag H) (Hermitian path)
  const double2* Hraw;      // N*N, as given (general path)
  const double2* Hadj;      // N*N, conjugate transpose of Hraw
};

template <typename TB>
struct Rkf45Bufs {
  TB* k[5];                  // k0..k4 (bulk precision)
  TB* tmp_out;               // next stage input
  double* rho_new;           // y5 (FP64)
  unsigned long long* err;   // max |y4 - y5| as the bits of a double
};

// RKF45 weights already multiplied by the current dt (host keeps PHI's
// in-place rescaling).  b[s][i]: stage-(s+1) input weights; c4: weights of
// k0,k2,k3,k4 in y4; c5: weights of k0,k2,k3,k4,k5 in y5.
struct Rkf45Coefs {
  double b[5][5];
  double c4[4];
  double c5[5];
};

// Dynamic LDS of one block; `a.tile` must already be the launch's tile.
size_t RhsSharedBytes(const RhsArgs& a, bool fp32);
// LDS bytes of a tile of `tile` nodes (with_h: H staged in LDS, i.e. every
// GEMM form except the matrix-core one).
size_t TileLdsBytes(int N, int D, int L, int tile, bool fp32, bool with_h);
// LDS bytes of a block of the persistent kernels (two tiles in flight).
size_t StreamTileLdsBytes(int N, int D, int L, int tile, bool fp32);

// Stage orders of an RK4 step: PHI's classic order (FP64 accumulator), the
// nested (Horner) order and its increment form (heom_stream.h; persistent
// kernels only).
// kRk4ClassicBulk: PHI's classic order with a bulk-precision accumulator and
// a bulk copy of rho (internal: the learned closure's FP32 runs, whose RHS
// is not linear; persistent kernels only).
enum { kRk4Classic = 0, kRk4Nested = 1, kRk4Increment = 2, kRk4ClassicBulk = 3 };
// One RK4 stage.  aux: the classic form's FP64 accumulator, or the
// increment form's K (bulk precision); unused by the nested form.
hipError_t LaunchRk4Stage(const RhsArgs& a, bool fp32, int stage, int form, const void* in,
                          double* rho, void* aux, void* yout, double dt, hipStream_t s);
hipError_t LaunchRkf45Stage(const RhsArgs& a, bool fp32, int stage, const void* in,
                            const double* rho, const Rkf45Bufs<float>& bf,
                            const Rkf45Bufs<double>& bd, const Rkf45Coefs& c, hipStream_t s);
// PHI's Get on any complex input, FP64.  active: the adaptive truncation's
// node flags (inactive neighbours' terms skipped), or null.
hipError_t LaunchRhsGeneral(const RhsArgs& a, const double* x, double* y, int trace_row,
                            hipStream_t s, const unsigned char* active = nullptr);
// Blocks of LaunchRhsGeneralDot (one partial record each), 0 when its LDS
// form does not fit (then it returns hipErrorInvalidValue).
int RhsGeneralDotBlocks(const RhsArgs& a);
// LaunchRhsGeneral (no active flags) fused with the next reductions of the
// steady-state solver, per block: dot 1 = sum conj(w) y (2 doubles),
// dot 3 = sums of conj(y) w, |y|^2, conj(w2) y and conj(w2) w (7 doubles).
hipError_t LaunchRhsGeneralDot(const RhsArgs& a, const double* x, double* y, int trace_row,
                               int dot, const double* w, const double* w2, double* partial,
                               hipStream_t s);
hipError_t LaunchSnapshot(const double* src, double* hist, int* counter, int ne, int capacity,
                          hipStream_t s);
// hist[slot] = mean over the members of their node 0 (fixed summation order).
hipError_t LaunchSnapshotMean(const double* src, const int* member_offsets, int members,
                              double* hist, int* counter, int ne, int capacity, hipStream_t s);
// dst (members x ne complex) = node 0 of every member.
hipError_t LaunchGatherNode0(const double* src, const int* member_offsets, int members, int ne,
                             double* dst, hipStream_t s);
hipError_t LaunchFillU64(unsigned long long* p, unsigned long long v, int n, hipStream_t s);
hipError_t LaunchCastToFloat(const double* src, float* dst, size_t n, hipStream_t s);
// Basis change of every ADO of a slab: a_n <- R a_n R^dagger (R N x N,
// column-major, on the device), in FP64 arithmetic; buf holds FP64 or FP32
// interleaved complex ADOs.
hipError_t LaunchRotateNodes(void* buf, bool fp32, const double2* R, int N, int nodes,
                             hipStream_t s);
hipError_t LaunchDot(const double* u, const double* v, size_t n, int mode, double* partial,
                     int nblocks, double* out, hipStream_t s);
hipError_t LaunchLincomb(double* y, size_t n, int nv, const double* const* x, const double2* c,
                         hipStream_t s);

// BiCGSTAB with its scalars on the device (no host round trip inside an
// iteration).  An iteration is four kernels: v = A p (with the partials of
// rhat . v), the s update (which first closes the previous iteration), t =
// A s (with the partials of t . s, |t|^2, rhat . t, rhat . s) and the
// x/r/p update (with the partials of |r|^2).  The dot products land in this
// struct; the update kernels derive their coefficients from it; the finish
// kernel closes the last iteration of a batch.
struct BicgScalars {
  double2 rho_old, alpha, omega;  // from the previous iteration (1, 1, 1 at start)
  double2 rho_new, rv, ts, tt, rn2;  // dot products of the current iteration
  double2 rho_next;               // rhat . r of the new r (the next rho_new)
  double res;                     // |r| after the last finished iteration
  int done;                       // converged, breakdown (omega = 0) or max_iter
  int it;                         // finished iterations
};
// Reductions are two-pass with a fixed order: a kernel writes one partial
// per block, and the kernel that needs the value sums the partials itself
// (every block the same way), so no single-block reduction kernel runs.
//
// Partials of conj(u) . v (2 doubles per block).
hipError_t LaunchDotPartial(const double* u, const double* v, size_t n, double* partial,
                            int nblocks, hipStream_t s);
// s = r - alpha v for record cur (rv = sum of the nb partials of rhat . v,
// kept in cur, alpha = rho_new / rv); close = true first closes the
// previous iteration (record prev, its nbx |r|^2 partials) into cur, as
// LaunchBicgFinish does.  Nothing once the (closed) record is done.
hipError_t LaunchBicgS(bool close, double* y, size_t n, const double* r, const double* v,
                       const BicgScalars* prev, BicgScalars* cur, const double* rv_part, int nb,
                       const double* xr_part, int nbx, double* hist, double tol, int max_iter,
                       hipStream_t s);
// Partials of conj(t) s, |t|^2, conj(rhat) t, conj(rhat) s (7 doubles per
// block): the matvec's dot 3 on its own.
hipError_t LaunchBicgDots(const double* t, const double* sv, const double* rhat, size_t n,
                          double* partial7, int nblocks, hipStream_t s);
// From those partials: omega = ts / tt (0 if tt = 0), rho_next = rhat . s -
// omega rhat . t (= rhat . r_new), beta; x += alpha p + omega s, r = s -
// omega t, p = r + beta (p - omega v), partials of |r|^2 (1 double per
// block); ts, tt, rho_next kept in st.  Nothing once st->done.
hipError_t LaunchBicgXRP(double* x, double* r, double* p, const double* sv, const double* t,
                         const double* v, size_t n, BicgScalars* st, const double* ts_part,
                         int nb_ts, double* partial, int nblocks, hipStream_t s);
// Closing an iteration: from record prev and the |r|^2 partials, the record
// next of the following iteration (rn2; alpha, omega, rho_old <- this
// iteration's, rho_new <- rho_next; res = |r|; hist[it] = res; done when
// !(res > tol), omega = 0 or it = max_iter).  One block.
hipError_t LaunchBicgFinish(const BicgScalars* prev, BicgScalars* next, const double* xr_part,
                            int nblocks, double* hist, double tol, int max_iter, hipStream_t s);

// BiCGSTAB(ell) with its scalars on the device (Sleijpen and Fokkema 1993,
// Algorithm 3.1).  Dot products land in dot[0..3] through
// LaunchBicglDotStep or LaunchBicglDot2Step, whose reduction kernel then
// runs a scalar step on one thread to derive the scalars from them; the vector
// updates read their coefficient from this record.  stop (breakdown,
// convergence or the step cap) turns every later step and update into a
// no-op, so whole cycles can be enqueued ahead of the host's check.
constexpr int kBicglMaxEll = 8;
struct BicglScalars {
  double2 dot[4];
  double2 rho0, alpha, omega, beta;
  double2 tau[kBicglMaxEll + 1][kBicglMaxEll + 1];
  double2 gp[kBicglMaxEll + 1], g[kBicglMaxEll + 1], gpp[kBicglMaxEll + 1];
  double sigma[kBicglMaxEll + 1];
  double res;   // |r_0| after the last finished cycle
  int stop;     // broke || converged || it >= max_iter
  int broke;
  int it;       // BiCG steps of the finished cycles
  int cycles;   // finished cycles
};
// Scalar steps (one thread):
//   kBlInit   res = sqrt(dot[2]), rho0 = 1, alpha = 0, omega = 1, hist[0];
//             then the first cycle's start
//   kBlBeta   rho1 = dot[0]: beta = alpha rho1 / rho0, rho0 = rho1
//   kBlAlpha  alpha = rho0 / dot[0]
//   kBlTau    tau[i][j] = dot[0] / sigma[i]
//   kBlSigma  sigma[j] = dot[0], gp[j] = dot[1] / sigma[j]; for j = i
//             (= ell) also g, gpp from tau and gp, omega = g[ell]
//   kBlEnd    res = sqrt(dot[2]), it += ell (= i), hist[cycles]; stop; then
//             the next cycle's start
// A cycle's start is rho0 = -omega rho0 and its first BiCG step's beta with
// rho1 = dot[3] = (rtilde, r_0).  A zero rho0 or dot[0] (alpha) or a
// non-positive sigma, or omega = 0 at the end, is a breakdown.
enum { kBlInit = 0, kBlBeta, kBlAlpha, kBlTau, kBlSigma, kBlEnd };
// LaunchDot's reduction of conj(u) . v into S->dot[k] followed, in the same
// one-block kernel, by the scalar step op.
hipError_t LaunchBicglDotStep(const double* u, const double* v, size_t n, double* partial,
                              int nblocks, BicglScalars* S, int k, int op, int i, int j,
                              double tol, int max_iter, double* hist, int hist_cap,
                              hipStream_t s);
// The reduction of nb complex partials (as LaunchRhsGeneralDot's dot 1
// writes them) into S->dot[k], then the scalar step op.
hipError_t LaunchBicglReduceStep(const double* partial, int nb, BicglScalars* S, int k, int op,
                                 int i, int j, double tol, int max_iter, double* hist,
                                 int hist_cap, hipStream_t s);
// Two dot products in one pass, conj(a) . b into S->dot[k] and conj(c) . d
// into S->dot[k + 1] (k = 0..2), then the scalar step op; partial holds 4
// doubles per block (nblocks blocks).
hipError_t LaunchBicglDot2Step(const double* a, const double* b, const double* c,
                               const double* d, size_t n, double* partial, int nblocks,
                               BicglScalars* S, int k, int op, int i, int j, double tol,
                               int max_iter, double* hist, int hist_cap, hipStream_t s);
// Several updates with one coefficient, one pass: y_p = a_p + sign_p *
// (*c) * b_p for p < count (count <= kBicglMaxEll + 1; no y_p may be an
// input of another pair); nothing when *stop.
struct ZPairs {
  double* y[kBicglMaxEll + 1];
  const double* a[kBicglMaxEll + 1];
  const double* b[kBicglMaxEll + 1];
  double sign[kBicglMaxEll + 1];
  int count;
};
hipError_t LaunchZaxpyPairs(const ZPairs& pr, size_t n, const double2* c, const int* stop,
                            hipStream_t s);
// One output from several terms: y = a + sum_t sign_t * (*c_t) * b_t over
// t < count (count <= kBicglMaxEll; y may alias a, no b_t); nothing when
// *stop.
struct ZTerms {
  const double* b[kBicglMaxEll];
  const double2* c[kBicglMaxEll];
  double sign[kBicglMaxEll];
  int count;
};
hipError_t LaunchZcombDev(double* y, const double* a, const ZTerms& t, size_t n, const int* stop,
                          hipStream_t s);
// y = a + sign * (*c) * b elementwise over n complex values (y may alias a
// or b); nothing when *stop.
hipError_t LaunchZaxpyDev(double* y, const double* a, const double* b, size_t n,
                          const double2* c, double sign, const int* stop, hipStream_t s);

}  // namespace phi

#endif  // PHI_AMD_HEOM_KERNELS_H_
