// Device engine of the MI355X HEOM integrator.  See engine.h.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

#include "filter_kernels.h"

namespace phi {

namespace {

template <typename T>
hipError_t Upload(T** dst, const T* src, size_t n) {
  hipError_t e = hipMalloc((void**)dst, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return e;
  if (n) e = hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

void Free(void* p) {
  if (p) (void)hipFree(p);
}

constexpr int kHistoryCapacity = 4096;

}  // namespace

int Engine::Check(hipError_t e, const char* where) {
  if (e == hipSuccess) return kOk;
  err_ = std::string(where) + ": " + hipGetErrorString(e);
  return kErrHip;
}

int Engine::Create(const PhiParameters& p, Engine** out, std::string* why) {
  *out = nullptr;
  // The host tables first (sizes are bounded there), then the device.
  Engine* e = new Engine;
  std::string msg = BuildHierarchy(p, &e->h_);
  if (!msg.empty()) {
    *why = msg;
    delete e;
    return kErrInvalidParameters;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    *why = "no HIP device is visible";
    delete e;
    return kErrNoDevice;
  }
  const HierarchyTables& h = e->h_;
  const int N = h.N, NE = N * N;
  if ((int)p.time_indepenent_hamiltonian.size() != NE) {
    *why = "a time-independent Hamiltonian is required";
    delete e;
    return kErrUnsupported;
  }
  int rc = e->Check(hipStreamCreateWithFlags(&e->stream_, hipStreamNonBlocking), "stream");
  if (!rc) rc = e->Check(hipGetDevice(&e->device_), "device");
  // H as given, its adjoint, and H shifted by the mean of its diagonal: the
  // shift leaves H rho - rho H^dag unchanged and removes the large common
  // site energy before the bulk GEMM (FP32 cancellation).
  std::vector<double2> Hraw(NE), Hadj(NE), Hs(NE);
  double mean = 0;
  for (int a = 0; a < N; ++a) mean += p.time_indepenent_hamiltonian[(size_t)a * (N + 1)].real();
  mean /= N;
  for (int a = 0; a < N; ++a)
    for (int b = 0; b < N; ++b) {
      const std::complex<double> v = p.time_indepenent_hamiltonian[(size_t)a + (size_t)b * N];
      Hraw[a + b * N] = make_double2(v.real(), v.imag());
      Hs[a + b * N] = make_double2(v.real() - (a == b ? mean : 0.0), v.imag());
      Hadj[b + a * N] = make_double2(v.real(), -v.imag());
    }
  std::vector<double2> pc(h.prev_col.size()), pr(h.prev_row.size());
  for (size_t i = 0; i < pc.size(); ++i) {
    pc[i] = make_double2(h.prev_col[i].real(), h.prev_col[i].imag());
    pr[i] = make_double2(h.prev_row[i].real(), h.prev_row[i].imag());
  }
  if (!rc) rc = e->Check(Upload(&e->d_next_, h.next.data(), h.next.size()), "upload next");
  if (!rc) rc = e->Check(Upload(&e->d_prev_, h.prev.data(), h.prev.size()), "upload prev");
  if (!rc) rc = e->Check(Upload(&e->d_occ_, h.occ.data(), h.occ.size()), "upload occ");
  if (!rc) rc = e->Check(Upload(&e->d_mu_, h.mu.data(), h.mu.size()), "upload mu");
  if (!rc) rc = e->Check(Upload(&e->d_site_ptr_, h.site_ptr.data(), h.site_ptr.size()), "upload site_ptr");
  if (!rc) rc = e->Check(Upload(&e->d_site_modes_, h.site_modes.data(), h.site_modes.size()), "upload site_modes");
  if (!rc) rc = e->Check(Upload(&e->d_next_s_, h.next_s.data(), h.next_s.size()), "upload next_s");
  if (!rc) rc = e->Check(Upload(&e->d_prev_col_, pc.data(), pc.size()), "upload prev_col");
  if (!rc) rc = e->Check(Upload(&e->d_prev_row_, pr.data(), pr.size()), "upload prev_row");
  if (!rc) rc = e->Check(Upload(&e->d_delta_, h.delta.data(), h.delta.size()), "upload delta");
  if (!rc) rc = e->Check(Upload(&e->d_H_, Hs.data(), Hs.size()), "upload H");
  if (!rc) rc = e->Check(Upload(&e->d_Hraw_, Hraw.data(), Hraw.size()), "upload Hraw");
  if (!rc) rc = e->Check(Upload(&e->d_Hadj_, Hadj.data(), Hadj.size()), "upload Hadj");
  // Edge records of the persistent kernels (one per node, site and mode slot
  // of that site), when every site has at most two modes.
  int mps = 0;
  for (int b = 0; b < N; ++b) mps = std::max(mps, h.site_ptr[b + 1] - h.site_ptr[b]);
  if (mps == 0) mps = 1;
  e->mps_ = mps;
  if (!rc && mps <= 2) rc = e->UploadEdges(0, 0, 0);
  e->slab_nodes_ = h.num_nodes;
  e->rho0_init_.assign(p.initial_density_matrix.begin(), p.initial_density_matrix.end());
  const size_t bytes = e->state_elems() * 2 * sizeof(double);
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_rho_, bytes), "alloc rho");
  if (!rc) rc = e->Check(hipMemset(e->d_rho_, 0, bytes), "zero rho");
  e->hist_capacity_ = kHistoryCapacity;
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_hist_, (size_t)kHistoryCapacity * NE * 2 * sizeof(double)), "alloc hist");
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_hist_count_, sizeof(int)), "alloc hist count");
  if (!rc) rc = e->Check(hipMemset(e->d_hist_count_, 0, sizeof(int)), "zero hist count");
  e->red_blocks_ = 1024;
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_partial_, e->red_blocks_ * 2 * sizeof(double)), "alloc partial");
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_dot_, 4 * sizeof(double)), "alloc dot");
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_err_, sizeof(unsigned long long)), "alloc err");
  // Adaptive truncation: every node starts active (hierarchy_node.cc:16).
  // The vacuum updater ignores the flags.
  if (!rc && p.filter_tolerance > 0 && !h.vacuum) {
    e->filter_ = p.filter_tolerance;
    rc = e->Check(hipMalloc((void**)&e->d_active_, h.num_nodes), "alloc flags");
    if (!rc) rc = e->Check(hipMemset(e->d_active_, 1, h.num_nodes), "set flags");
  }
  // Diagonal / full coupling operators (the vacuum updater ignores them).
  if (!rc && (p.is_diagonal_bath_coupling || p.is_full_bath_coupling) && !h.vacuum) {
    e->bath_model_ = p.is_full_bath_coupling ? 2 : 1;
    std::vector<double2> corr(h.D);
    for (int j = 0; j < h.D; ++j) corr[j] = make_double2(h.corr[j].real(), h.corr[j].imag());
    rc = e->Check(Upload(&e->d_alpha_, h.alpha.data(), h.alpha.size()), "upload alpha");
    if (!rc) rc = e->Check(Upload(&e->d_corr_, corr.data(), corr.size()), "upload corr");
    const size_t per = e->bath_model_ == 1 ? (size_t)N : (size_t)NE;
    if (!rc && p.bath_coupling_op.size() == (size_t)h.M * per) {
      std::vector<std::complex<double>> ops(p.bath_coupling_op.begin(), p.bath_coupling_op.end());
      rc = e->SetBathOps(ops.data());
      if (!rc) rc = e->Synchronize();
    }
  }
  e->h_hcur_.assign(p.time_indepenent_hamiltonian.begin(), p.time_indepenent_hamiltonian.end());
  if (!rc && p.use_time_local_truncation) {
    if ((int)e->h_hcur_.size() != NE) {
      e->err_ = "time-local truncation needs the system Hamiltonian";
      rc = kErrInvalidParameters;
    } else {
      rc = e->InitTimeLocal(p);
    }
  }
  if (rc) {
    *why = e->err_;
    delete e;
    return rc;
  }
  RhsArgs& a = e->args_;
  a.N = N;
  a.NE = NE;
  a.D = h.D;
  a.L = h.L;
  a.num_nodes = h.num_nodes;
  // Nodes per 256-thread block: enough to give every thread an element.
  a.tile = std::max(1, std::min(64, 256 / NE));
  a.vacuum = h.vacuum ? 1 : 0;
  a.real_h = 1;
  for (const auto& v : p.time_indepenent_hamiltonian)
    if (v.imag() != 0) a.real_h = 0;
  a.mN = ((1ull << 32) + N - 1) / N;
  a.mNE = ((1ull << 32) + NE - 1) / NE;
  a.inv_hbar = 1.0 / p.hbar;
  a.next = e->d_next_;
  a.prev = e->d_prev_;
  a.occ = e->d_occ_;
  a.mu = e->d_mu_;
  a.site_ptr = e->d_site_ptr_;
  a.site_modes = e->d_site_modes_;
  a.next_s = e->d_next_s_;
  a.prev_col = e->d_prev_col_;
  a.prev_row = e->d_prev_row_;
  a.delta = e->d_delta_;
  a.H = e->d_H_;
  a.Hraw = e->d_Hraw_;
  a.Hadj = e->d_Hadj_;
  a.edge = e->d_edge_;
  a.mps = e->mps_;
  rc = e->SetLaunch(-1, 0, 0, 0);
  if (rc) {
    *why = e->err_;
    delete e;
    return rc;
  }
  e->fp32_ = p.bulk_precision == kBulkFp32;
  *out = e;
  return kOk;
}

// Edge records: node n's q-th mode j of site b -> {n + e_j, n - e_j, j*L + n_j}.
// With a closure, the nodes first .. first + nb - 1 of the last tier point
// their (missing) n + e_j at virtual slot vbase + (n - first)*D + j.  The
// records go to a new device array, *out.
int Engine::BuildEdges(int first, int nb, int vbase, int4** out) {
  const HierarchyTables& h = h_;
  const int N = h.N, mps = mps_;
  std::vector<int4> edge((size_t)h.num_nodes * N * mps);
  for (int n = 0; n < h.num_nodes; ++n)
    for (int b = 0; b < N; ++b)
      for (int q = 0; q < mps; ++q) {
        int4 r = make_int4(-1, -1, -1, 0);
        if (!h.vacuum && q < h.site_ptr[b + 1] - h.site_ptr[b]) {
          const int j = h.site_modes[h.site_ptr[b] + q];
          const size_t o = (size_t)n * h.D + j;
          int nx = h.next[o];
          if (nb > 0 && n >= first && n < first + nb && nx < 0)
            nx = vbase + (n - first) * h.D + j;
          r = make_int4(nx, h.prev[o], j * h.L + h.occ[o], 0);
        }
        edge[((size_t)n * N + b) * mps + q] = r;
      }
  *out = nullptr;
  const int rc = Check(Upload(out, edge.data(), edge.size()), "upload edges");
  if (rc) {
    Free(*out);
    *out = nullptr;
  }
  return rc;
}

int Engine::UploadEdges(int first, int nb, int vbase) {
  int4* e = nullptr;
  const int rc = BuildEdges(first, nb, vbase, &e);
  if (rc) return rc;
  Free(d_edge_);
  d_edge_ = e;
  args_.edge = d_edge_;
  return kOk;
}

int Engine::CreateEnsemble(const PhiParameters& base, int members,
                           const std::complex<double>* H, const int* depths, Engine** out,
                           std::string* why) {
  *out = nullptr;
  const int N = base.num_states, NE = N * N;
  if (members < 1 || !H || !depths || N < 1) {
    *why = "an ensemble needs at least one member, its Hamiltonians and depths";
    return kErrInvalidArgument;
  }
  int lh_max = 0;
  for (int m = 0; m < members; ++m) {
    if (depths[m] < 1) {
      *why = "ensemble depths (HierarchyTruncation) must be >= 1";
      return kErrInvalidArgument;
    }
    lh_max = std::max(lh_max, depths[m]);
  }
  PhiParameters p = base;
  p.hierarchy_truncation_level = lh_max;
  p.time_indepenent_hamiltonian.assign(H, H + NE);
  Engine* e = nullptr;
  int rc = Create(p, &e, why);
  if (rc) return rc;
  if (e->h_.vacuum) {
    *why = "an ensemble needs a hierarchy (HierarchyTruncation >= 2 for some member)";
    delete e;
    return kErrUnsupported;
  }
  if (N > 32 || e->mps_ > 2 || !e->d_edge_) {
    *why = "ensembles run on the persistent kernels (NumStates <= 32, at most two bath "
           "modes per site)";
    delete e;
    return kErrUnsupported;
  }
  if (e->explicit_path()) {
    *why = "FilterTolerance, diagonal/full coupling operators and time-local truncation are "
           "not available for an ensemble";
    delete e;
    return kErrUnsupported;
  }
  // A member of depth 1 would need the vacuum equation (no dephasing).
  for (int m = 0; m < members; ++m)
    if (depths[m] < 2) {
      *why = "ensemble members need HierarchyTruncation >= 2";
      delete e;
      return kErrUnsupported;
    }
  e->members_ = members;
  e->mcount_.resize(members);
  e->moff_.resize(members);
  long long total = 0;
  for (int m = 0; m < members; ++m) {
    e->mcount_[m] = e->h_.tier_start[depths[m]];
    e->moff_[m] = (int)total;
    total += e->mcount_[m];
  }
  e->mreal_ = e->mcount_;
  if (total * NE > (1LL << 31) - 1) {
    *why = "ensemble slab too large (more than 2^31 complex elements)";
    delete e;
    return kErrUnsupported;
  }
  e->slab_nodes_ = (int)total;
  // Every member's H, shifted by its own mean diagonal; all real or not.
  std::vector<double2> Hm((size_t)members * NE);
  int real_h = 1;
  for (int m = 0; m < members; ++m) {
    const std::complex<double>* Hr = H + (size_t)m * NE;
    double mean = 0;
    for (int a = 0; a < N; ++a) mean += Hr[(size_t)a * (N + 1)].real();
    mean /= N;
    for (int i = 0; i < NE; ++i) {
      const bool diag = i % (N + 1) == 0;
      Hm[(size_t)m * NE + i] = make_double2(Hr[i].real() - (diag ? mean : 0.0), Hr[i].imag());
      if (Hr[i].imag() != 0) real_h = 0;
    }
  }
  Free(e->d_rho_);
  e->d_rho_ = nullptr;
  const size_t bytes = e->state_elems() * 2 * sizeof(double);
  rc = e->Check(Upload(&e->d_Hm_, Hm.data(), Hm.size()), "upload member H");
  if (!rc) rc = e->Check(Upload(&e->d_moff_, e->moff_.data(), e->moff_.size()), "upload offsets");
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_rho_, bytes), "alloc ensemble rho");
  if (!rc) rc = e->Check(hipMemset(e->d_rho_, 0, bytes), "zero ensemble rho");
  // rho_0 of every member from the base initial density matrix.
  if (!rc && (int)base.initial_density_matrix.size() == NE) {
    for (int m = 0; m < members && !rc; ++m)
      rc = e->Check(hipMemcpy(e->d_rho_ + (size_t)e->moff_[m] * NE * 2,
                              base.initial_density_matrix.data(), NE * sizeof(double2),
                              hipMemcpyHostToDevice), "initial state");
  }
  if (rc) {
    *why = e->err_;
    delete e;
    return rc;
  }
  RhsArgs& a = e->args_;
  a.num_nodes = e->slab_nodes_;
  a.H = e->d_Hm_;
  a.real_h = real_h;
  rc = e->SetLaunch(2, 0, 0, 0);
  if (rc) {
    *why = e->err_;
    delete e;
    return rc;
  }
  *out = e;
  return kOk;
}

// Tile descriptors of an ensemble for the current tile size: each member's
// nodes cut into tiles of at most `tile` nodes, members in order.
int Engine::BuildTiles() {
  if (members_ == 0) return kOk;
  const int T = args_.col_tile[fp32_ ? 1 : 0];
  std::vector<TileDesc> t;
  for (int m = 0; m < members_; ++m)
    for (int lb = 0; lb < mreal_[m]; lb += T)
      t.push_back(TileDesc{moff_[m] + lb, lb, std::min(T, mreal_[m] - lb), m, mcount_[m], 0, 0, 0});
  int rc = kOk;
  if ((int)t.size() > tiles_cap_) {
    Free(d_tiles_);
    d_tiles_ = nullptr;
    tiles_cap_ = 0;
    rc = Check(hipMalloc((void**)&d_tiles_, t.size() * sizeof(TileDesc)), "alloc tiles");
    if (rc) return rc;
    tiles_cap_ = (int)t.size();
  }
  rc = Check(hipMemcpy(d_tiles_, t.data(), t.size() * sizeof(TileDesc), hipMemcpyHostToDevice),
             "upload tiles");
  if (rc) return rc;
  args_.tiles = d_tiles_;
  args_.ntiles = (int)t.size();
  tiles_for_ = T;
  return kOk;
}

int Engine::GetMemberRho0(std::complex<double>* host) {
  const int NE = N() * N();
  int rc = kOk;
  if (members_ == 0) return GetNode0(host);
  // One gather kernel and one copy (a copy per member costs a round trip each).
  const size_t bytes = (size_t)members_ * NE * sizeof(double2);
  if (!d_mrho_) rc = Check(hipMalloc((void**)&d_mrho_, bytes), "alloc member rho0");
  if (!rc) rc = Check(LaunchGatherNode0(d_rho_, d_moff_, members_, NE, d_mrho_, stream_),
                      "gather member rho0");
  if (!rc) rc = Check(hipMemcpyAsync(host, d_mrho_, bytes, hipMemcpyDeviceToHost, stream_),
                      "member rho0");
  return rc ? rc : Synchronize();
}

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  DropGraph();
  void* ptrs[] = {d_next_, d_prev_, d_occ_, d_mu_, d_site_ptr_, d_site_modes_, d_next_s_,
                  d_prev_col_, d_prev_row_, d_delta_, d_H_, d_Hraw_, d_Hadj_, d_edge_, d_rho_, d_aux_,
                  d_y_[0], d_y_[1], d_k_[0], d_k_[1], d_k_[2], d_k_[3], d_k_[4], d_tmp_,
                  d_err_, d_hist_, d_hist_count_, d_partial_, d_dot_, d_moff_, d_Hm_,
                  d_tiles_, d_cbmi_, d_csite_, d_cslot_, d_cx_, d_cs_, d_cy_, d_cpack_, d_cfall_, d_active_, d_kf_[0], d_kf_[1],
                  d_kf_[2], d_kf_[3], d_kf_[4], d_kf_[5], d_rot_,
                  d_ops_, d_alpha_, d_corr_, d_tlq_, d_tlsite_, d_mrho_};
  for (void* p : ptrs) Free(p);
  if (h_stage_) (void)hipHostFree(h_stage_);
  if (h_ops_) (void)hipHostFree(h_ops_);
  if (h_tlstage_) (void)hipHostFree(h_tlstage_);
  delete tl_;
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::set_fp32(bool v) { fp32_ = v; }

namespace {
// Dynamic LDS a stage-kernel block may take.
constexpr size_t kLdsBudget = 64 * 1024;
}  // namespace

int Engine::StreamThreads(int t, bool fp32, int threads, int* th) {
  const int N = h_.N, NE = N * N;
  const int n = threads > 0 ? threads : std::min(1024, (t * NE + 63) / 64 * 64);
  const int ept = (t * NE + n - 1) / n;
  if (ept > 4 || (ept > 1 && n > 256) ||
      StreamTileLdsBytes(N, h_.D, h_.L, t, fp32) > kLdsBudget) {
    err_ = "tile does not fit the persistent kernels (at most 4 elements per thread, "
           "256 threads when more than one)";
    return kErrInvalidArgument;
  }
  // H's MFMA A-fragments: ceil(N/16) row tiles x ceil(N/4) k-steps x 64 lanes
  const int frags = ((N + 15) / 16) * ((N + 3) / 4) * 64;
  if (members_ && 2 * n < frags) {
    err_ = "an ensemble block needs at least half as many threads as H has MFMA "
           "fragment elements";
    return kErrInvalidArgument;
  }
  *th = n;
  return kOk;
}

int Engine::SetLaunch(int col_nm, int tile, int threads, int blocks) {
  RhsArgs a = args_;
  const int N = h_.N, D = h_.D, L = h_.L, NE = N * N;
  const bool stream_ok = N <= 32 && mps_ <= 2 && d_edge_ != nullptr;
  if (col_nm < 0) {
    // Default: the persistent matrix-core kernels where they apply, else
    // matrix cores whenever H fits two 16-row tiles.
    col_nm = stream_ok ? 2 : (N <= 32 ? 1 : 0);
  }
  if (col_nm < 0 || col_nm > 2) {
    err_ = "kernel form must be 0 (element), 1 (matrix cores) or 2 (persistent)";
    return kErrInvalidArgument;
  }
  if (members_ && col_nm != 2) {
    err_ = "ensembles run on the persistent kernels (form 2) only";
    return kErrInvalidArgument;
  }
  if (col_nm >= 1 && N > 32) {
    err_ = "the matrix-core forms need NumStates <= 32";
    return kErrInvalidArgument;
  }
  if (col_nm == 2 && !stream_ok) {
    err_ = "the persistent form needs at most two bath modes per site";
    return kErrInvalidArgument;
  }
  const int max_threads = col_nm == 2 ? 1024 : 256;
  if (threads != 0 && (threads < 64 || threads > max_threads || threads % 64 != 0)) {
    err_ = col_nm == 2 ? "threads per block must be a multiple of 64 up to 1024"
                       : "threads per block must be 64, 128, 192 or 256";
    return kErrInvalidArgument;
  }
  if (blocks < 0) {
    err_ = "blocks must be >= 0";
    return kErrInvalidArgument;
  }
  const bool with_h = col_nm == 0;
  for (int f = 0; f < 2; ++f) {
    int t, th = threads > 0 ? threads : 256;
    if (col_nm == 2) {
      // Default: about 512 elements per block, one per thread.  With fewer
      // threads than elements a thread takes up to 4 (blocks <= 256 then).
      t = tile > 0 ? tile : std::max(1, 512 / NE);
      while (t > 1 && ((size_t)t * NE > 1024 ||
                       StreamTileLdsBytes(N, D, L, t, f == 1) > kLdsBudget))
        --t;
      const int rt = StreamThreads(t, f == 1, threads, &th);
      if (rt) return rt;
      // Kernels without a persistent form (RKF45) run the one-shot
      // matrix-core kernels with 256 threads and their default tile.
      int t1 = std::max(1, std::min(64, 256 / NE));
      while (t1 > 1 && TileLdsBytes(N, D, L, t1, f == 1, false) > kLdsBudget) --t1;
      a.v1_tile[f] = t1;
    } else {
      // Nodes per block: enough elements for every thread, fewer when the
      // tile's LDS exceeds the budget.
      t = tile > 0 ? tile : std::max(1, std::min(64, th / NE));
      while (t > 1 && TileLdsBytes(N, D, L, t, f == 1, with_h) > kLdsBudget) --t;
      if (TileLdsBytes(N, D, L, t, f == 1, with_h) > kLdsBudget) {
        err_ = "NumStates too large for one node's LDS tile (64 KB budget)";
        return kErrUnsupported;
      }
      a.v1_tile[f] = t;
    }
    a.col_tile[f] = t;
    a.col_threads[f] = th;
    a.col_grid[f] = blocks;
  }
  a.col_nm = col_nm;
  a.tile = a.col_tile[0];
  args_ = a;
  for (int f = 0; f < 2; ++f)
    for (int s = 0; s < 4; ++s) {
      stage_threads_[f][s] = a.col_threads[f];
      stage_grid_[f][s] = blocks;
    }
  return EnsureTiles();
}

int Engine::SetStageLaunch(int stage, int threads, int blocks) {
  if (stage < 0 || stage > 3) {
    err_ = "RK4 stage must be 0..3";
    return kErrInvalidArgument;
  }
  if (args_.col_nm != 2) {
    err_ = "per-stage shapes apply to the persistent kernels (form 2)";
    return kErrInvalidArgument;
  }
  if (threads != 0 && (threads < 64 || threads > 1024 || threads % 64 != 0)) {
    err_ = "threads per block must be a multiple of 64 up to 1024";
    return kErrInvalidArgument;
  }
  if (blocks < 0) {
    err_ = "blocks must be >= 0";
    return kErrInvalidArgument;
  }
  int th[2];
  for (int f = 0; f < 2; ++f) {
    const int rt = StreamThreads(args_.col_tile[f], f == 1, threads, &th[f]);
    if (rt) return rt;
  }
  for (int f = 0; f < 2; ++f) {
    stage_threads_[f][stage] = th[f];
    stage_grid_[f][stage] = blocks;
  }
  return kOk;
}

void Engine::StageLaunch(int threads[4], int blocks[4]) const {
  const int f = fp32_ ? 1 : 0;
  for (int s = 0; s < 4; ++s) {
    threads[s] = args_.col_nm == 2 ? stage_threads_[f][s] : args_.col_threads[f];
    blocks[s] = args_.col_nm == 2 ? stage_grid_[f][s] : args_.col_grid[f];
  }
}

void Engine::SetStageShapes(const int threads[4], const int blocks[4]) {
  const int f = fp32_ ? 1 : 0;
  for (int s = 0; s < 4; ++s) {
    stage_threads_[f][s] = threads[s];
    stage_grid_[f][s] = blocks[s];
  }
}

RhsArgs Engine::StageArgs(int s) const {
  RhsArgs a = args_;
  if (a.col_nm == 2)
    for (int f = 0; f < 2; ++f) {
      a.col_threads[f] = stage_threads_[f][s];
      a.col_grid[f] = stage_grid_[f][s];
    }
  return a;
}

int Engine::Unsupported(const char* what) {
  err_ = std::string(what) + " is not available for an ensemble (RK4 only)";
  return kErrUnsupported;
}

namespace {

struct TuneKey {
  int dev, N, D, L, nodes, mps, members;
  bool fp32;
  int form;
  bool operator<(const TuneKey& o) const {
    return std::tie(dev, N, D, L, nodes, mps, members, fp32, form) <
           std::tie(o.dev, o.N, o.D, o.L, o.nodes, o.mps, o.members, o.fp32, o.form);
  }
};
struct TuneVal {
  int col_nm, tile, threads, blocks;
  int stage_threads[4], stage_blocks[4];  // per RK4 stage (persistent kernels)
  double ms;
  bool graphs;
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

}  // namespace

int Engine::Autotune(int steps, double dt, double* best_ms, int shape_out[4]) {
  if (explicit_path()) {
    err_ = "the autotuner covers the fused stage kernels (not adaptive truncation, "
           "diagonal/full coupling operators or time-local truncation)";
    return kErrUnsupported;
  }
  if (steps < 1) steps = 1;
  int rc = AllocBulk();
  if (rc) return rc;
  const int N = h_.N, NE = N * N;
  const int f = fp32_ ? 1 : 0;
  int dev = 0;
  (void)hipGetDevice(&dev);
  const TuneKey key{dev, N, h_.D, h_.L, slab_nodes_, mps_, members_, fp32_, rk4_form_};
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      const TuneVal& v = it->second;
      rc = SetLaunch(v.col_nm, v.tile, v.threads, v.blocks);
      if (!rc) SetStageShapes(v.stage_threads, v.stage_blocks);
      if (!rc && !graphs_pinned_) {
        use_graphs_ = v.graphs;
        if (!use_graphs_) DropGraph();
      }
      if (!rc) {
        *best_ms = v.ms;
        shape_out[0] = v.col_nm;
        shape_out[1] = v.tile;
        shape_out[2] = v.threads;
        shape_out[3] = v.blocks;
      }
      return rc;
    }
  }
  int cus = 1;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // Candidates: the persistent kernels over tiles of up to 1024 elements,
  // one element per thread or 2..4 per thread with 64..256 threads, and
  // grids of 0 (all resident) or 1..4 blocks per CU; the one-shot forms as
  // a reference point.
  struct Cand { int col_nm, tile, threads, blocks; };
  std::vector<Cand> cands;
  const bool stream_ok = N <= 32 && mps_ <= 2 && d_edge_ != nullptr;
  const int grids[] = {0, cus, 2 * cus, 3 * cus, 4 * cus};
  if (stream_ok) {
    const int tiles[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 48, 64, 96, 128, 192, 256};
    for (int t : tiles) {
      if ((long long)t * NE > 1024) break;
      std::vector<int> ths{0};
      for (int th = 64; th <= 256; th += 64) {
        const int ept = (t * NE + th - 1) / th;
        if (ept >= 2 && ept <= 4) ths.push_back(th);
      }
      for (int th : ths)
        for (int b : grids) cands.push_back({2, t, th, b});
    }
  }
  if (N <= 32) cands.push_back({1, 0, 0, 0});
  cands.push_back({0, 0, 0, 0});
  // Save the state; tuning steps advance it.
  const size_t bytes = state_elems() * 2 * sizeof(double);
  double* saved = nullptr;
  rc = Check(hipMalloc((void**)&saved, bytes), "alloc tune state");
  if (rc) return rc;
  rc = Check(hipMemcpyAsync(saved, d_rho_, bytes, hipMemcpyDeviceToDevice, stream_), "save state");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = Check(hipEventCreate(&e0), "event");
  if (!rc) rc = Check(hipEventCreate(&e1), "event");
  // Milliseconds per step of `steps` eager steps after one untimed step.
  auto time_steps = [&](float* per_step) {
    int r = Rk4Step(dt, false);
    if (!r) r = Check(hipEventRecord(e0, stream_), "record");
    for (int s = 0; s < steps && !r; ++s) r = Rk4Step(dt, false);
    if (!r) r = Check(hipEventRecord(e1, stream_), "record");
    if (!r) r = Check(hipEventSynchronize(e1), "event sync");
    float ms = 0;
    if (!r) r = Check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    *per_step = ms / steps;
    return r;
  };
  const RhsArgs keep = args_;
  int keep_th[2][4], keep_g[2][4];
  std::memcpy(keep_th, stage_threads_, sizeof(keep_th));
  std::memcpy(keep_g, stage_grid_, sizeof(keep_g));
  const bool keep_graphs = use_graphs_;
  TuneVal best{keep.col_nm, keep.col_tile[f], keep.col_threads[f], keep.col_grid[f],
               {}, {}, 1e300, use_graphs_};
  StageLaunch(best.stage_threads, best.stage_blocks);
  for (const Cand& c : cands) {
    if (rc) break;
    if (SetLaunch(c.col_nm, c.tile, c.threads, c.blocks) != kOk) continue;
    float ms = 0;
    rc = time_steps(&ms);
    if (!rc && ms < best.ms) {
      best.col_nm = c.col_nm;
      best.tile = args_.col_tile[f];
      best.threads = args_.col_threads[f];
      best.blocks = args_.col_grid[f];
      StageLaunch(best.stage_threads, best.stage_blocks);
      best.ms = ms;
    }
  }
  // Stages differ in traffic (the first and last move the FP64 state), so
  // each stage in turn takes the threads and grid, with the chosen tile,
  // that make whole steps fastest given the other stages' shapes (a change
  // kept only when it gains more than 0.5%).
  if (!rc && best.col_nm == 2) {
    rc = SetLaunch(best.col_nm, best.tile, best.threads, best.blocks);
    float cur = 0;
    if (!rc) rc = time_steps(&cur);
    for (int s = 0; s < 4 && !rc; ++s) {
      for (int th = 0; th <= 256 && !rc; th += 64)
        for (int b : grids) {
          if (rc || SetStageLaunch(s, th, b) != kOk) continue;
          int rth[4], rb[4];
          StageLaunch(rth, rb);
          if (rth[s] == best.stage_threads[s] && rb[s] == best.stage_blocks[s]) continue;
          float ms = 0;
          rc = time_steps(&ms);
          if (!rc && ms < 0.995f * cur) {
            cur = ms;
            best.stage_threads[s] = rth[s];
            best.stage_blocks[s] = rb[s];
          }
        }
      SetStageShapes(best.stage_threads, best.stage_blocks);
    }
    if (!rc) best.ms = cur;
  }
  // Graph blocks or eager launches for the chosen shape: short stage kernels
  // can run faster eagerly (the graph's node-to-node latency), long ones
  // do not care, tiny ones need the graph when the host cannot keep up.
  if (!rc && !graphs_pinned_) {
    rc = SetLaunch(best.col_nm, best.tile, best.threads, best.blocks);
    if (!rc) SetStageShapes(best.stage_threads, best.stage_blocks);
  }
  if (!rc && !graphs_pinned_) {
    // One eager head step, then four graph blocks (Rk4Steps' layout).
    const long long n = 4LL * kGraphSteps + 1;
    double t[2] = {0, 0};
    for (int g = 0; g < 2 && !rc; ++g) {
      use_graphs_ = g == 1;
      if (!use_graphs_) DropGraph();
      rc = Rk4Steps(n, dt, 0, 0);  // warm-up (and graph capture)
      if (!rc) rc = Check(hipEventRecord(e0, stream_), "record");
      if (!rc) rc = Rk4Steps(n, dt, 0, 0);
      if (!rc) rc = Check(hipEventRecord(e1, stream_), "record");
      if (!rc) rc = Check(hipEventSynchronize(e1), "event sync");
      float ms = 0;
      if (!rc) rc = Check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
      t[g] = ms / n;
    }
    if (!rc) {
      best.graphs = t[1] <= t[0];
      best.ms = best.graphs ? t[1] : t[0];
      use_graphs_ = best.graphs;
      if (!use_graphs_) DropGraph();
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  const int rc_restore = Check(hipMemcpyAsync(d_rho_, saved, bytes, hipMemcpyDeviceToDevice,
                                              stream_), "restore state");
  const int rc_sync = Synchronize();
  Free(saved);
  if (!rc) rc = rc_restore ? rc_restore : rc_sync;
  if (rc) {
    args_ = keep;
    std::memcpy(stage_threads_, keep_th, sizeof(keep_th));
    std::memcpy(stage_grid_, keep_g, sizeof(keep_g));
    use_graphs_ = keep_graphs;
    DropGraph();
    return rc;
  }
  rc = SetLaunch(best.col_nm, best.tile, best.threads, best.blocks);
  if (rc) return rc;
  SetStageShapes(best.stage_threads, best.stage_blocks);
  {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[key] = best;
  }
  *best_ms = best.ms;
  shape_out[0] = best.col_nm;
  shape_out[1] = best.tile;
  shape_out[2] = best.threads;
  shape_out[3] = best.blocks;
  return kOk;
}

int Engine::AllocBulk() {
  const int rt = EnsureTiles();
  if (rt) return rt;
  if (bulk_alloc_ && bulk_is_fp32_ == fp32_) return kOk;
  Free(d_y_[0]);
  Free(d_y_[1]);
  d_y_[0] = d_y_[1] = nullptr;
  for (auto& k : d_k_) { Free(k); k = nullptr; }
  Free(d_tmp_);
  d_tmp_ = nullptr;
  rkf45_alloc_ = false;
  const size_t bytes = state_elems() * 2 * (fp32_ ? sizeof(float) : sizeof(double));
  int rc = Check(hipMalloc(&d_y_[0], bytes), "alloc y0");
  if (!rc) rc = Check(hipMalloc(&d_y_[1], bytes), "alloc y1");
  if (!rc && !d_aux_) rc = Check(hipMalloc((void**)&d_aux_, state_elems() * 2 * sizeof(double)), "alloc aux");
  if (rc) return rc;
  bulk_alloc_ = true;
  bulk_is_fp32_ = fp32_;
  return kOk;
}

int Engine::SetState(const std::complex<double>* host) {
  int rc = Check(hipMemcpyAsync(d_rho_, host, state_elems() * sizeof(std::complex<double>),
                                hipMemcpyHostToDevice, stream_), "set state");
  return rc ? rc : Synchronize();
}

int Engine::GetState(std::complex<double>* host) {
  int rc = Check(hipMemcpyAsync(host, d_rho_, state_elems() * sizeof(std::complex<double>),
                                hipMemcpyDeviceToHost, stream_), "get state");
  return rc ? rc : Synchronize();
}

int Engine::GetNode0(std::complex<double>* host) {
  int rc = Check(hipMemcpyAsync(host, d_rho_, (size_t)N() * N() * sizeof(std::complex<double>),
                                hipMemcpyDeviceToHost, stream_), "get node 0");
  return rc ? rc : Synchronize();
}

int Engine::Synchronize() { return Check(hipStreamSynchronize(stream_), "synchronize"); }

int Engine::RhsDevice(const double* x, double* y, int trace_row) {
  if (members_) return Unsupported("the general RHS");
  return Check(LaunchRhsGeneral(args_, x, y, trace_row, stream_), "rhs_general");
}

int Engine::Rhs(const std::complex<double>* x, std::complex<double>* y) {
  if (members_) return Unsupported("the general RHS");
  const size_t bytes = state_elems() * sizeof(std::complex<double>);
  double *dx = nullptr, *dy = nullptr;
  int rc = Check(hipMalloc((void**)&dx, bytes), "alloc rhs in");
  if (!rc) rc = Check(hipMalloc((void**)&dy, bytes), "alloc rhs out");
  if (!rc) rc = RhsStaged(x, y, dx, dy);
  Free(dx);
  Free(dy);
  return rc;
}

int Engine::RhsStaged(const std::complex<double>* x, std::complex<double>* y, double* dx,
                      double* dy) {
  if (members_) return Unsupported("the general RHS");
  const size_t bytes = state_elems() * sizeof(std::complex<double>);
  int rc = Check(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, stream_), "rhs upload");
  // Under the filter the current flags apply.
  if (!rc) rc = OperatorRhs(dx, dy);
  if (!rc) rc = Check(hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, stream_), "rhs download");
  if (!rc) rc = Synchronize();
  return rc;
}

int Engine::RhsNodes(const int* nodes, int count, std::complex<double>* y) {
  if (members_) return Unsupported("the general RHS");
  const int NE = h_.N * h_.N;
  for (int i = 0; i < count; ++i)
    if (nodes[i] < 0 || nodes[i] >= num_nodes()) {
      err_ = "node id " + std::to_string(nodes[i]) + " outside 0.." +
             std::to_string(num_nodes() - 1);
      return kErrInvalidArgument;
    }
  const size_t bytes = state_elems() * sizeof(std::complex<double>);
  const size_t out_bytes = (size_t)count * NE * sizeof(std::complex<double>);
  double *dy = nullptr, *dout = nullptr;
  int* dids = nullptr;
  int rc = Check(hipMalloc((void**)&dy, bytes), "alloc rhs out");
  if (!rc) rc = Check(hipMalloc((void**)&dout, out_bytes), "alloc node rhs");
  if (!rc) rc = Check(hipMalloc((void**)&dids, count * sizeof(int)), "alloc node ids");
  if (!rc)
    rc = Check(hipMemcpyAsync(dids, nodes, count * sizeof(int), hipMemcpyHostToDevice, stream_),
               "node ids upload");
  // The general operator on the live state (current flags and H, as PHI's
  // Get reads the node's linked neighbours), then the requested blocks.
  if (!rc) rc = OperatorRhs(d_rho_, dy);
  if (!rc) rc = Check(LaunchGatherNode0(dy, dids, count, NE, dout, stream_), "gather nodes");
  if (!rc) rc = Check(hipMemcpyAsync(y, dout, out_bytes, hipMemcpyDeviceToHost, stream_),
                      "node rhs download");
  if (!rc) rc = Synchronize();
  Free(dy);
  Free(dout);
  Free(dids);
  return rc;
}

int Engine::StepForm() const {
  // The closed RHS is not linear: PHI's classic stage order then, with the
  // bulk accumulator when the bulk is FP32 (the stage inputs are FP32 values
  // either way).
  if (closure_) return fp32_ && args_.col_nm == 2 ? kRk4ClassicBulk : kRk4Classic;
  if (rk4_form_ == kRk4Increment && args_.col_nm != 2) return kRk4Nested;
  return rk4_form_;
}

void Engine::StageBuffers(int form, const void* in[4], void* out[4], void** aux) const {
  if (form == kRk4Increment) {
    // K = L rho in y0 (read again by stages 2 and 3), D in y1, D' in aux
    // (FP64-sized, unused by this form otherwise).
    in[0] = d_rho_, in[1] = d_y_[0], in[2] = d_y_[1], in[3] = d_aux_;
    out[0] = d_y_[0], out[1] = d_y_[1], out[2] = d_aux_, out[3] = nullptr;
    *aux = d_y_[0];
    return;
  }
  in[0] = d_rho_, in[1] = d_y_[0], in[2] = d_y_[1], in[3] = d_y_[0];
  out[0] = d_y_[0], out[1] = d_y_[1], out[2] = d_y_[0], out[3] = nullptr;
  *aux = d_aux_;
}

int Engine::Rk4Step(double dt, bool snapshot) {
  const int form = StepForm();
  const void* in[4];
  void* out[4];
  void* aux;
  StageBuffers(form, in, out, &aux);
  int rc = kOk;
  for (int s = 0; s < 4 && !rc; ++s) {
    if (closure_) rc = ClosurePredict(in[s], s > 0 && fp32_);
    if (!rc)
      rc = Check(LaunchRk4Stage(StageArgs(s), fp32_, s, form, in[s], d_rho_, aux, out[s], dt,
                                stream_), "rk4 stage");
  }
  if (!rc && snapshot) {
    // A single hierarchy records rho_0; an ensemble the member average.
    if (members_)
      rc = Check(LaunchSnapshotMean(d_rho_, d_moff_, members_, d_hist_, d_hist_count_, N() * N(),
                                    hist_capacity_, stream_), "snapshot mean");
    else
      rc = Check(LaunchSnapshot(d_rho_, d_hist_, d_hist_count_, N() * N(), hist_capacity_, stream_),
                 "snapshot");
  }
  return rc;
}

void Engine::DropGraph() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  graph_exec_ = nullptr;
}

// A block of `steps` RK4 steps (with node-0 snapshots after steps j with
// j % snap_every == 0) captured once and replayed; re-captured when any
// launch argument changes.  Only kernels are captured.
int Engine::GraphBlock(int steps, int snap_every, double dt) {
  GraphKey key{steps, snap_every, dt, fp32_, StepForm(), args_.col_nm,
               args_.col_tile[0], args_.col_tile[1], args_.col_threads[0], args_.col_threads[1],
               args_.col_grid[0], args_.col_grid[1], d_y_[0], d_rho_, d_tiles_, {}};
  for (int f = 0; f < 2; ++f)
    for (int s = 0; s < 4; ++s) {
      key.stage[f * 8 + s] = stage_threads_[f][s];
      key.stage[f * 8 + 4 + s] = stage_grid_[f][s];
    }
  if (graph_exec_ && key == graph_key_) return kOk;
  DropGraph();
  // The persistent kernels query their occupancy on first launch; callers
  // run at least one eager step with the same shape before capturing
  // (Rk4Steps does), so the capture holds only kernel launches.
  hipGraph_t g = nullptr;
  int rc = Check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
  if (rc) return rc;
  for (int j = 0; j < steps && !rc; ++j) rc = Rk4Step(dt, snap_every > 0 && j % snap_every == 0);
  const int rc_end = Check(hipStreamEndCapture(stream_, &g), "end capture");
  if (!rc) rc = rc_end;
  if (!rc) rc = Check(hipGraphInstantiate(&graph_exec_, g, nullptr, nullptr, 0), "instantiate");
  if (g) (void)hipGraphDestroy(g);
  if (rc) {
    graph_exec_ = nullptr;
    return rc;
  }
  graph_key_ = key;
  return kOk;
}

int Engine::Rk4Steps(long long nsteps, double dt, long long step_offset, int snap_every) {
  last_graph_steps_ = last_eager_steps_ = 0;
  int rc = AllocBulk();
  if (rc) return rc;
  auto snap = [&](long long s) { return snap_every > 0 && (step_offset + s) % snap_every == 0; };
  if (explicit_path()) {
    for (long long s = 0; s < nsteps && !rc; ++s) rc = ExplicitRk4Step(0.0, dt, snap(s), nullptr);
    last_eager_steps_ = nsteps;
    return rc;
  }
  // Graph blocks of G steps, G a multiple of the snapshot period, aligned so
  // that every block starts on a snapshot step.
  int G = 0;
  if (use_graphs_) {
    G = kGraphSteps;
    if (snap_every > 0) G = std::max(1, kGraphSteps / snap_every) * snap_every;
  }
  long long s = 0;
  // Eager head: at least one step (settles launch parameters), then up to
  // the first aligned step.
  while (s < nsteps && (s == 0 || (G > 0 && snap_every > 0 && (step_offset + s) % G != 0))) {
    rc = Rk4Step(dt, snap(s));
    if (rc) return rc;
    ++s;
  }
  if (G > 0 && nsteps - s >= 2LL * G) {
    rc = GraphBlock(G, snap_every, dt);
    if (rc) return rc;
    for (; s + G <= nsteps; s += G) {
      rc = Check(hipGraphLaunch(graph_exec_, stream_), "graph launch");
      if (rc) return rc;
      last_graph_steps_ += G;
    }
  }
  for (; s < nsteps; ++s) {
    rc = Rk4Step(dt, snap(s));
    if (rc) return rc;
  }
  last_eager_steps_ = nsteps - last_graph_steps_;
  return kOk;
}

int Engine::Rk4StepsTimed(long long nsteps, double dt, double* ms, int snap_every) {
  *ms = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = Check(hipEventCreate(&e0), "event");
  if (!rc) rc = Check(hipEventCreate(&e1), "event");
  if (!rc) rc = Check(hipEventRecord(e0, stream_), "record");
  if (!rc) rc = Rk4Steps(nsteps, dt, 0, snap_every);
  if (!rc) rc = Check(hipEventRecord(e1, stream_), "record");
  if (!rc) rc = Check(hipEventSynchronize(e1), "event sync");
  float t = 0;
  if (!rc) rc = Check(hipEventElapsedTime(&t, e0, e1), "elapsed");
  if (!rc) *ms = t;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

int Engine::Rk4StageTimes(long long nsteps, double dt, double ms[4]) {
  if (explicit_path()) {
    err_ = "stage timing covers the fused stage kernels (not adaptive truncation, "
           "diagonal/full coupling operators or time-local truncation)";
    return kErrUnsupported;
  }
  int rc = AllocBulk();
  if (rc) return rc;
  for (int s = 0; s < 4; ++s) ms[s] = 0;
  if (nsteps < 1) return kOk;
  // Stage s is launched nsteps times back to back between one event pair
  // (an event between every kernel would add its own gap to each), after
  // one untimed launch.  The stages run in RK4 order on RK4's buffers; the
  // repeated last stage moves rho, so the state is saved and restored.
  const size_t bytes = state_elems() * 2 * sizeof(double);
  double* saved = nullptr;
  rc = Check(hipMalloc((void**)&saved, bytes), "alloc timing state");
  if (rc) return rc;
  rc = Check(hipMemcpyAsync(saved, d_rho_, bytes, hipMemcpyDeviceToDevice, stream_), "save state");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = Check(hipEventCreate(&e0), "event");
  if (!rc) rc = Check(hipEventCreate(&e1), "event");
  const int form = StepForm();
  const void* in[4];
  void* out[4];
  void* aux;
  StageBuffers(form, in, out, &aux);
  for (int s = 0; s < 4 && !rc; ++s) {
    for (long long i = 0; i <= nsteps && !rc; ++i) {
      if (i == 1) rc = Check(hipEventRecord(e0, stream_), "record");
      if (!rc) rc = Check(LaunchRk4Stage(StageArgs(s), fp32_, s, form, in[s], d_rho_, aux, out[s],
                                         dt, stream_), "rk4 stage");
    }
    if (!rc) rc = Check(hipEventRecord(e1, stream_), "record");
    if (!rc) rc = Check(hipEventSynchronize(e1), "event sync");
    float t = 0;
    if (!rc) rc = Check(hipEventElapsedTime(&t, e0, e1), "elapsed");
    ms[s] = t;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  const int rc_restore = Check(hipMemcpyAsync(d_rho_, saved, bytes, hipMemcpyDeviceToDevice,
                                              stream_), "restore state");
  const int rc_sync = Synchronize();
  Free(saved);
  return rc ? rc : (rc_restore ? rc_restore : rc_sync);
}

// BiCGSTAB(ell) (Sleijpen and Fokkema, ETNA 1, 1993, Algorithm 3.1) on the
// steady-state system of SteadyState, complex, inner products (a, b) =
// a^H b: each cycle takes ell BiCG steps (u_0..u_ell, r_0..r_ell and their
// images under A), then minimises |r_0 - sum_j gamma_j r_j| over the
// Krylov polynomial of degree ell (modified Gram-Schmidt on r_1..r_ell).
// Scalars on the host: every inner product is one reduction and one
// round trip.  PHI's BicgstabLSteadyState (hierarchy_integrator.cc:
// 3112-3489) follows the same outline.
int Engine::SteadyStateL(int ell, double tol, int max_iter, int* iters, double* resid,
                         std::vector<double>* history) {
  if (members_) return Unsupported("the steady state");
  if (bath_model_ != 0) {
    err_ = "the steady state with diagonal/full coupling operators is not supported";
    return kErrUnsupported;
  }
  if (tl_) {
    err_ = "the steady state with time-local truncation is not supported";
    return kErrUnsupported;
  }
  if (ell < 1 || ell > kBicglMaxEll) {
    err_ = "BiCGSTAB(ell) takes ell = 1 .. 8";
    return kErrInvalidArgument;
  }
  if (max_iter < 0) {
    err_ = "BiCGSTAB(ell) takes max_iter >= 0";
    return kErrInvalidArgument;
  }
  const size_t n = state_elems();
  const size_t bytes = n * 2 * sizeof(double);
  // r_0..r_ell, u_0..u_ell, rtilde
  std::vector<double*> vr(ell + 1, nullptr), vu(ell + 1, nullptr);
  double* rt = nullptr;
  int rc = kOk;
  for (auto& p : vr)
    if (!rc) rc = Check(hipMalloc((void**)&p, bytes), "alloc bicgstab(l)");
  for (auto& p : vu)
    if (!rc) rc = Check(hipMalloc((void**)&p, bytes), "alloc bicgstab(l)");
  if (!rc) rc = Check(hipMalloc((void**)&rt, bytes), "alloc bicgstab(l)");
  auto cleanup = [&]() {
    for (auto p : vr) Free(p);
    for (auto p : vu) Free(p);
    Free(rt);
  };
  if (rc) {
    cleanup();
    return rc;
  }
  double* x = d_rho_;
  // Scalars and the per-cycle residuals on the device.
  BicglScalars* S = nullptr;
  double* hist = nullptr;
  // (the residual history keeps its first 2^20 cycles)
  const int hist_cap = (int)std::min<long long>((long long)max_iter / ell + 2, 1 << 20);
  rc = Check(hipMalloc((void**)&S, sizeof(BicglScalars)), "alloc bicgstab(l) scalars");
  if (!rc) rc = Check(hipMalloc((void**)&hist, (size_t)hist_cap * sizeof(double)), "alloc hist");
  // dot product into S->dot[k], then the scalar step op in the same kernel
  auto dot_step = [&](const double* a, const double* b, int k, int op, int i, int j) -> int {
    return Check(LaunchBicglDotStep(a, b, n, d_partial_, red_blocks_, S, k, op, i, j, tol,
                                    max_iter, hist, hist_cap, stream_),
                 "bicgstab(l) step");
  };
  // (a, b) into S->dot[k] and (c, d) into S->dot[k + 1] in one pass, then
  // the step op; the partials of 4 doubles per block fill d_partial_
  // (2 * red_blocks_ doubles) with red_blocks_ / 2 blocks.
  auto dot2_step = [&](const double* a, const double* b, const double* c, const double* d, int k,
                       int op, int i, int j) -> int {
    return Check(LaunchBicglDot2Step(a, b, c, d, n, d_partial_, red_blocks_ / 2, S, k, op, i, j,
                                     tol, max_iter, hist, hist_cap, stream_),
                 "bicgstab(l) step");
  };
  // y = a + sign * c * b, c a scalar of S
  auto axpy = [&](double* y, const double* a, const double* b, const double2* c,
                  double sign) -> int {
    return Check(LaunchZaxpyDev(y, a, b, n, c, sign, &S->stop, stream_), "axpy");
  };
  // y = A x with (rtilde, y) into S->dot[0], then the step op: the matvec
  // writes the dot product's partials itself when its LDS form fits.
  const int mb = RhsGeneralDotBlocks(args_);
  double* pmv = nullptr;
  if (!rc && mb > 0)
    rc = Check(hipMalloc((void**)&pmv, (size_t)mb * 2 * sizeof(double)), "alloc matvec partials");
  auto matvec_dot = [&](const double* xin, double* y, int op) -> int {
    if (mb > 0) {
      int e = Check(LaunchRhsGeneralDot(args_, xin, y, 1, 1, rt, nullptr, pmv, stream_),
                    "rhs_general dot");
      if (!e)
        e = Check(LaunchBicglReduceStep(pmv, mb, S, 0, op, 0, 0, tol, max_iter, hist, hist_cap,
                                        stream_),
                  "bicgstab(l) step");
      return e;
    }
    int e = RhsDevice(xin, y, 1);
    if (!e) e = dot_step(rt, y, 0, op, 0, 0);
    return e;
  };
  auto pairs = [&](const ZPairs& pr, const double2* c) -> int {
    return Check(LaunchZaxpyPairs(pr, n, c, &S->stop, stream_), "axpy pairs");
  };
  auto comb = [&](double* y, const ZTerms& t) -> int {
    return Check(LaunchZcombDev(y, y, t, n, &S->stop, stream_), "lincomb terms");
  };
  // r_0 = b - A x (b = e_0 of the trace row), rtilde = r_0, u_0 = 0
  if (!rc) rc = RhsDevice(x, vu[0], 1);
  if (!rc) {
    const double* xs[1] = {vu[0]};
    const double2 cs[1] = {make_double2(-1., 0.)};
    rc = Check(LaunchLincomb(vr[0], n, 1, xs, cs, stream_), "lincomb");
  }
  double r00[2];
  if (!rc) rc = Check(hipMemcpyAsync(r00, vr[0], sizeof(r00), hipMemcpyDeviceToHost, stream_), "r0");
  if (!rc) rc = Synchronize();
  if (!rc) {
    r00[0] += 1.0;
    rc = Check(hipMemcpyAsync(vr[0], r00, sizeof(r00), hipMemcpyHostToDevice, stream_), "r0 set");
  }
  if (!rc) rc = Check(hipMemcpyAsync(rt, vr[0], bytes, hipMemcpyDeviceToDevice, stream_), "rtilde");
  if (!rc) rc = Check(hipMemsetAsync(vu[0], 0, bytes, stream_), "u0");
  // |r_0|^2 into dot[2] and the first BiCG step's (rtilde, r_0) into dot[3];
  // every cycle ends the same way.
  if (!rc) rc = dot2_step(vr[0], vr[0], rt, vr[0], 2, kBlInit, 0, 0);
  // One cycle: ell BiCG steps, then the minimal-residual part (the first
  // step's beta comes from the previous kBlInit / kBlEnd).
  auto enqueue_cycle = [&]() -> int {
    int e = kOk;
    for (int j = 0; j < ell && !e; ++j) {
      // (beta: from kBlInit / kBlEnd for j = 0, else from the matvec that
      // made r_j)
      // u_i = r_i - beta u_i, i <= j
      ZPairs up{};
      for (int i = 0; i <= j; ++i) {
        up.y[i] = vu[i];
        up.a[i] = vr[i];
        up.b[i] = vu[i];
        up.sign[i] = -1.;
      }
      up.count = j + 1;
      if (!e) e = pairs(up, &S->beta);
      if (!e) e = matvec_dot(vu[j], vu[j + 1], kBlAlpha);
      // r_i -= alpha u_{i+1}, i <= j, and x += alpha u_0
      ZPairs rp{};
      for (int i = 0; i <= j; ++i) {
        rp.y[i] = vr[i];
        rp.a[i] = vr[i];
        rp.b[i] = vu[i + 1];
        rp.sign[i] = -1.;
      }
      rp.y[j + 1] = x;
      rp.a[j + 1] = x;
      rp.b[j + 1] = vu[0];
      rp.sign[j + 1] = 1.;
      rp.count = j + 2;
      if (!e) e = pairs(rp, &S->alpha);
      // r_{j+1} = A r_j, and the next BiCG step's beta from (rtilde, r_{j+1})
      if (!e)
        e = j + 1 < ell ? matvec_dot(vr[j], vr[j + 1], kBlBeta) : RhsDevice(vr[j], vr[j + 1], 1);
    }
    // modified Gram-Schmidt
    for (int j = 1; j <= ell && !e; ++j) {
      for (int i = 1; i < j && !e; ++i) {
        e = dot_step(vr[i], vr[j], 0, kBlTau, i, j);
        if (!e) e = axpy(vr[j], vr[j], vr[i], &S->tau[i][j], -1.);
      }
      // (for j = ell this also solves for g, gpp and omega)
      if (!e) e = dot2_step(vr[j], vr[j], vr[j], vr[0], 0, kBlSigma, ell, j);
    }
    // x += g_1 r_0 + sum gpp_j r_j (before r_0 changes), then
    // r_0 -= gp_ell r_ell + sum gp_j r_j and u_0 -= g_ell u_ell + sum g_j u_j
    ZTerms tx{}, tr{}, tu{};
    tx.b[0] = vr[0], tx.c[0] = &S->g[1], tx.sign[0] = 1.;
    tr.b[0] = vr[ell], tr.c[0] = &S->gp[ell], tr.sign[0] = -1.;
    tu.b[0] = vu[ell], tu.c[0] = &S->g[ell], tu.sign[0] = -1.;
    for (int j = 1; j <= ell - 1; ++j) {
      tx.b[j] = vr[j], tx.c[j] = &S->gpp[j], tx.sign[j] = 1.;
      tr.b[j] = vr[j], tr.c[j] = &S->gp[j], tr.sign[j] = -1.;
      tu.b[j] = vu[j], tu.c[j] = &S->g[j], tu.sign[j] = -1.;
    }
    tx.count = tr.count = tu.count = ell;
    if (!e) e = comb(x, tx);
    if (!e) e = comb(vr[0], tr);
    if (!e) e = comb(vu[0], tu);
    if (!e) e = dot2_step(vr[0], vr[0], rt, vr[0], 2, kBlEnd, ell, 0);
    return e;
  };
  // Batches of cycles (about 16 BiCG steps) between the host's checks,
  // captured once as a hipGraph and replayed; cycles after the stop are
  // no-ops.
  const int batch = std::max(1, 16 / ell);
  auto enqueue_batch = [&]() -> int {
    int e = kOk;
    for (int b = 0; b < batch && !e; ++b) e = enqueue_cycle();
    return e;
  };
  BicglScalars hs{};
  if (!rc) rc = Check(hipMemcpyAsync(&hs, S, sizeof(hs), hipMemcpyDeviceToHost, stream_), "scalars");
  if (!rc) rc = Synchronize();
  hipGraphExec_t exec = nullptr;
  while (!rc && !hs.stop) {
    if (use_graphs_ && !exec) {
      hipGraph_t g = nullptr;
      rc = Check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
                 "begin bicgstab(l) capture");
      if (!rc) {
        rc = enqueue_batch();
        const int rc_end = Check(hipStreamEndCapture(stream_, &g), "end bicgstab(l) capture");
        if (!rc) rc = rc_end;
      }
      if (!rc) rc = Check(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0),
                          "instantiate bicgstab(l) batch");
      if (g) (void)hipGraphDestroy(g);
      if (rc) exec = nullptr;
    }
    if (!rc) rc = exec ? Check(hipGraphLaunch(exec, stream_), "bicgstab(l) batch") : enqueue_batch();
    if (!rc) rc = Check(hipMemcpyAsync(&hs, S, sizeof(hs), hipMemcpyDeviceToHost, stream_), "scalars");
    if (!rc) rc = Synchronize();
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (!rc && history) {
    std::vector<double> h(std::min(hs.cycles + 1, hist_cap));
    rc = Check(hipMemcpyAsync(h.data(), hist, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                              stream_), "residual history");
    if (!rc) rc = Synchronize();
    if (!rc) history->insert(history->end(), h.begin(), h.end());
  }
  const int rc_sync = Synchronize();
  Free(S);
  Free(hist);
  Free(pmv);
  cleanup();
  if (!rc) rc = rc_sync;
  if (iters) *iters = hs.it;
  if (resid) *resid = hs.res;
  if (rc) return rc;
  return hs.res <= tol ? kOk : kErrNotConverged;
}

int Engine::HistoryCount() {
  int n = 0;
  int rc = Check(hipMemcpyAsync(&n, d_hist_count_, sizeof(int), hipMemcpyDeviceToHost, stream_), "hist count");
  if (!rc) rc = Synchronize();
  return rc ? rc : n;
}

int Engine::DrainHistory(std::vector<std::complex<double>>* out) {
  int n = HistoryCount();
  if (n < 0) return n;
  n = std::min(n, hist_capacity_);
  const size_t ne = (size_t)N() * N();
  out->resize((size_t)n * ne);
  int rc = kOk;
  if (n > 0)
    rc = Check(hipMemcpyAsync(out->data(), d_hist_, (size_t)n * ne * sizeof(std::complex<double>),
                              hipMemcpyDeviceToHost, stream_), "drain hist");
  if (!rc) rc = Check(hipMemsetAsync(d_hist_count_, 0, sizeof(int), stream_), "reset hist");
  if (!rc) rc = Synchronize();
  return rc ? rc : n;
}

int Engine::Rkf45Attempt(const Rkf45Coefs& c, double* err, double t, const double* a,
                         HamiltonianSource* src) {
  if (members_) return Unsupported("RKF45");
  if (src && filter_on()) {
    err_ = "FilterTolerance with a time-dependent Hamiltonian is not supported";
    return kErrUnsupported;
  }
  if (src && closure_) {
    err_ = "a closure with a time-dependent Hamiltonian is not supported";
    return kErrUnsupported;
  }
  int rc = AllocBulk();
  if (rc) return rc;
  if (!rkf45_alloc_) {
    const size_t bytes = state_elems() * 2 * (fp32_ ? sizeof(float) : sizeof(double));
    for (auto& k : d_k_)
      if (!rc) rc = Check(hipMalloc(&k, bytes), "alloc k");
    if (!rc) rc = Check(hipMalloc(&d_tmp_, bytes), "alloc tmp");
    if (rc) return rc;
    rkf45_alloc_ = true;
  }
  if (explicit_path()) return ExplicitRkf45Attempt(c, err, t, a, src);
  Rkf45Bufs<float> bf{};
  Rkf45Bufs<double> bd{};
  for (int i = 0; i < 5; ++i) {
    bf.k[i] = (float*)d_k_[i];
    bd.k[i] = (double*)d_k_[i];
  }
  // Stage inputs ping-pong between d_tmp_ and d_y_[0].
  void* tin[2] = {d_tmp_, d_y_[0]};
  bf.rho_new = bd.rho_new = d_aux_;
  bf.err = bd.err = d_err_;
  rc = Check(LaunchFillU64(d_err_, 0ull, 1, stream_), "zero err");
  for (int s = 0; s < 6 && !rc; ++s) {
    if (src && src->Update(s == 0 ? t : t + a[s - 1])) {
      rc = SetHamiltonian(src->H());
      const std::complex<double>* R = src->Rotation();
      if (!rc && R) {
        // rho (FP64); the stage input and k0..k(s-1) (bulk precision).
        void* bufs[7];
        bool f32[7];
        int nb = 0;
        bufs[nb] = d_rho_;
        f32[nb++] = false;
        if (s > 0) {
          bufs[nb] = tin[(s - 1) & 1];
          f32[nb++] = fp32_;
          for (int j = 0; j < s; ++j)
            if (!(s == 5 && j == 1)) {
              bufs[nb] = d_k_[j];
              f32[nb++] = fp32_;
            }
        }
        rc = Rotate(R, nb, bufs, f32);
      }
      ++stage_slot_;
      if (rc) break;
    }
    bf.tmp_out = (float*)tin[s & 1];
    bd.tmp_out = (double*)tin[s & 1];
    const void* in = s == 0 ? (const void*)d_rho_ : tin[(s - 1) & 1];
    rc = Check(LaunchRkf45Stage(args_, fp32_, s, in, d_rho_, bf, bd, c, stream_), "rkf45 stage");
  }
  unsigned long long bits = 0;
  if (!rc) rc = Check(hipMemcpyAsync(&bits, d_err_, sizeof(bits), hipMemcpyDeviceToHost, stream_), "err");
  if (!rc) rc = Synchronize();
  // The attempt's staging slots are free again.
  stage_slot_ = 0;
  if (rc) return rc;
  std::memcpy(err, &bits, sizeof(double));
  return kOk;
}

int Engine::Rkf45Accept() {
  std::swap(d_rho_, d_aux_);
  if (tl_) {
    // Time-local matrices to t + dt (a4 = dt), saved for the next step.
    int rc = TlStep(tl_a_[3] - tl_a_[4]);
    tl_->Store();
    if (rc) return rc;
  }
  if (!filter_on()) return kOk;
  return Check(LaunchUpdateActive(d_rho_, d_active_, N() * N(), num_nodes(), filter_, stream_),
               "flag update");
}

int Engine::SteadyState(double tol, int max_iter, int* iters, double* resid,
                        std::vector<double>* history) {
  if (members_) return Unsupported("the steady state");
  if (bath_model_ != 0) {
    err_ = "the steady state with diagonal/full coupling operators is not supported";
    return kErrUnsupported;
  }
  if (tl_) {
    // PHI switches the option off for the steady state (phi_parameters.cc:
    // 833-837); the C ABI can still ask for it.
    err_ = "the steady state with time-local truncation is not supported";
    return kErrUnsupported;
  }
  // Unpreconditioned complex BiCGSTAB (van der Vorst 1992) on A x = b with
  // b = e_(node 0, element (0,0)); inner products conjugate the first
  // argument; all scalars and reductions in FP64.
  const size_t n = state_elems();
  const size_t bytes = n * 2 * sizeof(double);
  double* v[7] = {nullptr};  // r, rhat, p, vv, s, t, tmp
  int rc = kOk;
  for (auto& p : v)
    if (!rc) rc = Check(hipMalloc((void**)&p, bytes), "alloc bicgstab");
  auto cleanup = [&]() { for (auto p : v) Free(p); };
  if (rc) { cleanup(); return rc; }
  double *r = v[0], *rh = v[1], *pp = v[2], *vv = v[3], *s = v[4], *t = v[5], *x = d_rho_;
  typedef std::complex<double> Z;
  auto dot = [&](const double* a, const double* b, Z* out) -> int {
    int e = Check(LaunchDot(a, b, n, 0, d_partial_, red_blocks_, d_dot_, stream_), "dot");
    double h[2];
    if (!e) e = Check(hipMemcpyAsync(h, d_dot_, sizeof(h), hipMemcpyDeviceToHost, stream_), "dot d2h");
    if (!e) e = Synchronize();
    *out = Z(h[0], h[1]);
    return e;
  };
  auto lin = [&](double* y, int nv, const double* const* xs, const Z* cs) -> int {
    double2 c[4];
    for (int i = 0; i < nv; ++i) c[i] = make_double2(cs[i].real(), cs[i].imag());
    return Check(LaunchLincomb(y, n, nv, xs, c, stream_), "lincomb");
  };
  // r = b - A x
  rc = RhsDevice(x, t, 1);
  {
    const double* xs[1] = {t};
    Z cs[1] = {Z(-1, 0)};
    if (!rc) rc = lin(r, 1, xs, cs);
  }
  double one[2] = {1.0, 0.0};
  double rr0[2];
  if (!rc) rc = Check(hipMemcpyAsync(rr0, r, sizeof(rr0), hipMemcpyDeviceToHost, stream_), "r0");
  if (!rc) rc = Synchronize();
  if (!rc) {
    rr0[0] += one[0];
    rr0[1] += one[1];
    rc = Check(hipMemcpyAsync(r, rr0, sizeof(rr0), hipMemcpyHostToDevice, stream_), "r0 set");
  }
  if (!rc) rc = Check(hipMemcpyAsync(rh, r, bytes, hipMemcpyDeviceToDevice, stream_), "rhat");
  // p0 = r0 (the first iteration's p; later ones come with the x/r update)
  if (!rc) rc = Check(hipMemcpyAsync(pp, r, bytes, hipMemcpyDeviceToDevice, stream_), "p0");
  Z rn2;
  if (!rc) rc = dot(r, r, &rn2);
  double res = std::sqrt(rn2.real());
  // The iterations keep their scalars on the device (BicgScalars): the dot
  // products land there, the update kernels derive alpha, omega and beta
  // from them, and the next iteration's s update closes each iteration (a
  // one-block kernel closes the last of a batch).  Four kernels per
  // iteration: v = A p, s, t = A s (with rhat . t and rhat . s, so that
  // rhat . r_new = rhat . s - omega rhat . t is known before r_new exists),
  // and x, r and p in one pass.  The host checks after every kBatch
  // iterations; once converged (or broken down, or at max_iter) the
  // updates and the closing kernel do nothing, so x and the count are those
  // of stopping right there.
  constexpr int kBatch = 16;
  BicgScalars* st = nullptr;
  double* hist = nullptr;
  double* part3 = nullptr;
  // Two records alternate: iteration k works on st[k % 2], and closing it
  // writes st[(k + 1) % 2] while other blocks may still read st[k % 2].
  if (!rc) rc = Check(hipMalloc((void**)&st, 2 * sizeof(BicgScalars)), "alloc bicgstab scalars");
  // Partials: rhat.v (2 per block), t.s / t.t / rhat.t / rhat.s (7), |r|^2
  // (1).  When the matvec's LDS form fits, it writes the rhat.v and the
  // seven-value partials itself (one record per matvec block; mb blocks).
  constexpr int kRedBlocks = 512;
  const int mb = RhsGeneralDotBlocks(args_);
  const int nb_dot = mb > 0 ? mb : kRedBlocks;
  const size_t cap = (size_t)std::max(kRedBlocks, mb);
  if (!rc)
    rc = Check(hipMalloc((void**)&part3, cap * 10 * sizeof(double)), "alloc bicgstab partials");
  double* p_rv = part3;
  double* p_ts = part3 + 2 * cap;
  double* p_xr = part3 + 9 * cap;
  if (!rc)
    rc = Check(hipMalloc((void**)&hist, (size_t)(std::max(max_iter, 0) + 1) * sizeof(double)),
               "alloc bicgstab history");
  BicgScalars hs{};
  hs.rho_old = hs.alpha = hs.omega = make_double2(1., 0.);
  hs.res = res;
  hs.done = !(res > tol) || max_iter <= 0 ? 1 : 0;
  hs.it = 0;
  if (!rc) rc = Check(hipMemcpyAsync(st, &hs, sizeof(hs), hipMemcpyHostToDevice, stream_),
                      "upload bicgstab scalars");
  if (!rc) rc = Check(hipMemcpyAsync(hist, &res, sizeof(double), hipMemcpyHostToDevice, stream_),
                      "upload residual");
  auto dev_dot = [&](const double* a, const double* b, double2* out) -> int {
    return Check(LaunchDot(a, b, n, 0, d_partial_, red_blocks_, reinterpret_cast<double*>(out),
                           stream_), "dot");
  };
  // rho for the first iteration; later ones come with the r update.
  if (!rc && !hs.done) rc = dev_dot(rh, r, &st->rho_new);
  // One batch: kBatch iterations, then the last one closes into the record
  // the next batch starts from.  kBatch is even, so every batch launches the
  // same kernels on the same buffers and records; in the matvec's LDS form
  // (kernels only) the batch is captured once as a hipGraph and replayed,
  // which removes most of the gaps between its 4 kBatch + 1 launches.
  auto enqueue_batch = [&](long long k0) -> int {
    int e = kOk;
    long long k = k0;
    for (int b = 0; b < kBatch && !e; ++b, ++k) {
      BicgScalars* cur = st + (k & 1);
      BicgScalars* prev = st + ((k + 1) & 1);
      // v = A p with the rhat.v partials
      if (mb > 0)
        e = Check(LaunchRhsGeneralDot(args_, pp, vv, 1, 1, rh, nullptr, p_rv, stream_),
                  "A p, rhat.v");
      if (!e && mb == 0) e = RhsDevice(pp, vv, 1);
      if (!e && mb == 0)
        e = Check(LaunchDotPartial(rh, vv, n, p_rv, kRedBlocks, stream_), "rhat.v");
      // (close the previous iteration, then) s = r - alpha v
      if (!e)
        e = Check(LaunchBicgS(b > 0, s, n, r, vv, prev, cur, p_rv, nb_dot, p_xr, kRedBlocks,
                              hist, tol, max_iter, stream_), "s update");
      // t = A s with the t.s, t.t, rhat.t, rhat.s partials
      if (!e && mb > 0)
        e = Check(LaunchRhsGeneralDot(args_, s, t, 1, 3, s, rh, p_ts, stream_),
                  "A s, t.s, t.t, rhat.t, rhat.s");
      if (!e && mb == 0) e = RhsDevice(s, t, 1);
      if (!e && mb == 0)
        e = Check(LaunchBicgDots(t, s, rh, n, p_ts, kRedBlocks, stream_), "t and s dots");
      // x += alpha p + omega s, r = s - omega t, p = r + beta (p - omega v),
      // partials of |r|^2
      if (!e)
        e = Check(LaunchBicgXRP(x, r, pp, s, t, vv, n, cur, p_ts, nb_dot, p_xr, kRedBlocks,
                                stream_),
                  "x/r/p update");
    }
    if (!e)
      e = Check(LaunchBicgFinish(st + ((k + 1) & 1), st + (k & 1), p_xr, kRedBlocks, hist, tol,
                                 max_iter, stream_), "finish");
    return e;
  };
  hipGraphExec_t batch_exec = nullptr;
  long long k = 0;  // iterations enqueued
  while (!rc && !hs.done) {
    if (mb > 0 && use_graphs_ && !batch_exec) {
      hipGraph_t g = nullptr;
      rc = Check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
                 "begin bicgstab capture");
      if (!rc) {
        rc = enqueue_batch(k);
        const int rc_end = Check(hipStreamEndCapture(stream_, &g), "end bicgstab capture");
        if (!rc) rc = rc_end;
      }
      if (!rc) rc = Check(hipGraphInstantiate(&batch_exec, g, nullptr, nullptr, 0),
                          "instantiate bicgstab batch");
      if (g) (void)hipGraphDestroy(g);
      if (rc) batch_exec = nullptr;
    }
    if (!rc)
      rc = batch_exec ? Check(hipGraphLaunch(batch_exec, stream_), "bicgstab batch")
                      : enqueue_batch(k);
    k += kBatch;
    if (!rc) rc = Check(hipMemcpyAsync(&hs, st + (k & 1), sizeof(hs), hipMemcpyDeviceToHost,
                                       stream_), "bicgstab scalars");
    if (!rc) rc = Synchronize();
  }
  if (batch_exec) (void)hipGraphExecDestroy(batch_exec);
  const int it = hs.it;
  if (!rc) res = hs.res;
  if (!rc && history) {
    std::vector<double> h(it + 1);
    rc = Check(hipMemcpyAsync(h.data(), hist, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                              stream_), "residual history");
    if (!rc) rc = Synchronize();
    if (!rc) history->insert(history->end(), h.begin(), h.end());
  }
  Free(st);
  Free(hist);
  Free(part3);
  cleanup();
  if (iters) *iters = it;
  if (resid) *resid = res;
  if (rc) return rc;
  return res <= tol ? kOk : kErrNotConverged;
}

}  // namespace phi
