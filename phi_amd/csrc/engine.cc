// Device engine of the MI355X HEOM integrator.  See engine.h.
#include "engine.h"

#include <algorithm>
#include <cmath>
#include <cstring>

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
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    *why = "no HIP device is visible";
    return kErrNoDevice;
  }
  Engine* e = new Engine;
  std::string msg = BuildHierarchy(p, &e->h_);
  if (!msg.empty()) {
    *why = msg;
    delete e;
    return kErrInvalidParameters;
  }
  const HierarchyTables& h = e->h_;
  const int N = h.N, NE = N * N;
  if ((int)p.time_indepenent_hamiltonian.size() != NE) {
    *why = "a time-independent Hamiltonian is required";
    delete e;
    return kErrUnsupported;
  }
  int rc = e->Check(hipStreamCreateWithFlags(&e->stream_, hipStreamNonBlocking), "stream");
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
  if (mps <= 2) {
    std::vector<int4> edge((size_t)h.num_nodes * N * mps);
    for (int n = 0; n < h.num_nodes; ++n)
      for (int b = 0; b < N; ++b)
        for (int q = 0; q < mps; ++q) {
          int4 r = make_int4(-1, -1, -1, 0);
          if (!h.vacuum && q < h.site_ptr[b + 1] - h.site_ptr[b]) {
            const int j = h.site_modes[h.site_ptr[b] + q];
            const size_t o = (size_t)n * h.D + j;
            r = make_int4(h.next[o], h.prev[o], j * h.L + h.occ[o], 0);
          }
          edge[((size_t)n * N + b) * mps + q] = r;
        }
    if (!rc) rc = e->Check(Upload(&e->d_edge_, edge.data(), edge.size()), "upload edges");
  }
  e->mps_ = mps;
  const size_t bytes = e->state_elems() * 2 * sizeof(double);
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_rho_, bytes), "alloc rho");
  if (!rc) rc = e->Check(hipMemset(e->d_rho_, 0, bytes), "zero rho");
  e->hist_capacity_ = kHistoryCapacity;
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_hist_, (size_t)kHistoryCapacity * NE * 2 * sizeof(double)), "alloc hist");
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_hist_count_, sizeof(int)), "alloc hist count");
  if (!rc) rc = e->Check(hipMemset(e->d_hist_count_, 0, sizeof(int)), "zero hist count");
  e->red_blocks_ = 1024;
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_partial_, e->red_blocks_ * 2 * sizeof(double)), "alloc partial");
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_dot_, 2 * sizeof(double)), "alloc dot");
  if (!rc) rc = e->Check(hipMalloc((void**)&e->d_err_, sizeof(unsigned long long)), "alloc err");
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

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  void* ptrs[] = {d_next_, d_prev_, d_occ_, d_mu_, d_site_ptr_, d_site_modes_, d_next_s_,
                  d_prev_col_, d_prev_row_, d_delta_, d_H_, d_Hraw_, d_Hadj_, d_edge_, d_rho_, d_aux_,
                  d_y_[0], d_y_[1], d_k_[0], d_k_[1], d_k_[2], d_k_[3], d_k_[4], d_tmp_,
                  d_err_, d_hist_, d_hist_count_, d_partial_, d_dot_};
  for (void* p : ptrs) Free(p);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void Engine::set_fp32(bool v) { fp32_ = v; }

int Engine::SetLaunch(int col_nm, int tile, int threads, int blocks) {
  RhsArgs& a = args_;
  const int N = h_.N, D = h_.D, L = h_.L, NE = N * N;
  const size_t kLdsBudget = 64 * 1024;
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
      if (threads == 0) th = std::min(1024, (t * NE + 63) / 64 * 64);
      const int ept = (t * NE + th - 1) / th;
      if (ept > 4 || (ept > 1 && th > 256) ||
          StreamTileLdsBytes(N, D, L, t, f == 1) > kLdsBudget) {
        err_ = "tile does not fit the persistent kernels (at most 4 elements per thread, "
               "256 threads when more than one)";
        return kErrInvalidArgument;
      }
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
        err_ = "NumStates too large for the LDS tile of this build";
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
  return kOk;
}

int Engine::AllocBulk() {
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
  return Check(LaunchRhsGeneral(args_, x, y, trace_row, stream_), "rhs_general");
}

int Engine::Rhs(const std::complex<double>* x, std::complex<double>* y) {
  const size_t bytes = state_elems() * sizeof(std::complex<double>);
  double *dx = nullptr, *dy = nullptr;
  int rc = Check(hipMalloc((void**)&dx, bytes), "alloc rhs in");
  if (!rc) rc = Check(hipMalloc((void**)&dy, bytes), "alloc rhs out");
  if (!rc) rc = Check(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, stream_), "rhs upload");
  if (!rc) rc = RhsDevice(dx, dy, 0);
  if (!rc) rc = Check(hipMemcpyAsync(y, dy, bytes, hipMemcpyDeviceToHost, stream_), "rhs download");
  if (!rc) rc = Synchronize();
  Free(dx);
  Free(dy);
  return rc;
}

int Engine::Rk4Steps(long long nsteps, double dt, long long step_offset, int snap_every) {
  int rc = AllocBulk();
  if (rc) return rc;
  const bool hn = rk4_nested_;
  for (long long s = 0; s < nsteps; ++s) {
    // Stage inputs alternate between the two bulk buffers so that a stage
    // never overwrites what its own neighbour gathers read.
    rc = Check(LaunchRk4Stage(args_, fp32_, 0, hn, d_rho_, d_rho_, d_aux_, d_y_[0], dt, stream_), "rk4 s0");
    if (!rc) rc = Check(LaunchRk4Stage(args_, fp32_, 1, hn, d_y_[0], d_rho_, d_aux_, d_y_[1], dt, stream_), "rk4 s1");
    if (!rc) rc = Check(LaunchRk4Stage(args_, fp32_, 2, hn, d_y_[1], d_rho_, d_aux_, d_y_[0], dt, stream_), "rk4 s2");
    if (!rc) rc = Check(LaunchRk4Stage(args_, fp32_, 3, hn, d_y_[0], d_rho_, d_aux_, nullptr, dt, stream_), "rk4 s3");
    if (rc) return rc;
    const long long i = step_offset + s;
    if (snap_every > 0 && i % snap_every == 0) {
      rc = Check(LaunchSnapshot(d_rho_, d_hist_, d_hist_count_, N() * N(), hist_capacity_, stream_), "snapshot");
      if (rc) return rc;
    }
  }
  return kOk;
}

int Engine::Rk4StageTimes(long long nsteps, double dt, double ms[4]) {
  int rc = AllocBulk();
  if (rc) return rc;
  hipEvent_t ev[5];
  for (auto& e : ev)
    if (!rc) rc = Check(hipEventCreate(&e), "event");
  if (rc) return rc;
  for (int s = 0; s < 4; ++s) ms[s] = 0;
  const void* in[4] = {d_rho_, d_y_[0], d_y_[1], d_y_[0]};
  void* out[4] = {d_y_[0], d_y_[1], d_y_[0], nullptr};
  for (long long i = 0; i < nsteps && !rc; ++i) {
    rc = Check(hipEventRecord(ev[0], stream_), "record");
    for (int s = 0; s < 4 && !rc; ++s) {
      rc = Check(LaunchRk4Stage(args_, fp32_, s, rk4_nested_, in[s], d_rho_, d_aux_, out[s], dt,
                                stream_), "rk4 stage");
      if (!rc) rc = Check(hipEventRecord(ev[s + 1], stream_), "record");
    }
    if (!rc) rc = Check(hipEventSynchronize(ev[4]), "event sync");
    for (int s = 0; s < 4 && !rc; ++s) {
      float t = 0;
      rc = Check(hipEventElapsedTime(&t, ev[s], ev[s + 1]), "elapsed");
      ms[s] += t;
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  return rc;
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

int Engine::Rkf45Attempt(const Rkf45Coefs& c, double* err) {
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
    bf.tmp_out = (float*)tin[s & 1];
    bd.tmp_out = (double*)tin[s & 1];
    const void* in = s == 0 ? (const void*)d_rho_ : tin[(s - 1) & 1];
    rc = Check(LaunchRkf45Stage(args_, fp32_, s, in, d_rho_, bf, bd, c, stream_), "rkf45 stage");
  }
  unsigned long long bits = 0;
  if (!rc) rc = Check(hipMemcpyAsync(&bits, d_err_, sizeof(bits), hipMemcpyDeviceToHost, stream_), "err");
  if (!rc) rc = Synchronize();
  if (rc) return rc;
  std::memcpy(err, &bits, sizeof(double));
  return kOk;
}

void Engine::Rkf45Accept() { std::swap(d_rho_, d_aux_); }

int Engine::SteadyState(double tol, int max_iter, int* iters, double* resid,
                        std::vector<double>* history) {
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
  if (!rc) rc = Check(hipMemsetAsync(pp, 0, bytes, stream_), "p0");
  if (!rc) rc = Check(hipMemsetAsync(vv, 0, bytes, stream_), "v0");
  Z rho_old(1, 0), alpha(1, 0), omega(1, 0), tmp;
  Z rn2;
  if (!rc) rc = dot(r, r, &rn2);
  double res = std::sqrt(rn2.real());
  if (history) history->push_back(res);
  int it = 0;
  while (!rc && res > tol && it < max_iter) {
    ++it;
    Z rho_new;
    rc = dot(rh, r, &rho_new);
    if (rc) break;
    const Z beta = (rho_new / rho_old) * (alpha / omega);
    {  // p = r + beta (p - omega v)
      const double* xs[3] = {r, pp, vv};
      Z cs[3] = {Z(1, 0), beta, -beta * omega};
      rc = lin(pp, 3, xs, cs);
    }
    if (!rc) rc = RhsDevice(pp, vv, 1);
    Z rv;
    if (!rc) rc = dot(rh, vv, &rv);
    if (rc) break;
    alpha = rho_new / rv;
    {  // s = r - alpha v
      const double* xs[2] = {r, vv};
      Z cs[2] = {Z(1, 0), -alpha};
      rc = lin(s, 2, xs, cs);
    }
    if (!rc) rc = RhsDevice(s, t, 1);
    Z ts, tt;
    if (!rc) rc = dot(t, s, &ts);
    if (!rc) rc = dot(t, t, &tt);
    if (rc) break;
    omega = tt.real() > 0 ? ts / tt : Z(0, 0);
    {  // x = x + alpha p + omega s
      const double* xs[3] = {x, pp, s};
      Z cs[3] = {Z(1, 0), alpha, omega};
      rc = lin(x, 3, xs, cs);
    }
    {  // r = s - omega t
      const double* xs[2] = {s, t};
      Z cs[2] = {Z(1, 0), -omega};
      if (!rc) rc = lin(r, 2, xs, cs);
    }
    if (!rc) rc = dot(r, r, &rn2);
    res = std::sqrt(rn2.real());
    if (history) history->push_back(res);
    rho_old = rho_new;
    if (omega == Z(0, 0)) break;
  }
  cleanup();
  if (iters) *iters = it;
  if (resid) *resid = res;
  if (rc) return rc;
  return res <= tol ? kOk : kErrNotConverged;
}

}  // namespace phi
