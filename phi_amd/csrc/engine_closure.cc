// Learned truncation closure of the device engine: training samples from a
// deep ensemble, and the coupling of a trained network to a shallow one.
// See engine.h.
#include <algorithm>
#include <vector>

#include "closure_kernels.h"
#include "engine.h"
#include "mlp.h"

namespace phi {

namespace {

template <typename T>
hipError_t Up(T** dst, const T* src, size_t n) {
  hipError_t e = hipMalloc((void**)dst, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) return e;
  if (n) e = hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice);
  return e;
}

void FreeD(void* p) {
  if (p) (void)hipFree(p);
}

// Multi-indices / L of the nodes first .. first + nb - 1 (nb x D floats).
std::vector<float> BoundaryIndex(const HierarchyTables& h, int first, int nb, int L) {
  std::vector<float> b((size_t)nb * h.D);
  const float inv = 1.0f / (float)std::max(L, 1);
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < h.D; ++j) b[(size_t)i * h.D + j] = h.occ[(size_t)(first + i) * h.D + j] * inv;
  return b;
}

}  // namespace

int Engine::ClosureSamples(int shallow_lh, float* X, float* T, long long cap_rows,
                           long long* rows) {
  const HierarchyTables& h = h_;
  const int N = h.N, NE = N * N, D = h.D;
  if (shallow_lh < 2 || shallow_lh >= h.Lh) {
    err_ = "the sample depth must be at least 2 and below this hierarchy's depth";
    return kErrInvalidArgument;
  }
  const int B = shallow_lh - 1;  // boundary tier
  const int first = h.tier_start[B], nb = h.tier_start[B + 1] - h.tier_start[B];
  const int members = std::max(members_, 1);
  // Every child must be integrated in every member.
  for (int m = 0; m < members; ++m) {
    const int cnt = members_ ? mreal_[m] : h.num_nodes;
    if (cnt < h.tier_start[B + 2]) {
      err_ = "every member must be at least one tier deeper than the sample depth";
      return kErrInvalidArgument;
    }
  }
  const long long nrows = (long long)members * nb;
  *rows = nrows;
  if (nrows > cap_rows) {
    err_ = "sample buffer too small";
    return kErrInvalidArgument;
  }
  std::vector<int> children((size_t)nb * D);
  for (int i = 0; i < nb; ++i)
    for (int j = 0; j < D; ++j) children[(size_t)i * D + j] = h.next[(size_t)(first + i) * D + j];
  const std::vector<float> bmi = BoundaryIndex(h, first, nb, B);
  std::vector<int> moff = members_ ? moff_ : std::vector<int>{0};
  const int n_in = 2 * NE + D, n_out = 2 * NE * D;
  float *dX = nullptr, *dT = nullptr, *dS = nullptr, *dB = nullptr;
  int *dC = nullptr, *dM = nullptr;
  int rc = Check(Up(&dC, children.data(), children.size()), "upload children");
  if (!rc) rc = Check(Up(&dB, bmi.data(), bmi.size()), "upload indices");
  if (!rc) rc = Check(Up(&dM, moff.data(), moff.size()), "upload offsets");
  if (!rc) rc = Check(hipMalloc((void**)&dX, (size_t)nrows * n_in * 4), "alloc X");
  if (!rc) rc = Check(hipMalloc((void**)&dT, (size_t)nrows * n_out * 4), "alloc T");
  if (!rc) rc = Check(hipMalloc((void**)&dS, (size_t)nrows * 4), "alloc S");
  if (!rc)
    rc = Check(LaunchClosureInputs(d_rho_, false, dM, members, first, nb, NE, D, dB, dX, dS,
                                   stream_), "closure inputs");
  if (!rc)
    rc = Check(LaunchClosureTargets(d_rho_, dM, members, nb, NE, D, dC, dS, dT, stream_),
               "closure targets");
  if (!rc) rc = Check(hipMemcpyAsync(X, dX, (size_t)nrows * n_in * 4, hipMemcpyDeviceToHost,
                                     stream_), "download X");
  if (!rc) rc = Check(hipMemcpyAsync(T, dT, (size_t)nrows * n_out * 4, hipMemcpyDeviceToHost,
                                     stream_), "download T");
  if (!rc) rc = Synchronize();
  void* ptrs[] = {dX, dT, dS, dB, dC, dM};
  for (void* p : ptrs) FreeD(p);
  return rc;
}

int Engine::AttachClosure(Mlp* mlp) {
  const HierarchyTables& h = h_;
  const int N = h.N, NE = N * N, D = h.D;
  if (!mlp) {
    err_ = "no network";
    return kErrInvalidArgument;
  }
  if (!members_) {
    err_ = "the closure couples to an ensemble (one member for a single hierarchy)";
    return kErrUnsupported;
  }
  if (closure_) {
    err_ = "a closure is already attached";
    return kErrInvalidArgument;
  }
  for (int m = 0; m < members_; ++m)
    if (mreal_[m] != h.num_nodes) {
      err_ = "the closure needs every member at the same depth";
      return kErrUnsupported;
    }
  if (h.Lh < 2) {
    err_ = "the closure needs a hierarchy (HierarchyTruncation >= 2)";
    return kErrUnsupported;
  }
  if (mlp->n_in() != 2 * NE + D || mlp->n_out() != 2 * NE * D) {
    err_ = "network sizes do not match 2 N^2 + D inputs and 2 N^2 D outputs";
    return kErrInvalidArgument;
  }
  const int B = h.Lh - 1;
  const int first = h.tier_start[B], nb = h.tier_start[B + 1] - h.tier_start[B];
  const int vbase = h.num_nodes;
  const int span = h.num_nodes + nb * D;
  const long long total = (long long)members_ * span;
  if (total * NE > kMaxStateElems) {
    err_ = "ensemble slab with closure slots too large";
    return kErrUnsupported;
  }
  // Everything that can fail runs first, into new buffers; the engine
  // changes only once all of them exist, so a failed attach leaves it as
  // it was (state, offsets, edges and tiles).
  std::vector<int> moff(members_), mcount(members_, span);
  for (int m = 0; m < members_; ++m) moff[m] = m * span;
  FusedNet fnet{};
  void* pack = nullptr;
  double* rho = nullptr;
  int* dmoff = nullptr;
  int4* edge = nullptr;
  float *cbmi = nullptr, *cx = nullptr, *cs = nullptr, *cy = nullptr;
  int *csite = nullptr, *cslot = nullptr;
  unsigned long long* cfall = nullptr;
  // The stage kernels read slot j only in column site(j), so the fused pass
  // evaluates just those 2 N D of the network's 2 N^2 D outputs and writes
  // them straight into the slots, interleaved complex: packed output
  // 2 (j N + w) + part is the network's output j 2 N^2 + part N^2 +
  // site(j) N + w (element w of that column), at float offset
  // 2 (j N^2 + site(j) N + w) + part from the row's first slot.
  std::vector<int> perm((size_t)2 * N * D), slotoff((size_t)2 * N * D);
  for (int j = 0; j < D; ++j)
    for (int w = 0; w < N; ++w)
      for (int part = 0; part < 2; ++part) {
        const size_t o = (size_t)2 * (j * N + w) + part;
        const int e = h.site_of_mode[j] * N + w;  // column-major element
        perm[o] = j * 2 * NE + part * NE + e;
        slotoff[o] = 2 * (j * NE + e) + part;
      }
  int rc = mlp->Pack(&fnet, &pack, perm.data(), (int)perm.size());
  if (rc) {
    err_ = "closure network: " + mlp->last_error();
    return rc == kErrUnsupported ? kErrInvalidArgument : rc;
  }
  const size_t bytes = (size_t)total * NE * 2 * sizeof(double);
  rc = Check(hipMalloc((void**)&rho, bytes), "alloc closure rho");
  if (!rc) rc = Check(hipMemsetAsync(rho, 0, bytes, stream_), "zero closure rho");
  if (!rc && (int)rho0_init_.size() == NE)
    for (int m = 0; m < members_ && !rc; ++m)
      rc = Check(hipMemcpyAsync(rho + (size_t)moff[m] * NE * 2, rho0_init_.data(),
                                NE * sizeof(double2), hipMemcpyHostToDevice, stream_),
                 "initial state");
  if (!rc) rc = Check(Up(&dmoff, moff.data(), moff.size()), "upload offsets");
  if (!rc) rc = BuildEdges(first, nb, vbase, &edge);
  const std::vector<float> bmi = BoundaryIndex(h, first, nb, B);
  const long long nrows = (long long)members_ * nb;
  if (!rc) rc = Check(Up(&cbmi, bmi.data(), bmi.size()), "upload boundary indices");
  if (!rc) rc = Check(Up(&csite, h.site_of_mode.data(), (size_t)D), "upload mode sites");
  if (!rc) rc = Check(Up(&cslot, slotoff.data(), slotoff.size()), "upload slot offsets");
  if (!rc) rc = Check(hipMalloc((void**)&cx, (size_t)nrows * mlp->n_in() * 4), "alloc X");
  if (!rc) rc = Check(hipMalloc((void**)&cs, (size_t)nrows * 4), "alloc S");
  if (!rc) rc = Check(hipMalloc((void**)&cy, (size_t)nrows * mlp->n_out() * 4), "alloc Y");
  if (!rc) rc = Check(hipMalloc((void**)&cfall, sizeof(unsigned long long)), "alloc fallback count");
  if (!rc) rc = Check(hipMemsetAsync(cfall, 0, sizeof(unsigned long long), stream_), "zero count");
  if (!rc) rc = Synchronize();
  if (rc) {
    void* ptrs[] = {pack, rho, dmoff, edge, cbmi, csite, cslot, cx, cs, cy, cfall};
    for (void* q : ptrs) FreeD(q);
    return rc;
  }
  // Commit: the new slab (each member's nodes followed by its virtual
  // slots) replaces the old one; stage buffers are reallocated on demand.
  DropGraph();
  FreeD(d_rho_);
  d_rho_ = rho;
  FreeD(d_aux_);
  d_aux_ = nullptr;
  FreeD(d_y_[0]);
  FreeD(d_y_[1]);
  d_y_[0] = d_y_[1] = nullptr;
  for (auto& k : d_k_) {
    FreeD(k);
    k = nullptr;
  }
  FreeD(d_tmp_);
  d_tmp_ = nullptr;
  bulk_alloc_ = false;
  rkf45_alloc_ = false;
  FreeD(d_moff_);
  d_moff_ = dmoff;
  FreeD(d_edge_);
  d_edge_ = edge;
  args_.edge = d_edge_;
  void* old[] = {d_cbmi_, d_csite_, d_cslot_, d_cx_, d_cs_, d_cy_, d_cpack_, d_cfall_};
  for (void* q : old) FreeD(q);
  d_cfall_ = cfall;
  d_cbmi_ = cbmi;
  d_csite_ = csite;
  d_cslot_ = cslot;
  d_cx_ = cx;
  d_cs_ = cs;
  d_cy_ = cy;
  d_cpack_ = pack;
  fused_ = fnet;
  moff_ = moff;
  mcount_ = mcount;
  slab_nodes_ = (int)total;
  args_.num_nodes = slab_nodes_;
  closure_ = mlp;
  closure_first_ = first;
  closure_nb_ = nb;
  closure_vbase_ = vbase;
  tiles_for_ = 0;
  return EnsureTiles();
}

// Predictions for the virtual slots of the stage input `in` (one fused
// launch; the unfused path, kept for reference and for networks whose
// activations do not fit one CU's LDS, runs the inputs kernel, the
// network's GEMMs and the scatter kernel).
int Engine::ClosurePredict(const void* in, bool in_fp32) {
  const int NE = h_.N * h_.N, D = h_.D;
  const ClosureGuard guard{closure_ymax_, d_cfall_};
  if (closure_fused_ && FusedClosureFits(fused_) && NE <= 64 && D <= 64)
    return Check(LaunchClosureFused(fused_, const_cast<void*>(in), in_fp32, d_moff_, members_,
                                    closure_first_, closure_nb_, NE, D, d_cbmi_, closure_vbase_,
                                    d_cslot_, guard, stream_), "closure fused");
  const long long nrows = (long long)members_ * closure_nb_;
  int rc = Check(LaunchClosureInputs(in, in_fp32, d_moff_, members_, closure_first_, closure_nb_,
                                     NE, D, d_cbmi_, d_cx_, d_cs_, stream_), "closure inputs");
  if (!rc) {
    rc = closure_->ForwardDevice(d_cx_, (int)nrows, d_cy_, stream_);
    if (rc) err_ = "closure network: " + closure_->last_error();
  }
  if (!rc)
    rc = Check(LaunchClosureScatter(d_cy_, d_cs_, d_moff_, members_, closure_nb_, h_.N, D,
                                    d_csite_, closure_vbase_, const_cast<void*>(in), in_fp32,
                                    guard, stream_),
               "closure scatter");
  return rc;
}

int Engine::ClosureFallback(double ymax, long long* fallen) {
  if (!closure_) {
    err_ = "no closure is attached";
    return kErrInvalidArgument;
  }
  if (ymax > 0) {
    closure_ymax_ = (float)std::min(ymax, 3.0e38);
    DropGraph();  // the threshold is a launch argument
  }
  if (!fallen) return kOk;
  unsigned long long n = 0;
  int rc = Check(hipMemcpyAsync(&n, d_cfall_, sizeof(n), hipMemcpyDeviceToHost, stream_),
                 "fallback count");
  if (!rc) rc = Synchronize();
  *fallen = (long long)n;
  return rc;
}

}  // namespace phi
