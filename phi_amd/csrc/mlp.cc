// Host side of the truncation-closure network.  See mlp.h.
#include "mlp.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>

#include "closure_kernels.h"

namespace phi {

namespace {

void FreeDev(void* p) {
  if (p) (void)hipFree(p);
}

}  // namespace

int Mlp::Check(hipError_t e, const char* where) {
  if (e == hipSuccess) return 0;
  err_ = std::string(where) + ": " + hipGetErrorString(e);
  return -4;
}

int Mlp::Create(int n_in, int n_hidden, int n_out, int max_batch, Mlp** out, std::string* why) {
  *out = nullptr;
  if (n_in < 1 || n_hidden < 1 || n_out < 1 || max_batch < 1) {
    *why = "network sizes and batch must be positive";
    return -1;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    *why = "no HIP device is visible";
    return -7;
  }
  Mlp* m = new Mlp;
  m->n_in_ = n_in;
  m->n_hid_ = n_hidden;
  m->n_out_ = n_out;
  m->max_batch_ = max_batch;
  const long long sizes[6] = {(long long)n_hidden * n_in, n_hidden, (long long)n_hidden * n_hidden,
                              n_hidden, (long long)n_out * n_hidden, n_out};
  long long o = 0;
  for (int i = 0; i < 6; ++i) {
    m->off_[i] = o;
    o += sizes[i];
  }
  m->nparam_ = o;
  const size_t pb = (size_t)o * sizeof(float);
  const size_t hb = (size_t)max_batch * n_hidden * sizeof(float);
  int rc = m->Check(hipStreamCreateWithFlags(&m->stream_, hipStreamNonBlocking), "stream");
  float** bufs[] = {&m->d_p_, &m->d_g_, &m->d_m_, &m->d_v_};
  for (float** b : bufs)
    if (!rc) rc = m->Check(hipMalloc((void**)b, pb), "alloc params");
  if (!rc) rc = m->Check(hipMemset(m->d_m_, 0, pb), "zero m");
  if (!rc) rc = m->Check(hipMemset(m->d_v_, 0, pb), "zero v");
  if (!rc) rc = m->Check(hipMemset(m->d_p_, 0, pb), "zero p");
  if (!rc) rc = m->Check(hipMalloc((void**)&m->d_x_, (size_t)max_batch * n_in * 4), "alloc x");
  if (!rc) rc = m->Check(hipMalloc((void**)&m->d_t_, (size_t)max_batch * n_out * 4), "alloc t");
  float** acts[] = {&m->d_z1_, &m->d_h1_, &m->d_z2_, &m->d_h2_, &m->d_dh2_, &m->d_dh1_};
  for (float** b : acts)
    if (!rc) rc = m->Check(hipMalloc((void**)b, hb), "alloc activations");
  if (!rc) rc = m->Check(hipMalloc((void**)&m->d_y_, (size_t)max_batch * n_out * 4), "alloc y");
  if (!rc) rc = m->Check(hipMalloc((void**)&m->d_dy_, (size_t)max_batch * n_out * 4), "alloc dy");
  m->partial_slots_ = 1;
  if (!rc)
    rc = m->Check(hipMalloc((void**)&m->d_partial_, m->partial_blocks_ * sizeof(double)),
                  "alloc partial");
  if (rc) {
    *why = m->err_;
    delete m;
    return rc;
  }
  *out = m;
  return 0;
}

Mlp::~Mlp() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  void* ptrs[] = {d_p_, d_g_, d_m_, d_v_, d_x_, d_t_, d_z1_, d_h1_, d_z2_, d_h2_, d_y_,
                  d_dy_, d_dh2_, d_dh1_, d_partial_};
  for (void* p : ptrs) FreeDev(p);
  if (stream_) (void)hipStreamDestroy(stream_);
}

namespace {

// float32 -> bfloat16 bits, round to nearest even (as the device kernels).
unsigned short Bf16(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

}  // namespace

int Mlp::Pack(FusedNet* out, void** owned, const int* out_perm, int n_packed) {
  *owned = nullptr;
  if (n_hid_ % 32 != 0) {
    err_ = "the fused closure pass needs a hidden width that is a multiple of 32";
    return -3;
  }
  if (!out_perm || n_packed <= 0) n_packed = n_out_;
  std::vector<float> p(nparam_);
  int rc = GetParams(p.data());
  if (rc) return rc;
  const int k1 = (n_in_ + 63) / 64 * 64;
  const int hid = n_hid_;
  // Rows padded to multiples of 256: the fused pass streams whole groups of
  // 256 output columns.
  auto rows256 = [](int n) { return (n + 255) / 256 * 256; };
  const size_t n1 = (size_t)rows256(hid) * k1, n2 = (size_t)rows256(hid) * hid,
               n3 = (size_t)rows256(n_packed) * hid;
  // One buffer: the bfloat16 weights, then (16-byte aligned) the float32
  // biases b1, b2, b3 (b3 zero padded like W3's rows: the fused pass reads
  // the biases of 4 output columns at once).
  const size_t nw = (n1 + n2 + n3 + 7) / 8 * 8;
  const size_t nb = (size_t)2 * hid + rows256(n_packed);
  std::vector<unsigned short> w(nw + 2 * nb, 0);
  for (int o = 0; o < hid; ++o)
    for (int i = 0; i < n_in_; ++i) w[(size_t)o * k1 + i] = Bf16(p[off_[0] + (size_t)o * n_in_ + i]);
  for (int o = 0; o < hid; ++o)
    for (int i = 0; i < hid; ++i) w[n1 + (size_t)o * hid + i] = Bf16(p[off_[2] + (size_t)o * hid + i]);
  for (int o = 0; o < n_packed; ++o) {
    const int src = out_perm ? out_perm[o] : o;
    if (src < 0 || src >= n_out_) {
      err_ = "output permutation entry out of range";
      return -3;
    }
    for (int i = 0; i < hid; ++i)
      w[n1 + n2 + (size_t)o * hid + i] = Bf16(p[off_[4] + (size_t)src * hid + i]);
  }
  float* bias = reinterpret_cast<float*>(w.data() + nw);
  std::memcpy(bias, p.data() + off_[1], hid * sizeof(float));
  std::memcpy(bias + hid, p.data() + off_[3], hid * sizeof(float));
  for (int o = 0; o < n_packed; ++o) bias[2 * hid + o] = p[off_[5] + (out_perm ? out_perm[o] : o)];
  unsigned short* d = nullptr;
  rc = Check(hipMalloc((void**)&d, w.size() * sizeof(unsigned short)), "alloc pack");
  if (!rc) rc = Check(hipMemcpy(d, w.data(), w.size() * sizeof(unsigned short),
                                hipMemcpyHostToDevice), "upload pack");
  if (rc) {
    FreeDev(d);
    return rc;
  }
  const float* db = reinterpret_cast<const float*>(d + nw);
  out->w1 = d;
  out->w2 = d + n1;
  out->w3 = d + n1 + n2;
  out->b1 = db;
  out->b2 = db + hid;
  out->b3 = db + 2 * hid;
  out->k1 = k1;
  out->hid = hid;
  out->n_out = n_packed;
  out->width = std::max(k1, hid);
  *owned = d;
  return 0;
}

int Mlp::SetParams(const float* flat) {
  int rc = Check(hipMemcpyAsync(d_p_, flat, nparam_ * sizeof(float), hipMemcpyHostToDevice,
                                stream_), "set params");
  if (!rc) rc = Check(hipStreamSynchronize(stream_), "sync");
  adam_t_ = 0;
  if (!rc) rc = Check(hipMemset(d_m_, 0, nparam_ * sizeof(float)), "zero m");
  if (!rc) rc = Check(hipMemset(d_v_, 0, nparam_ * sizeof(float)), "zero v");
  return rc;
}

int Mlp::GetParams(float* flat) {
  int rc = Check(hipMemcpyAsync(flat, d_p_, nparam_ * sizeof(float), hipMemcpyDeviceToHost,
                                stream_), "get params");
  return rc ? rc : Check(hipStreamSynchronize(stream_), "sync");
}

// The three layers for `batch` rows of x; activations kept for backward.
int Mlp::ForwardBatch(const float* x, int batch, float* y, hipStream_t s) {
  mlp::GemmArgs g{};
  g.M = batch;
  g.N = n_hid_;
  g.K = n_in_;
  g.A = x;
  g.sa_i = n_in_;
  g.sa_k = 1;
  g.B = d_p_ + off_[0];
  g.sb_k = 1;
  g.sb_j = n_in_;
  g.C = d_h1_;
  g.sc_i = n_hid_;
  g.bias = d_p_ + off_[1];
  g.Z = d_z1_;
  g.gelu = 1;
  int rc = Check(LaunchGemmBf16(g, s), "layer 1");
  g.K = n_hid_;
  g.A = d_h1_;
  g.sa_i = n_hid_;
  g.B = d_p_ + off_[2];
  g.sb_j = n_hid_;
  g.C = d_h2_;
  g.bias = d_p_ + off_[3];
  g.Z = d_z2_;
  if (!rc) rc = Check(LaunchGemmBf16(g, s), "layer 2");
  g.N = n_out_;
  g.A = d_h2_;
  g.B = d_p_ + off_[4];
  g.sb_j = n_hid_;
  g.C = y;
  g.sc_i = n_out_;
  g.bias = d_p_ + off_[5];
  g.Z = nullptr;
  g.gelu = 0;
  if (!rc) rc = Check(LaunchGemmBf16(g, s), "layer 3");
  return rc;
}

int Mlp::ForwardDevice(const float* x, int batch, float* y, hipStream_t s) {
  int rc = 0;
  for (int r0 = 0; r0 < batch && !rc; r0 += max_batch_) {
    const int b = std::min(max_batch_, batch - r0);
    rc = ForwardBatch(x + (size_t)r0 * n_in_, b, y + (size_t)r0 * n_out_, s);
  }
  return rc;
}

int Mlp::Forward(const float* x, int batch, float* y) {
  if (batch < 1 || batch > max_batch_) {
    err_ = "batch outside 1 .. max_batch";
    return -1;
  }
  int rc = Check(hipMemcpyAsync(d_x_, x, (size_t)batch * n_in_ * 4, hipMemcpyHostToDevice,
                                stream_), "upload x");
  if (!rc) rc = ForwardBatch(d_x_, batch, d_y_, stream_);
  if (!rc) rc = Check(hipMemcpyAsync(y, d_y_, (size_t)batch * n_out_ * 4, hipMemcpyDeviceToHost,
                                     stream_), "download y");
  if (!rc) rc = Check(hipStreamSynchronize(stream_), "sync");
  return rc;
}

// Gradients of the batch in d_x_/d_t_ into d_g_; the loss partial sums of
// the batch go to partial slot `slot`.
int Mlp::BackwardBatch(const float* x, const float* t, int batch, int slot) {
  hipStream_t s = stream_;
  int rc = ForwardBatch(x, batch, d_y_, s);
  const long long ny = (long long)batch * n_out_;
  if (!rc)
    rc = Check(LaunchMseGrad(d_y_, t, d_dy_, ny, d_partial_ + (size_t)slot * partial_blocks_,
                             partial_blocks_, s), "loss");
  // Layer 3: dW3 = dY^T H2, db3 = colsum dY, dH2 = dY W3.
  mlp::GemmArgs g{};
  auto weight_grad = [&](const float* dz, int nout, const float* h, int nin, long long woff) {
    mlp::GemmArgs w{};
    w.M = nout;
    w.N = nin;
    w.K = batch;
    w.A = dz;
    w.sa_i = 1;
    w.sa_k = nout;
    w.B = h;
    w.sb_k = nin;
    w.sb_j = 1;
    w.C = d_g_ + woff;
    w.sc_i = nin;
    return Check(LaunchGemmBf16(w, s), "weight grad");
  };
  auto input_grad = [&](const float* dz, int nout, long long woff, int nin, float* dh) {
    mlp::GemmArgs a{};
    a.M = batch;
    a.N = nin;
    a.K = nout;
    a.A = dz;
    a.sa_i = nout;
    a.sa_k = 1;
    a.B = d_p_ + woff;
    a.sb_k = nin;
    a.sb_j = 1;
    a.C = dh;
    a.sc_i = nin;
    return Check(LaunchGemmBf16(a, s), "input grad");
  };
  (void)g;
  if (!rc) rc = weight_grad(d_dy_, n_out_, d_h2_, n_hid_, off_[4]);
  if (!rc) rc = Check(LaunchColumnSums(d_dy_, batch, n_out_, n_out_, d_g_ + off_[5], s), "db3");
  if (!rc) rc = input_grad(d_dy_, n_out_, off_[4], n_hid_, d_dh2_);
  // Layer 2.
  const long long nh = (long long)batch * n_hid_;
  if (!rc) rc = Check(LaunchGeluBackward(d_dh2_, d_z2_, d_dh2_, nh, s), "gelu 2");
  if (!rc) rc = weight_grad(d_dh2_, n_hid_, d_h1_, n_hid_, off_[2]);
  if (!rc) rc = Check(LaunchColumnSums(d_dh2_, batch, n_hid_, n_hid_, d_g_ + off_[3], s), "db2");
  if (!rc) rc = input_grad(d_dh2_, n_hid_, off_[2], n_hid_, d_dh1_);
  // Layer 1.
  if (!rc) rc = Check(LaunchGeluBackward(d_dh1_, d_z1_, d_dh1_, nh, s), "gelu 1");
  if (!rc) rc = weight_grad(d_dh1_, n_hid_, x, n_in_, off_[0]);
  if (!rc) rc = Check(LaunchColumnSums(d_dh1_, batch, n_hid_, n_hid_, d_g_ + off_[1], s), "db1");
  return rc;
}

int Mlp::Gradients(const float* x, const float* t, int batch, double* loss, float* grads) {
  if (batch < 1 || batch > max_batch_) {
    err_ = "batch outside 1 .. max_batch";
    return -1;
  }
  int rc = Check(hipMemcpyAsync(d_x_, x, (size_t)batch * n_in_ * 4, hipMemcpyHostToDevice,
                                stream_), "upload x");
  if (!rc) rc = Check(hipMemcpyAsync(d_t_, t, (size_t)batch * n_out_ * 4, hipMemcpyHostToDevice,
                                     stream_), "upload t");
  if (!rc) rc = BackwardBatch(d_x_, d_t_, batch, 0);
  std::vector<double> part(partial_blocks_);
  if (!rc) rc = Check(hipMemcpyAsync(part.data(), d_partial_, part.size() * sizeof(double),
                                     hipMemcpyDeviceToHost, stream_), "loss");
  if (!rc) rc = Check(hipMemcpyAsync(grads, d_g_, nparam_ * sizeof(float), hipMemcpyDeviceToHost,
                                     stream_), "download grads");
  if (!rc) rc = Check(hipStreamSynchronize(stream_), "sync");
  if (!rc) {
    double s = 0;
    for (double v : part) s += v;
    *loss = s / ((double)batch * n_out_);
  }
  return rc;
}

int Mlp::AdamStep(float lr) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  ++adam_t_;
  const double c1 = 1.0 - std::pow((double)b1, adam_t_);
  const double c2 = 1.0 - std::pow((double)b2, adam_t_);
  const float lr_t = (float)(lr * std::sqrt(c2) / c1);
  const float eps_t = (float)(eps * std::sqrt(c2));
  return Check(LaunchAdam(d_p_, d_g_, d_m_, d_v_, nparam_, lr_t, b1, b2, eps_t, stream_), "adam");
}

int Mlp::Train(const float* X, const float* T, long long samples, int batch, int epochs,
               float lr, float lr_final, unsigned seed, double* epoch_loss) {
  if (samples < 1 || batch < 1 || batch > max_batch_ || epochs < 1) {
    err_ = "bad training sizes";
    return -1;
  }
  float *dX = nullptr, *dT = nullptr;
  int* dIdx = nullptr;
  const long long nb = samples / batch;  // full batches per epoch
  if (nb < 1) {
    err_ = "fewer samples than one batch";
    return -1;
  }
  int rc = Check(hipMalloc((void**)&dX, (size_t)samples * n_in_ * 4), "alloc data x");
  if (!rc) rc = Check(hipMalloc((void**)&dT, (size_t)samples * n_out_ * 4), "alloc data t");
  if (!rc) rc = Check(hipMalloc((void**)&dIdx, (size_t)samples * sizeof(int)), "alloc index");
  if (!rc) rc = Check(hipMemcpyAsync(dX, X, (size_t)samples * n_in_ * 4, hipMemcpyHostToDevice,
                                     stream_), "upload data x");
  if (!rc) rc = Check(hipMemcpyAsync(dT, T, (size_t)samples * n_out_ * 4, hipMemcpyHostToDevice,
                                     stream_), "upload data t");
  // One partial-sum slot per batch of an epoch.
  if (!rc && partial_slots_ < nb) {
    FreeDev(d_partial_);
    d_partial_ = nullptr;
    rc = Check(hipMalloc((void**)&d_partial_, (size_t)nb * partial_blocks_ * sizeof(double)),
               "alloc partials");
    partial_slots_ = rc ? 0 : (int)nb;
  }
  std::mt19937 gen(seed);
  std::vector<int> perm(samples);
  std::vector<double> part((size_t)nb * partial_blocks_);
  for (int ep = 0; ep < epochs && !rc; ++ep) {
    const float lr_ep = (lr_final > 0 && epochs > 1)
                            ? (float)(lr * std::pow((double)lr_final / lr, (double)ep / (epochs - 1)))
                            : lr;
    std::iota(perm.begin(), perm.end(), 0);
    std::shuffle(perm.begin(), perm.end(), gen);
    rc = Check(hipMemcpyAsync(dIdx, perm.data(), perm.size() * sizeof(int),
                              hipMemcpyHostToDevice, stream_), "upload permutation");
    for (long long b = 0; b < nb && !rc; ++b) {
      rc = Check(LaunchGatherRows(dX, dIdx + b * batch, batch, n_in_, d_x_, stream_), "gather x");
      if (!rc) rc = Check(LaunchGatherRows(dT, dIdx + b * batch, batch, n_out_, d_t_, stream_),
                          "gather t");
      if (!rc) rc = BackwardBatch(d_x_, d_t_, batch, (int)b);
      if (!rc) rc = AdamStep(lr_ep);
    }
    if (!rc) rc = Check(hipMemcpyAsync(part.data(), d_partial_, part.size() * sizeof(double),
                                       hipMemcpyDeviceToHost, stream_), "epoch loss");
    if (!rc) rc = Check(hipStreamSynchronize(stream_), "sync");
    if (!rc && epoch_loss) {
      double s = 0;
      for (double v : part) s += v;
      epoch_loss[ep] = s / ((double)nb * batch * n_out_);
    }
  }
  if (stream_) (void)hipStreamSynchronize(stream_);
  FreeDev(dX);
  FreeDev(dT);
  FreeDev(dIdx);
  return rc;
}

}  // namespace phi
