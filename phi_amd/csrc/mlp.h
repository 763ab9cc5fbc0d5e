// The truncation-closure network on the device: a 3-layer perceptron
// (n_in -> hidden -> hidden -> n_out, GELU) with bfloat16 GEMM operands and
// float32 accumulation, trained by Adam on the squared error.  Parameters
// are one flat float32 array: W1 (hidden x n_in, row-major), b1, W2, b2,
// W3 (n_out x hidden), b3.
#ifndef PHI_AMD_MLP_H_
#define PHI_AMD_MLP_H_

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mlp_kernels.h"

namespace phi {

class Mlp {
 public:
  static int Create(int n_in, int n_hidden, int n_out, int max_batch, Mlp** out,
                    std::string* why);
  ~Mlp();

  int n_in() const { return n_in_; }
  int n_hidden() const { return n_hid_; }
  int n_out() const { return n_out_; }
  int max_batch() const { return max_batch_; }
  long long param_count() const { return nparam_; }
  const std::string& last_error() const { return err_; }
  hipStream_t stream() const { return stream_; }

  int SetParams(const float* flat);
  int GetParams(float* flat);

  // y = f(x) for host arrays (batch <= max_batch).
  int Forward(const float* x, int batch, float* y);
  // Loss mean((f(x) - t)^2) and its gradient (flat, parameter layout).
  int Gradients(const float* x, const float* t, int batch, double* loss, float* grads);
  // Adam on host data: the samples are copied to the device once, then
  // `epochs` passes over shuffled minibatches (seeded permutation), the
  // learning rate going geometrically from lr to lr_final over the epochs
  // (lr_final <= 0: constant).  epoch_loss[e] = mean batch loss of epoch e.
  int Train(const float* X, const float* T, long long samples, int batch, int epochs, float lr,
            float lr_final, unsigned seed, double* epoch_loss);

  // Forward on device arrays: x (batch x n_in) -> y (batch x n_out), on the
  // given stream; activations of the first max_batch rows live in this
  // object's buffers, larger batches run in chunks.
  int ForwardDevice(const float* x, int batch, float* y, hipStream_t s);
  // A snapshot of the current parameters packed for the fused closure pass
  // (bfloat16 weights and float32 biases in one device buffer, *owned, which
  // the caller frees with hipFree; later training does not change it).
  // out_perm (n_packed entries, or null for all n_out in order): packed
  // output neuron o is neuron out_perm[o] of the network (a selection and
  // relabelling of the last layer's rows; n_packed <= 0 means n_out).
  int Pack(struct FusedNet* out, void** owned, const int* out_perm = nullptr,
           int n_packed = 0);

 private:
  Mlp() = default;
  int Check(hipError_t e, const char* where);
  int ForwardBatch(const float* x, int batch, float* y, hipStream_t s);
  int BackwardBatch(const float* x, const float* t, int batch, int partial_slot);
  int AdamStep(float lr);

  int n_in_ = 0, n_hid_ = 0, n_out_ = 0, max_batch_ = 0;
  long long nparam_ = 0;
  long long off_[6] = {0, 0, 0, 0, 0, 0};  // W1, b1, W2, b2, W3, b3
  hipStream_t stream_ = nullptr;
  std::string err_;
  int adam_t_ = 0;
  float* d_p_ = nullptr;   // parameters
  float* d_g_ = nullptr;   // gradients
  float* d_m_ = nullptr;   // Adam first moment
  float* d_v_ = nullptr;   // Adam second moment
  float *d_x_ = nullptr, *d_t_ = nullptr;  // staged batch
  float *d_z1_ = nullptr, *d_h1_ = nullptr, *d_z2_ = nullptr, *d_h2_ = nullptr, *d_y_ = nullptr;
  float *d_dy_ = nullptr, *d_dh2_ = nullptr, *d_dh1_ = nullptr;
  double* d_partial_ = nullptr;
  int partial_blocks_ = 64;
  int partial_slots_ = 0;
};

}  // namespace phi

#endif  // PHI_AMD_MLP_H_
