// Launchers of the closure data-movement kernels (closure_kernels.hip).
#ifndef PHI_AMD_CLOSURE_KERNELS_H_
#define PHI_AMD_CLOSURE_KERNELS_H_

#include <hip/hip_runtime.h>

namespace phi {

// The closure network packed for the fused inference pass: weights as
// bfloat16 rows (out x K, row-major; K = k1 for layer 1 and hid for layers
// 2-3, rows padded to multiples of 256 with zeros), biases float32.
struct FusedNet {
  const unsigned short* w1;
  const unsigned short* w2;
  const unsigned short* w3;
  const float* b1;
  const float* b2;
  const float* b3;
  int k1;     // inputs rounded up to a multiple of 64
  int hid;    // hidden width, a multiple of 32
  int n_out;
  int width;  // max(k1, hid): row width of the LDS activations
};

size_t FusedClosureLds(const FusedNet& net);
// Whether the fused pass's LDS fits one CU (hidden widths up to 256).
bool FusedClosureFits(const FusedNet& net);
// The fallback to plain truncation: a row whose outputs are not all finite
// and at most ymax in magnitude (they are the children's entries over the
// parent's largest entry) gets zero slots, i.e. PHI's truncation for that
// boundary node; fallen counts such rows.
struct ClosureGuard {
  float ymax;
  unsigned long long* fallen;
};

// Inputs, the three layers and the scatter in one launch (128 rows per block);
// slotoff[o] = float offset of packed output o from a row's first slot.
hipError_t LaunchClosureFused(const FusedNet& net, void* slab, bool fp32, const int* moff,
                              int members, int first, int nb, int ne, int D, const float* bmi,
                              int vbase, const int* slotoff, const ClosureGuard& guard,
                              hipStream_t s);

// Network inputs X (rows x (2 N^2 + D)) and scales S of the boundary nodes
// first .. first + nb - 1 of every member, from a stage input in FP32 or
// FP64; bmi = the boundary multi-indices / L (nb x D).
hipError_t LaunchClosureInputs(const void* in, bool in_fp32, const int* moff, int members,
                               int first, int nb, int ne, int D, const float* bmi, float* X,
                               float* S, hipStream_t s);
// Training targets T (rows x 2 N^2 D) from a deeper hierarchy's state.
hipError_t LaunchClosureTargets(const double* rho, const int* moff, int members, int nb, int ne,
                                int D, const int* children, const float* S, float* T,
                                hipStream_t s);
// Network outputs Y (rows x 2 N^2 D) -> column site[j] of the virtual
// neighbour slots vbase + b*D + j of each member (the column the stage
// kernels read).
hipError_t LaunchClosureScatter(const float* Y, const float* S, const int* moff, int members,
                                int nb, int N, int D, const int* site, int vbase, void* out,
                                bool out_fp32, const ClosureGuard& guard, hipStream_t s);

}  // namespace phi

#endif  // PHI_AMD_CLOSURE_KERNELS_H_
