// Element functions of the closure network, shared by the training kernels
// and the fused inference pass so that both evaluate the same function.
#ifndef PHI_AMD_MLP_ACT_H_
#define PHI_AMD_MLP_ACT_H_

#include <hip/hip_runtime.h>

namespace phi {
namespace act {

// float32 -> bfloat16 bits, round to nearest even (one v_cvt_pk_bf16_f32
// per pair on gfx950).
__device__ __forceinline__ unsigned short to_bf16(float f) {
  return __builtin_bit_cast(unsigned short, (__bf16)f);
}

// erf by Abramowitz and Stegun, Handbook of Mathematical Functions, formula
// 7.1.26: for x >= 0, erf(x) = 1 - (a1 t + a2 t^2 + a3 t^3 + a4 t^4 +
// a5 t^5) exp(-x^2) with t = 1 / (1 + p x), absolute error <= 1.5e-7 in
// exact arithmetic (5.3e-7 in float32 over [-10, 10], largest near 0 where
// 1 - p e cancels; GELU's error there is z/2 times that, far below the
// bfloat16 rounding of its result); extended as an odd function.  Branch
// free (one reciprocal, one exp2): the
// library erff takes a different polynomial on each side of |x| = 1, and a
// wavefront with lanes on both sides runs both.
__device__ __forceinline__ float erf_as(float x) {
  const float ax = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, ax, 1.0f));
  float p = fmaf(1.061405429f, t, -1.453152027f);
  p = fmaf(p, t, 1.421413741f);
  p = fmaf(p, t, -0.284496736f);
  p = fmaf(p, t, 0.254829592f);
  p *= t;
  const float sq = ax * ax;
  const float e = __builtin_amdgcn_exp2f(sq * -1.44269504088896341f);  // exp(-x^2)
  return copysignf(fmaf(-p, e, 1.0f), x);
}

// GELU(z) = z/2 (1 + erf(z / sqrt 2)).
__device__ __forceinline__ float gelu(float z) {
  const float h = 0.5f * z;
  return fmaf(h, erf_as(z * 0.70710678118654752f), h);
}

// GELU'(z) = (1 + erf(z / sqrt 2)) / 2 + z exp(-z^2 / 2) / sqrt(2 pi).
__device__ __forceinline__ float gelu_grad(float z) {
  const float sq = z * z;
  const float e = __builtin_amdgcn_exp2f(sq * -0.72134752044448170f);  // exp(-z^2/2)
  return fmaf(0.5f, erf_as(z * 0.70710678118654752f), 0.5f) + z * (e * 0.39894228040143268f);
}

}  // namespace act
}  // namespace phi

#endif  // PHI_AMD_MLP_ACT_H_
