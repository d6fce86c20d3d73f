// Per-element data steps of the solvers, shared by the elementwise kernels
// (solver_ops.hip) and the per-angle-block data step of the fused on-the-fly
// iteration (attenuation.hip), so both produce the same float32 bits.
// a = ax (+ bg), s = sigma, b = data, m = mask (1 when absent).
#pragma once

#include <cfloat>

namespace chip {

// DataFidelity_l2 update_dual + apply_proximal_dual (data_terms.py:217-230, 294-304)
__device__ __forceinline__ float elem_dual_l2(float p, float a, float s, float b) {
  float v = p + a * s;
  v = v - s * b;
  return v * (1.f / (1.f + s));
}

// DataFidelity_KL (data_terms.py:677-692)
__device__ __forceinline__ float elem_dual_kl(float p, float a, float s, float b) {
  const float v = p + a * s;
  const float sd = (4.f * s) * fmaxf(b, 0.f);
  const float dm = v - 1.f;
  return ((1.f + v) - sqrtf(dm * dm + sd)) / 2.f;
}

// SIRT residual (b - a) [* mask] [* sigma] (data_terms.py:136-161, solvers.py:954, 975); the mask and
// sigma products are applied by the caller in that order
__device__ __forceinline__ float elem_residual(float a, float b) { return b - a; }

// MLEM ratio b / max(a, eps) (solvers.py:793-797)
__device__ __forceinline__ float elem_mlem_ratio(float a, float b) { return b / fmaxf(a, FLT_EPSILON); }

}  // namespace chip
