// Shared host/device definitions of libcorrct_hip (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/corrct_hip.h"

namespace chip {

// ---------------------------------------------------------------------------
// error handling: per-thread last-error string, status codes
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

#define CHIP_CHECK_ARG(cond, msg)                 \
  do {                                            \
    if (!(cond)) {                                \
      ::chip::set_error(std::string(__func__) + ": " + (msg)); \
      return CHIP_EINVAL;                         \
    }                                             \
  } while (0)

#define CHIP_HIP(call)                                                                    \
  do {                                                                                    \
    hipError_t err_ = (call);                                                             \
    if (err_ != hipSuccess) {                                                             \
      ::chip::set_error(std::string(__func__) + ": " #call " -> " + hipGetErrorString(err_)); \
      return CHIP_EHIP;                                                                   \
    }                                                                                     \
  } while (0)

#define CHIP_LAUNCH_CHECK() CHIP_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// per-rotation affine inverse map of an order-1 warp (float32, skimage order)
//   X = (m00 * x + m01 * y) + m02,  Y = (m10 * x + m11 * y) + m12
// ---------------------------------------------------------------------------
struct Affine {
  float m00, m01, m02, m10, m11, m12;
};

// Host-side geometry plan. Device tables are allocated at plan creation.
struct Plan {
  int n = 0;          // slice side N (square), also detector pixels
  int n_angles = 0;   // A
  int n_det = 0;      // D
  int pad = 0;        // attenuation pad width e
  int np = 0;         // padded side N + 2e
  // fp geometry per angle: (cos, sin, m02, m12) of the radon inverse map
  std::vector<float> h_fp_geom;  // [A][4]
  std::vector<float> h_w;        // [A][D] = w_det[d] * w_ang[a][d]
  std::vector<Affine> h_rot_in_fwd, h_rot_in_back;    // [A]
  std::vector<Affine> h_rot_out_fwd, h_rot_out_back;  // [A][D]
  float4* d_fp_geom = nullptr;
  float2* d_ytab = nullptr;  // [A][N] (fl(sin * y), fl(cos * y)) row terms of the fp map
  float* d_w = nullptr;
  Affine* d_rot = nullptr;  // [A] in-fwd, [A] in-back, [A*D] out-fwd, [A*D] out-back
  // scratch for the attenuation builder (rotated-frame cumulative sums)
  float* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  int device = 0;
  // kernel selection (chip_plan_tune): fp 0 = gather, 1 = LDS-staged,
  // 2 = LDS-staged with register prefetch, 3 = prefetched and lean
  // (2 and 3 fall back to 1 when N % 4 != 0)
  int fp_variant = 3;
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// whole-sinogram fp over detectors [det0, det0 + nd), nd in {1, 2}; needs
// N % 4 == 0 and 16-byte aligned vol / maps (fp_stream.hip)
void launch_fp_stream(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch,
                      int det0, int nd, hipStream_t st);

// whole-sinogram exact adjoint of the fp, all detectors (bp_adjoint.hip)
void launch_bp_adjoint_tiles(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol,
                             int batch, hipStream_t st);

constexpr int kMaxDetPerPass = 4;

}  // namespace chip

struct chip_plan {
  chip::Plan p;
};
