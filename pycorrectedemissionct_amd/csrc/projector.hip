// Forward / back projection kernels of libcorrct_hip (gfx950).
//
// fp  -- radon discretisation of scikit-image 0.18.3 as used by corrct's
//        skimage backend (_projector_backends.py:312-333): sample (y, u) of
//        angle a reads the slice at
//          x' = (c u + s y) + m02,  y' = (-s u + c y) + m12   (float32, no FMA)
//        with an order-1 blend (0 outside), and the line is the sum over y.
//        The XRF form multiplies the slice by the map M[a, d] and the line by
//        w[a, d] (projectors.py:519-525).
// bp  -- (a) skimage: per angle, np.interp of the line at
//        t = (j - k) c - (i - k) s, 0 outside [-k, N-1-k] (iradon without
//        filter, _projector_backends.py:335-359), summed over angles and
//        detectors; optionally weighted by w[a, d] M[a, d] (is_symmetric);
//        (b) adjoint: the exact transpose of fp, as a gather over the fp
//        samples whose bilinear footprint covers each pixel.
//
// Work decomposition
//   fp: one workgroup = one angle x 32 detector bins; each wave covers an
//       8 (u) x 8 (y) patch of samples so a wave's 64 gathers touch a compact
//       rotated patch of the slice (few cache lines) instead of a 64-pixel
//       diagonal; partial line sums over y are reduced with cross-lane
//       shuffles. All detectors of the angle share the sample coordinates.
//   bp: one thread per pixel, 16 x 16 pixel tiles, loop over angles; the
//       per-angle geometry is wave-uniform (scalar loads).
#include "common.hpp"

namespace chip {

// ---------------------------------------------------------------------------
// fp
// ---------------------------------------------------------------------------

constexpr int kFpBinsPerBlock = 32;  // 4 waves x 8 bins
constexpr int kFpLanesY = 8;

template <int D>
__global__ __launch_bounds__(256) void fp_kernel(const float* __restrict__ vol, const float* __restrict__ maps,
                                                 int64_t maps_slice_stride, float* __restrict__ sino,
                                                 const float4* __restrict__ geom, const float* __restrict__ w,
                                                 int n, int n_angles, int n_det_total, int det0, int angle0,
                                                 int line_mode) {
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  const int du = lane & 7;
  const int dy = lane >> 3;
  const int u = blockIdx.x * kFpBinsPerBlock + wave * 8 + du;
  const int a = angle0 + blockIdx.y;
  const int b = blockIdx.z;

  const float4 g = geom[a];
  const float cs = g.x, sn = g.y, m02 = g.z, m12 = g.w;
  const size_t nn = static_cast<size_t>(n) * n;
  const float* v = vol + static_cast<size_t>(b) * nn;
  const float* mp[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    mp[d] = maps ? maps + static_cast<size_t>(b) * maps_slice_stride +
                       (static_cast<size_t>(a) * n_det_total + det0 + d) * nn
                 : nullptr;
  }

  float acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0.f;

  if (u < n) {
    const float uf = static_cast<float>(u);
    const float xu = __fmul_rn(cs, uf);
    const float yu = __fmul_rn(-sn, uf);
    for (int y = dy; y < n; y += kFpLanesY) {
      const float yf = static_cast<float>(y);
      const float xs = __fadd_rn(__fadd_rn(xu, __fmul_rn(sn, yf)), m02);
      const float ys = __fadd_rn(__fadd_rn(yu, __fmul_rn(cs, yf)), m12);
      if (!(xs > -1.f && xs < static_cast<float>(n) && ys > -1.f && ys < static_cast<float>(n))) continue;
      const float c0f = floorf(xs);
      const float r0f = floorf(ys);
      const float dc = xs - c0f;
      const float dr = ys - r0f;
      const int c0 = static_cast<int>(c0f);
      const int r0 = static_cast<int>(r0f);
      const bool okc0 = c0 >= 0, okc1 = c0 + 1 < n;
      const bool okr0 = r0 >= 0, okr1 = r0 + 1 < n;
      const int i00 = r0 * n + c0;
      const float v00 = (okr0 && okc0) ? v[i00] : 0.f;
      const float v01 = (okr0 && okc1) ? v[i00 + 1] : 0.f;
      const float v10 = (okr1 && okc0) ? v[i00 + n] : 0.f;
      const float v11 = (okr1 && okc1) ? v[i00 + n + 1] : 0.f;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        float t00 = v00, t01 = v01, t10 = v10, t11 = v11;
        if (maps) {
          const float* m = mp[d];
          t00 *= (okr0 && okc0) ? m[i00] : 0.f;
          t01 *= (okr0 && okc1) ? m[i00 + 1] : 0.f;
          t10 *= (okr1 && okc0) ? m[i00 + n] : 0.f;
          t11 *= (okr1 && okc1) ? m[i00 + n + 1] : 0.f;
        }
        const float top = fmaf(dc, t01 - t00, t00);
        const float bot = fmaf(dc, t11 - t10, t10);
        acc[d] += fmaf(dr, bot - top, top);
      }
    }
  }
  // reduce the 8 y-lanes that share a bin (lane bits 3..5)
#pragma unroll
  for (int d = 0; d < D; ++d) {
    float s = acc[d];
    s += __shfl_xor(s, 8);
    s += __shfl_xor(s, 16);
    s += __shfl_xor(s, 32);
    acc[d] = s;
  }
  if (dy == 0 && u < n) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int dd = det0 + d;
      const float wt = w[static_cast<size_t>(a) * n_det_total + dd];
      size_t o;
      if (line_mode) {
        o = static_cast<size_t>(dd) * n + u;  // line[d][u] for a single angle
      } else {
        o = ((static_cast<size_t>(b) * n_det_total + dd) * n_angles + a) * n + u;
      }
      sino[o] = acc[d] * wt;
    }
  }
}

template <int D>
static void launch_fp_group(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino,
                            int batch, int det0, int angle0, int n_ang, int line_mode, hipStream_t st) {
  dim3 grid((p.n + kFpBinsPerBlock - 1) / kFpBinsPerBlock, n_ang, batch);
  hipLaunchKernelGGL(fp_kernel<D>, grid, dim3(256), 0, st, vol, maps, stride, sino, p.d_fp_geom, p.d_w, p.n,
                     p.n_angles, p.n_det, det0, angle0, line_mode);
}

static void launch_fp(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch,
                      int angle0, int n_ang, int line_mode, hipStream_t st) {
  for (int det0 = 0; det0 < p.n_det; det0 += kMaxDetPerPass) {
    const int nd = std::min(kMaxDetPerPass, p.n_det - det0);
    switch (nd) {
      case 1: launch_fp_group<1>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, line_mode, st); break;
      case 2: launch_fp_group<2>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, line_mode, st); break;
      case 3: launch_fp_group<3>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, line_mode, st); break;
      default: launch_fp_group<4>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, line_mode, st); break;
    }
  }
}

// ---------------------------------------------------------------------------
// bp, skimage (iradon) discretisation
// ---------------------------------------------------------------------------

constexpr int kBpTile = 16;

// mode 0: sum over angles and detectors of interp (no maps, no weights)
// mode 1: sum of w[a,d] * M[a,d][i,j] * interp
template <int SYM>
__global__ __launch_bounds__(256) void bp_skimage_kernel(const float* __restrict__ sino, const float* __restrict__ maps,
                                                         int64_t maps_slice_stride, float* __restrict__ vol,
                                                         const float4* __restrict__ geom,
                                                         const float* __restrict__ w, int n, int n_angles,
                                                         int n_det, int angle0, int n_ang, int line_mode,
                                                         int accumulate) {
  const int j = blockIdx.x * kBpTile + (threadIdx.x & (kBpTile - 1));
  const int i = blockIdx.y * kBpTile + (threadIdx.x / kBpTile);
  const int b = blockIdx.z;
  if (i >= n || j >= n) return;
  const int k = n / 2;
  const float jc = static_cast<float>(j - k);
  const float ic = static_cast<float>(i - k);
  const float tlo = static_cast<float>(-k);
  const float thi = static_cast<float>(n - 1 - k);
  const size_t nn = static_cast<size_t>(n) * n;
  const size_t pix = static_cast<size_t>(i) * n + j;
  float acc = 0.f;
  for (int aa = 0; aa < n_ang; ++aa) {
    const int a = angle0 + aa;
    const float4 g = geom[a];
    const float t = fmaf(jc, g.x, -(ic * g.y));
    if (!(t >= tlo && t <= thi)) continue;
    const float tf = floorf(t);
    const float fr = t - tf;
    const int kk = static_cast<int>(tf) + k;  // in [0, n-1]
    const bool ok1 = kk + 1 < n;
    for (int d = 0; d < n_det; ++d) {
      const float* line = line_mode ? sino + static_cast<size_t>(d) * n
                                     : sino + ((static_cast<size_t>(b) * n_det + d) * n_angles + a) * n;
      const float s0 = line[kk];
      const float s1 = ok1 ? line[kk + 1] : 0.f;
      float val = fmaf(fr, s1 - s0, s0);
      if (SYM) {
        float f = w[static_cast<size_t>(a) * n_det + d];
        if (maps) f *= maps[static_cast<size_t>(b) * maps_slice_stride + (static_cast<size_t>(a) * n_det + d) * nn + pix];
        val *= f;
      }
      acc += val;
    }
  }
  float* o = vol + static_cast<size_t>(b) * nn + pix;
  *o = accumulate ? *o + acc : acc;
}

// ---------------------------------------------------------------------------
// bp, exact adjoint of fp
// ---------------------------------------------------------------------------

// For pixel (i, j) and angle a, the fp samples (y, u) whose bilinear taps
// include (i, j) lie within a rotated 2 x 2 square around the preimage
// (u*, y*) of the pixel; every lattice point of its 4 x 4 bounding box is
// tested with the same float32 coordinates fp uses, and contributes
// wx * wy * line[u] where wx, wy are the tap's 1-D blend weights.
__global__ __launch_bounds__(256) void bp_adjoint_kernel(const float* __restrict__ sino, const float* __restrict__ maps,
                                                         int64_t maps_slice_stride, float* __restrict__ vol,
                                                         const float4* __restrict__ geom,
                                                         const float* __restrict__ w, int n, int n_angles,
                                                         int n_det, int angle0, int n_ang, int line_mode,
                                                         int accumulate) {
  const int j = blockIdx.x * kBpTile + (threadIdx.x & (kBpTile - 1));
  const int i = blockIdx.y * kBpTile + (threadIdx.x / kBpTile);
  const int b = blockIdx.z;
  if (i >= n || j >= n) return;
  const size_t nn = static_cast<size_t>(n) * n;
  const size_t pix = static_cast<size_t>(i) * n + j;
  float acc = 0.f;
  for (int aa = 0; aa < n_ang; ++aa) {
    const int a = angle0 + aa;
    const float4 g = geom[a];
    const float cs = g.x, sn = g.y, m02 = g.z, m12 = g.w;
    // preimage of (x' = j, y' = i): u = c (j - m02) - s (i - m12), y = s (j - m02) + c (i - m12)
    const float px = static_cast<float>(j) - m02;
    const float py = static_cast<float>(i) - m12;
    const float us = cs * px - sn * py;
    const float ysm = sn * px + cs * py;
    const int u_lo = max(static_cast<int>(floorf(us - 1.5f)), 0);
    const int u_hi = min(static_cast<int>(floorf(us + 1.5f)) + 1, n - 1);
    const int y_lo = max(static_cast<int>(floorf(ysm - 1.5f)), 0);
    const int y_hi = min(static_cast<int>(floorf(ysm + 1.5f)) + 1, n - 1);
    for (int u = u_lo; u <= u_hi; ++u) {
      const float uf = static_cast<float>(u);
      const float xu = __fmul_rn(cs, uf);
      const float yu = __fmul_rn(-sn, uf);
      float kern = 0.f;
      for (int y = y_lo; y <= y_hi; ++y) {
        const float yf = static_cast<float>(y);
        const float xs = __fadd_rn(__fadd_rn(xu, __fmul_rn(sn, yf)), m02);
        const float ys = __fadd_rn(__fadd_rn(yu, __fmul_rn(cs, yf)), m12);
        const float c0f = floorf(xs);
        const float r0f = floorf(ys);
        const int c0 = static_cast<int>(c0f);
        const int r0 = static_cast<int>(r0f);
        const float dc = xs - c0f;
        const float dr = ys - r0f;
        const float wx = (j == c0) ? 1.f - dc : ((j == c0 + 1) ? dc : 0.f);
        const float wy = (i == r0) ? 1.f - dr : ((i == r0 + 1) ? dr : 0.f);
        kern = fmaf(wx, wy, kern);
      }
      if (kern != 0.f) {
        for (int d = 0; d < n_det; ++d) {
          const float* line = line_mode ? sino + static_cast<size_t>(d) * n
                                         : sino + ((static_cast<size_t>(b) * n_det + d) * n_angles + a) * n;
          float f = w[static_cast<size_t>(a) * n_det + d];
          if (maps) f *= maps[static_cast<size_t>(b) * maps_slice_stride + (static_cast<size_t>(a) * n_det + d) * nn + pix];
          acc = fmaf(kern * f, line[u], acc);
        }
      }
    }
  }
  float* o = vol + static_cast<size_t>(b) * nn + pix;
  *o = accumulate ? *o + acc : acc;
}

static void launch_bp(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol, int batch,
                      int mode, int angle0, int n_ang, int line_mode, int accumulate, hipStream_t st) {
  dim3 grid((p.n + kBpTile - 1) / kBpTile, (p.n + kBpTile - 1) / kBpTile, batch);
  if (mode == CHIP_BP_SKIMAGE) {
    hipLaunchKernelGGL(bp_skimage_kernel<0>, grid, dim3(256), 0, st, sino, maps, stride, vol, p.d_fp_geom, p.d_w,
                       p.n, p.n_angles, p.n_det, angle0, n_ang, line_mode, accumulate);
  } else if (mode == CHIP_BP_SKIMAGE_SYM) {
    hipLaunchKernelGGL(bp_skimage_kernel<1>, grid, dim3(256), 0, st, sino, maps, stride, vol, p.d_fp_geom, p.d_w,
                       p.n, p.n_angles, p.n_det, angle0, n_ang, line_mode, accumulate);
  } else {
    hipLaunchKernelGGL(bp_adjoint_kernel, grid, dim3(256), 0, st, sino, maps, stride, vol, p.d_fp_geom, p.d_w, p.n,
                       p.n_angles, p.n_det, angle0, n_ang, line_mode, accumulate);
  }
}

}  // namespace chip

using namespace chip;

extern "C" {

int chip_fp(const chip_plan* plan, const float* vol, const float* maps, int64_t maps_slice_stride, float* sino,
            int batch, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(vol != nullptr && sino != nullptr, "vol/sino is NULL");
  CHIP_CHECK_ARG(batch >= 1 && batch <= 65535, "batch must be in [1, 65535]");
  CHIP_CHECK_ARG(maps_slice_stride >= 0, "maps_slice_stride must be >= 0");
  const Plan& p = plan->p;
  CHIP_CHECK_ARG(p.n_angles <= 65535, "too many angles for one launch");
  launch_fp(p, vol, maps, maps_slice_stride, sino, batch, 0, p.n_angles, 0, as_stream(stream));
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_fp_angle(const chip_plan* plan, const float* vol, const float* maps, float* line, int angle,
                  void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(vol != nullptr && line != nullptr, "vol/line is NULL");
  const Plan& p = plan->p;
  CHIP_CHECK_ARG(angle >= 0 && angle < p.n_angles, "angle index out of range");
  launch_fp(p, vol, maps, 0, line, 1, angle, 1, 1, as_stream(stream));
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_bp(const chip_plan* plan, const float* sino, const float* maps, int64_t maps_slice_stride, float* vol,
            int batch, int mode, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(vol != nullptr && sino != nullptr, "vol/sino is NULL");
  CHIP_CHECK_ARG(batch >= 1 && batch <= 65535, "batch must be in [1, 65535]");
  CHIP_CHECK_ARG(mode >= CHIP_BP_SKIMAGE && mode <= CHIP_BP_ADJOINT, "unknown bp mode");
  CHIP_CHECK_ARG(maps_slice_stride >= 0, "maps_slice_stride must be >= 0");
  const Plan& p = plan->p;
  launch_bp(p, sino, maps, maps_slice_stride, vol, batch, mode, 0, p.n_angles, 0, 0, as_stream(stream));
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_bp_angle(const chip_plan* plan, const float* line, const float* maps, float* vol, int angle, int mode,
                  void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(vol != nullptr && line != nullptr, "vol/line is NULL");
  CHIP_CHECK_ARG(mode >= CHIP_BP_SKIMAGE && mode <= CHIP_BP_ADJOINT, "unknown bp mode");
  const Plan& p = plan->p;
  CHIP_CHECK_ARG(angle >= 0 && angle < p.n_angles, "angle index out of range");
  launch_bp(p, line, maps, 0, vol, 1, mode, angle, 1, 1, 0, as_stream(stream));
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

}  // extern "C"
