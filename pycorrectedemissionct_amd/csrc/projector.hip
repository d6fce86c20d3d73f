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

// skimage's separate float32 / float64 products and sums: no contraction of
// a * b + c into one rounding in this file (explicit fmaf / fma stay fused)
#pragma clang fp contract(off)

namespace chip {
namespace {
__device__ __forceinline__ float mul_rn(float a, float b) { return a * b; }
__device__ __forceinline__ float add_rn(float a, float b) { return a + b; }
}  // namespace
}  // namespace chip

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
                                                 int line_mode, int maps_a0) {
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
                       (static_cast<size_t>(a - maps_a0) * n_det_total + det0 + d) * nn
                 : nullptr;
  }

  float acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0.f;

  if (u < n) {
    const float uf = static_cast<float>(u);
    const float xu = mul_rn(cs, uf);
    const float yu = mul_rn(-sn, uf);
    for (int y = dy; y < n; y += kFpLanesY) {
      const float yf = static_cast<float>(y);
      const float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
      const float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
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
                            int batch, int det0, int angle0, int n_ang, int line_mode, hipStream_t st,
                            int maps_a0 = 0) {
  dim3 grid((p.n + kFpBinsPerBlock - 1) / kFpBinsPerBlock, n_ang, batch);
  hipLaunchKernelGGL(fp_kernel<D>, grid, dim3(256), 0, st, vol, maps, stride, sino, p.d_fp_geom, p.d_w, p.n,
                     p.n_angles, p.n_det, det0, angle0, line_mode, maps_a0);
}

// ---------------------------------------------------------------------------
// fp, LDS-staged (whole-sinogram path)
// ---------------------------------------------------------------------------
//
// One workgroup = one angle x 64 consecutive detector bins, walking the full
// ray length in chunks of 32 sample rows. Per chunk the axis-aligned
// bounding box of the chunk's bilinear footprint is loaded once, with
// coalesced row segments, as T_d = vol * M[a, d] into LDS (zero outside the
// slice), and the 64 x 32 samples gather their four taps from LDS. Thread
// (bin = tid % 64, group = tid / 64) accumulates 8 consecutive rows of each
// chunk in order; the four groups are summed in a fixed order at the end, so
// results are bitwise reproducible. The sample map is monotone in (u, y) in
// float32, so the chunk's footprint lies inside the box spanned by its four
// corner samples (+1 for the right/bottom taps): every tap index stays in the
// tile without per-tap bounds checks.
//
// Block order is XCD-aware: workgroup p runs on XCD p % 8, so logical block
// ids are assigned such that the 8 bin-blocks of one angle land on the same
// XCD one after another and share that XCD's L2 copy of the angle's maps.

constexpr int kLdsBins = 64;
constexpr int kLdsRows = 32;        // sample rows per chunk
constexpr int kLdsRowsPerThread = 8;
constexpr int kLdsTile = 76;        // >= 63|c| + 31|s| + 3 over all angles
constexpr int kLdsPitch = kLdsTile + 1;

__device__ __forceinline__ int xcd_logical_block(int p, int g) {
  const int q = g >> 3, r = g & 7;
  const int x = p & 7, k = p >> 3;
  return x * q + min(x, r) + k;
}

template <int D>
__global__ __launch_bounds__(256) void fp_lds_kernel(const float* __restrict__ vol, const float* __restrict__ maps,
                                                     int64_t maps_slice_stride, float* __restrict__ sino,
                                                     const float4* __restrict__ geom, const float* __restrict__ w,
                                                     int n, int n_angles, int n_det_total, int det0, int n_ublk) {
  extern __shared__ float lds[];  // [D][kLdsTile * kLdsPitch]
  const int g = gridDim.x;
  const int lb = xcd_logical_block(blockIdx.x, g);
  const int a = lb / n_ublk;
  const int u0 = (lb - a * n_ublk) * kLdsBins;
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int grp = tid >> 6;
  const int u = u0 + lane;

  const float4 gm = geom[a];
  const float cs = gm.x, sn = gm.y, m02 = gm.z, m12 = gm.w;
  const size_t nn = static_cast<size_t>(n) * n;
  const float* v = vol + static_cast<size_t>(b) * nn;
  const float* mbase = maps ? maps + static_cast<size_t>(b) * maps_slice_stride +
                                  (static_cast<size_t>(a) * n_det_total + det0) * nn
                            : nullptr;
  const float fn = static_cast<float>(n);

  // per-thread constant parts of the sample map
  const float uf = static_cast<float>(u);
  const float xu = mul_rn(cs, uf);
  const float yu = mul_rn(-sn, uf);
  // corner columns of the block (clamped to real bins)
  const int u_last = min(u0 + kLdsBins, n) - 1;
  const float xu_a = mul_rn(cs, static_cast<float>(u0)), yu_a = mul_rn(-sn, static_cast<float>(u0));
  const float xu_b = mul_rn(cs, static_cast<float>(u_last)), yu_b = mul_rn(-sn, static_cast<float>(u_last));

  float acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0.f;

  for (int yc = 0; yc < n; yc += kLdsRows) {
    const int y_last = min(yc + kLdsRows, n) - 1;
    // footprint box from the four corner samples
    float xmin, xmax, ymin, ymax;
    {
      const float ya = static_cast<float>(yc), yb = static_cast<float>(y_last);
      const float sya = mul_rn(sn, ya), syb = mul_rn(sn, yb);
      const float cya = mul_rn(cs, ya), cyb = mul_rn(cs, yb);
      const float x0 = add_rn(add_rn(xu_a, sya), m02), x1 = add_rn(add_rn(xu_b, sya), m02);
      const float x2 = add_rn(add_rn(xu_a, syb), m02), x3 = add_rn(add_rn(xu_b, syb), m02);
      const float y0 = add_rn(add_rn(yu_a, cya), m12), y1 = add_rn(add_rn(yu_b, cya), m12);
      const float y2 = add_rn(add_rn(yu_a, cyb), m12), y3 = add_rn(add_rn(yu_b, cyb), m12);
      xmin = fminf(fminf(x0, x1), fminf(x2, x3));
      xmax = fmaxf(fmaxf(x0, x1), fmaxf(x2, x3));
      ymin = fminf(fminf(y0, y1), fminf(y2, y3));
      ymax = fmaxf(fmaxf(y0, y1), fmaxf(y2, y3));
    }
    // taps span [floor(min), floor(max) + 1]; only [-1, n] can be non-trivial
    const int clo = max(static_cast<int>(floorf(fmaxf(xmin, -2.f))), -1);
    const int chi = min(static_cast<int>(floorf(fminf(xmax, fn + 1.f))) + 1, n);
    const int rlo = max(static_cast<int>(floorf(fmaxf(ymin, -2.f))), -1);
    const int rhi = min(static_cast<int>(floorf(fminf(ymax, fn + 1.f))) + 1, n);
    if (chi < 0 || clo > n - 1 || rhi < 0 || rlo > n - 1) continue;  // chunk misses the slice (uniform)
    const int tw = chi - clo + 1;
    const int th = rhi - rlo + 1;

    __syncthreads();  // previous chunk's gathers are done with the tile
    // stage T_d = vol * M_d over the box; rows spread over the 4 waves, 64 lanes per row segment
    for (int rr = grp; rr < th; rr += 4) {
      const int r = rlo + rr;
      const bool rin = r >= 0 && r < n;
      for (int cc = lane; cc < tw; cc += 64) {
        const int c = clo + cc;
        const bool in = rin && c >= 0 && c < n;
        const size_t off = static_cast<size_t>(r) * n + c;
        const float fv = in ? v[off] : 0.f;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const float mv = (in && mbase) ? mbase[static_cast<size_t>(d) * nn + off] : 1.f;
          lds[d * (kLdsTile * kLdsPitch) + rr * kLdsPitch + cc] = fv * mv;
        }
      }
    }
    __syncthreads();

    if (u < n) {
      const int yb0 = yc + grp * kLdsRowsPerThread;
#pragma unroll
      for (int k = 0; k < kLdsRowsPerThread; ++k) {
        const int y = yb0 + k;
        if (y >= n) break;
        const float yf = static_cast<float>(y);
        const float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
        const float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
        if (!(xs > -1.f && xs < fn && ys > -1.f && ys < fn)) continue;
        const float c0f = floorf(xs);
        const float r0f = floorf(ys);
        const float dc = xs - c0f;
        const float dr = ys - r0f;
        const int li = (static_cast<int>(r0f) - rlo) * kLdsPitch + (static_cast<int>(c0f) - clo);
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const float* t = lds + d * (kLdsTile * kLdsPitch) + li;
          const float t00 = t[0], t01 = t[1], t10 = t[kLdsPitch], t11 = t[kLdsPitch + 1];
          const float top = fmaf(dc, t01 - t00, t00);
          const float bot = fmaf(dc, t11 - t10, t10);
          acc[d] += fmaf(dr, bot - top, top);
        }
      }
    }
  }

  // fixed-order reduction of the 4 row groups
  __syncthreads();
#pragma unroll
  for (int d = 0; d < D; ++d) lds[(d * 4 + grp) * 64 + lane] = acc[d];
  __syncthreads();
  if (grp == 0 && u < n) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float s = ((lds[(d * 4 + 0) * 64 + lane] + lds[(d * 4 + 1) * 64 + lane]) + lds[(d * 4 + 2) * 64 + lane]) +
                      lds[(d * 4 + 3) * 64 + lane];
      const int dd = det0 + d;
      sino[((static_cast<size_t>(b) * n_det_total + dd) * n_angles + a) * n + u] =
          s * w[static_cast<size_t>(a) * n_det_total + dd];
    }
  }
}

// ---------------------------------------------------------------------------
// fp, LDS-staged with register prefetch (N % 4 == 0)
// ---------------------------------------------------------------------------
//
// Same decomposition as fp_lds_kernel, with the staging software-pipelined:
// the next chunk's footprint box is fetched into registers as float4 (its
// columns start at a multiple of 4, so with N % 4 == 0 every float4 is
// either wholly inside the slice or wholly outside) while the current chunk
// is gathered from LDS. Chunks whose box misses the slice are skipped; in
// chunks whose samples all fall inside (-1, N) the per-sample test is
// skipped too.

constexpr int kP2Rows = 76;    // box rows  <= 63|s| + 31|c| + 3
constexpr int kP2Slots = 20;   // float4 slots per box row (box cols <= 76 + 3 alignment)
constexpr int kP2Pitch = 81;   // odd pitch: row-strided gathers hit distinct banks
constexpr int kP2PerThread = (kP2Rows * kP2Slots + 255) / 256;  // 6

struct FpBox {
  int clo, chi, rlo, rhi;  // box (clo is 4-aligned), inclusive
  int valid_all;           // every sample of the chunk lies in (-1, N)^2
  int empty;
};

__device__ __forceinline__ FpBox fp_chunk_box(int yc, int n, float fn, float xu_a, float yu_a, float xu_b, float yu_b,
                                              float cs, float sn, float m02, float m12) {
  const int y_last = min(yc + kLdsRows, n) - 1;
  const float ya = static_cast<float>(yc), yb = static_cast<float>(y_last);
  const float sya = mul_rn(sn, ya), syb = mul_rn(sn, yb);
  const float cya = mul_rn(cs, ya), cyb = mul_rn(cs, yb);
  const float x0 = add_rn(add_rn(xu_a, sya), m02), x1 = add_rn(add_rn(xu_b, sya), m02);
  const float x2 = add_rn(add_rn(xu_a, syb), m02), x3 = add_rn(add_rn(xu_b, syb), m02);
  const float y0 = add_rn(add_rn(yu_a, cya), m12), y1 = add_rn(add_rn(yu_b, cya), m12);
  const float y2 = add_rn(add_rn(yu_a, cyb), m12), y3 = add_rn(add_rn(yu_b, cyb), m12);
  const float xmin = fminf(fminf(x0, x1), fminf(x2, x3));
  const float xmax = fmaxf(fmaxf(x0, x1), fmaxf(x2, x3));
  const float ymin = fminf(fminf(y0, y1), fminf(y2, y3));
  const float ymax = fmaxf(fmaxf(y0, y1), fmaxf(y2, y3));
  FpBox bx;
  bx.valid_all = (xmin > -1.f && xmax < fn && ymin > -1.f && ymax < fn) ? 1 : 0;
  const int clo = max(static_cast<int>(floorf(fmaxf(xmin, -2.f))), -1);
  bx.chi = min(static_cast<int>(floorf(fminf(xmax, fn + 1.f))) + 1, n);
  bx.rlo = max(static_cast<int>(floorf(fmaxf(ymin, -2.f))), -1);
  bx.rhi = min(static_cast<int>(floorf(fminf(ymax, fn + 1.f))) + 1, n);
  bx.clo = clo >= 0 ? (clo & ~3) : -4;
  bx.empty = (bx.chi < 0 || clo > n - 1 || bx.rhi < 0 || bx.rlo > n - 1) ? 1 : 0;
  return bx;
}

template <int D>
__global__ __launch_bounds__(256) void fp_lds2_kernel(const float* __restrict__ vol, const float* __restrict__ maps,
                                                      int64_t maps_slice_stride, float* __restrict__ sino,
                                                      const float4* __restrict__ geom, const float* __restrict__ w,
                                                      int n, int n_angles, int n_det_total, int det0, int n_ublk) {
  extern __shared__ float lds[];  // [D][kP2Rows * kP2Pitch]
  constexpr int kMapStride = kP2Rows * kP2Pitch;
  const int g = gridDim.x;
  const int lb = xcd_logical_block(blockIdx.x, g);
  const int a = lb / n_ublk;
  const int u0 = (lb - a * n_ublk) * kLdsBins;
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int grp = tid >> 6;
  const int u = u0 + lane;

  const float4 gm = geom[a];
  const float cs = gm.x, sn = gm.y, m02 = gm.z, m12 = gm.w;
  const size_t nn = static_cast<size_t>(n) * n;
  const float* v = vol + static_cast<size_t>(b) * nn;
  const float* mbase = maps ? maps + static_cast<size_t>(b) * maps_slice_stride +
                                  (static_cast<size_t>(a) * n_det_total + det0) * nn
                            : nullptr;
  const float fn = static_cast<float>(n);
  const float uf = static_cast<float>(u);
  const float xu = mul_rn(cs, uf);
  const float yu = mul_rn(-sn, uf);
  const int u_last = min(u0 + kLdsBins, n) - 1;
  const float xu_a = mul_rn(cs, static_cast<float>(u0)), yu_a = mul_rn(-sn, static_cast<float>(u0));
  const float xu_b = mul_rn(cs, static_cast<float>(u_last)), yu_b = mul_rn(-sn, static_cast<float>(u_last));

  float4 fr[kP2PerThread];
  float4 mr[D][kP2PerThread];

  auto issue = [&](const FpBox& bx) {
    const int th = bx.rhi - bx.rlo + 1;
    const int cw = bx.chi - bx.clo + 1;
#pragma unroll
    for (int i = 0; i < kP2PerThread; ++i) {
      const int s = tid + 256 * i;
      const int row = s / kP2Slots;
      const int c4 = s - row * kP2Slots;
      const int r = bx.rlo + row;
      const int c = bx.clo + 4 * c4;
      const bool in = row < th && 4 * c4 < cw && r >= 0 && r < n && c >= 0 && c < n;
      const size_t off = in ? static_cast<size_t>(r) * n + c : 0;
      fr[i] = in ? *reinterpret_cast<const float4*>(v + off) : make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
      for (int d = 0; d < D; ++d) {
        mr[d][i] = (in && mbase) ? *reinterpret_cast<const float4*>(mbase + static_cast<size_t>(d) * nn + off)
                                 : make_float4(1.f, 1.f, 1.f, 1.f);
      }
    }
  };
  auto stage = [&](const FpBox& bx) {
    const int th = bx.rhi - bx.rlo + 1;
    const int cw = bx.chi - bx.clo + 1;
#pragma unroll
    for (int i = 0; i < kP2PerThread; ++i) {
      const int s = tid + 256 * i;
      const int row = s / kP2Slots;
      const int c4 = s - row * kP2Slots;
      if (row < th && 4 * c4 < cw) {
        const int o = row * kP2Pitch + 4 * c4;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          float* t = lds + d * kMapStride + o;
          t[0] = fr[i].x * mr[d][i].x;
          t[1] = fr[i].y * mr[d][i].y;
          t[2] = fr[i].z * mr[d][i].z;
          t[3] = fr[i].w * mr[d][i].w;
        }
      }
    }
  };

  float acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0.f;

  int yc = 0;
  FpBox cur;
  for (;; yc += kLdsRows) {  // first chunk that meets the slice
    if (yc >= n) break;
    cur = fp_chunk_box(yc, n, fn, xu_a, yu_a, xu_b, yu_b, cs, sn, m02, m12);
    if (!cur.empty) break;
  }
  if (yc < n) issue(cur);
  while (yc < n) {
    __syncthreads();  // the previous chunk's gathers are done with the tile
    stage(cur);
    __syncthreads();
    // look ahead: next chunk that meets the slice; its loads overlap this chunk's gathers
    int yn = yc + kLdsRows;
    FpBox nxt;
    for (; yn < n; yn += kLdsRows) {
      nxt = fp_chunk_box(yn, n, fn, xu_a, yu_a, xu_b, yu_b, cs, sn, m02, m12);
      if (!nxt.empty) break;
    }
    if (yn < n) issue(nxt);

    if (u < n) {
      const int yb0 = yc + grp * kLdsRowsPerThread;
      const int base = cur.rlo * kP2Pitch + cur.clo;
      if (cur.valid_all && yb0 + kLdsRowsPerThread <= n) {
#pragma unroll
        for (int k = 0; k < kLdsRowsPerThread; ++k) {
          const float yf = static_cast<float>(yb0 + k);
          const float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
          const float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
          const float c0f = floorf(xs);
          const float r0f = floorf(ys);
          const float dc = xs - c0f;
          const float dr = ys - r0f;
          const int li = static_cast<int>(r0f) * kP2Pitch + static_cast<int>(c0f) - base;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const float* t = lds + d * kMapStride + li;
            const float t00 = t[0], t01 = t[1], t10 = t[kP2Pitch], t11 = t[kP2Pitch + 1];
            const float top = fmaf(dc, t01 - t00, t00);
            const float bot = fmaf(dc, t11 - t10, t10);
            acc[d] += fmaf(dr, bot - top, top);
          }
        }
      } else {
        for (int k = 0; k < kLdsRowsPerThread; ++k) {
          const int y = yb0 + k;
          if (y >= n) break;
          const float yf = static_cast<float>(y);
          const float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
          const float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
          if (!(xs > -1.f && xs < fn && ys > -1.f && ys < fn)) continue;
          const float c0f = floorf(xs);
          const float r0f = floorf(ys);
          const float dc = xs - c0f;
          const float dr = ys - r0f;
          const int li = static_cast<int>(r0f) * kP2Pitch + static_cast<int>(c0f) - base;
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const float* t = lds + d * kMapStride + li;
            const float t00 = t[0], t01 = t[1], t10 = t[kP2Pitch], t11 = t[kP2Pitch + 1];
            const float top = fmaf(dc, t01 - t00, t00);
            const float bot = fmaf(dc, t11 - t10, t10);
            acc[d] += fmaf(dr, bot - top, top);
          }
        }
      }
    }
    yc = yn;
    cur = nxt;
  }

  __syncthreads();
#pragma unroll
  for (int d = 0; d < D; ++d) lds[(d * 4 + grp) * 64 + lane] = acc[d];
  __syncthreads();
  if (grp == 0 && u < n) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float s = ((lds[(d * 4 + 0) * 64 + lane] + lds[(d * 4 + 1) * 64 + lane]) + lds[(d * 4 + 2) * 64 + lane]) +
                      lds[(d * 4 + 3) * 64 + lane];
      const int dd = det0 + d;
      sino[((static_cast<size_t>(b) * n_det_total + dd) * n_angles + a) * n + u] =
          s * w[static_cast<size_t>(a) * n_det_total + dd];
    }
  }
}

template <int D>
static void launch_fp_lds2_group(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino,
                                 int batch, int det0, hipStream_t st) {
  const int n_ublk = (p.n + kLdsBins - 1) / kLdsBins;
  dim3 grid(n_ublk * p.n_angles, batch);
  const size_t shmem = sizeof(float) * D * kP2Rows * kP2Pitch;
  hipLaunchKernelGGL(fp_lds2_kernel<D>, grid, dim3(256), shmem, st, vol, maps, stride, sino, p.d_fp_geom, p.d_w, p.n,
                     p.n_angles, p.n_det, det0, n_ublk);
}

template <int D>
static void launch_fp_lds_group(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino,
                                int batch, int det0, hipStream_t st) {
  const int n_ublk = (p.n + kLdsBins - 1) / kLdsBins;
  dim3 grid(n_ublk * p.n_angles, batch);
  const size_t shmem = sizeof(float) * D * kLdsTile * kLdsPitch;
  hipLaunchKernelGGL(fp_lds_kernel<D>, grid, dim3(256), shmem, st, vol, maps, stride, sino, p.d_fp_geom, p.d_w, p.n,
                     p.n_angles, p.n_det, det0, n_ublk);
}

static void launch_fp(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch,
                      int angle0, int n_ang, int line_mode, hipStream_t st) {
  const bool whole = !line_mode && angle0 == 0 && n_ang == p.n_angles;
  // the prefetching kernel reads float4: needs N % 4 == 0 and 16-byte aligned slices / maps
  const bool vec_ok = (p.n % 4 == 0) && (stride % 4 == 0) && (reinterpret_cast<uintptr_t>(vol) % 16 == 0) &&
                      (reinterpret_cast<uintptr_t>(maps) % 16 == 0);
  const int variant = (p.fp_variant >= 2 && !vec_ok) ? 1 : p.fp_variant;
  if (whole && variant >= 1) {
    // two detectors per pass keep the staged tiles within 64 KiB of LDS
    for (int det0 = 0; det0 < p.n_det; det0 += 2) {
      const bool two = p.n_det - det0 >= 2;
      if (variant == 6) {
        launch_fp_band(p, vol, maps, stride, sino, batch, det0, two ? 2 : 1, 0, p.n_angles, 0, st);
      } else if (variant >= 3) {
        launch_fp_stream(p, vol, maps, stride, sino, batch, det0, two ? 2 : 1, 0, p.n_angles, 0, st);
      } else if (variant == 2) {
        if (two) launch_fp_lds2_group<2>(p, vol, maps, stride, sino, batch, det0, st);
        else launch_fp_lds2_group<1>(p, vol, maps, stride, sino, batch, det0, st);
      } else {
        if (two) launch_fp_lds_group<2>(p, vol, maps, stride, sino, batch, det0, st);
        else launch_fp_lds_group<1>(p, vol, maps, stride, sino, batch, det0, st);
      }
    }
    return;
  }
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
// Sum over angles [angle0, angle0 + n_ang) and detectors of the line
// interpolated at pixel (i, j) of one slice: sino is the slice's [D][A][n]
// sinogram (line_mode: one [D][n] line set), maps its [A - maps_a0][D][n][n]
// block (SYM only, may be NULL).
template <int SYM>
__device__ __forceinline__ float bp_skimage_pixel(const float* __restrict__ sino, const float* __restrict__ maps,
                                                  const float4* __restrict__ geom, const float* __restrict__ w, int n,
                                                  int n_angles, int n_det, int angle0, int n_ang, int line_mode,
                                                  int maps_a0, int i, int j) {
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
    float t = fmaf(jc, g.x, -(ic * g.y));
    // np.interp is discontinuous at the ends of the line (0 outside): decide
    // membership with the reference's double-precision t (exact products of
    // the integer offsets with the float32 cos/sin, one rounding) whenever
    // the float32 t is within its error bound of an end
    if (fabsf(t - tlo) < 1e-3f || fabsf(t - thi) < 1e-3f) {
      const double td = fma(static_cast<double>(jc), static_cast<double>(g.x),
                            -(static_cast<double>(ic) * static_cast<double>(g.y)));
      if (!(td >= static_cast<double>(tlo) && td <= static_cast<double>(thi))) continue;
      t = fminf(fmaxf(t, tlo), thi);
    } else if (!(t >= tlo && t <= thi)) {
      continue;
    }
    const float tf = floorf(t);
    const float fr = t - tf;
    const int kk = static_cast<int>(tf) + k;  // in [0, n-1]
    const bool ok1 = kk + 1 < n;
    for (int d = 0; d < n_det; ++d) {
      const float* line = line_mode ? sino + static_cast<size_t>(d) * n
                                     : sino + (static_cast<size_t>(d) * n_angles + a) * n;
      const float s0 = line[kk];
      const float s1 = ok1 ? line[kk + 1] : 0.f;
      float val = fmaf(fr, s1 - s0, s0);
      if (SYM) {
        float f = w[static_cast<size_t>(a) * n_det + d];
        if (maps) f *= maps[(static_cast<size_t>(a - maps_a0) * n_det + d) * nn + pix];
        val *= f;
      }
      acc += val;
    }
  }
  return acc;
}

template <int SYM>
__global__ __launch_bounds__(256) void bp_skimage_kernel(const float* __restrict__ sino, const float* __restrict__ maps,
                                                         int64_t maps_slice_stride, float* __restrict__ vol,
                                                         const float4* __restrict__ geom,
                                                         const float* __restrict__ w, int n, int n_angles,
                                                         int n_det, int angle0, int n_ang, int line_mode,
                                                         int accumulate, int maps_a0) {
  const int j = blockIdx.x * kBpTile + (threadIdx.x & (kBpTile - 1));
  const int i = blockIdx.y * kBpTile + (threadIdx.x / kBpTile);
  const int b = blockIdx.z;
  if (i >= n || j >= n) return;
  const size_t nn = static_cast<size_t>(n) * n;
  const float* sb = line_mode ? sino : sino + static_cast<size_t>(b) * n_det * n_angles * n;
  const float* mb = maps ? maps + static_cast<size_t>(b) * maps_slice_stride : nullptr;
  const float acc = bp_skimage_pixel<SYM>(sb, mb, geom, w, n, n_angles, n_det, angle0, n_ang, line_mode, maps_a0, i, j);
  float* o = vol + static_cast<size_t>(b) * nn + static_cast<size_t>(i) * n + j;
  *o = accumulate ? *o + acc : acc;
}

void launch_bp_skimage_batch(const BatchPlan& bp, const float* sino, const float* maps, float* vol, int sym,
                             hipStream_t st) {
  // line windows staged per tile, as for single slices (bp_skimage_tile.hip)
  launch_bp_skimage_tiles_batch(bp, sino, maps, vol, sym, st);
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
      const float xu = mul_rn(cs, uf);
      const float yu = mul_rn(-sn, uf);
      float kern = 0.f;
      for (int y = y_lo; y <= y_hi; ++y) {
        const float yf = static_cast<float>(y);
        const float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
        const float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
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
                      int mode, int angle0, int n_ang, int line_mode, int accumulate, hipStream_t st,
                      const BpScratch& scr = BpScratch{}) {
  dim3 grid((p.n + kBpTile - 1) / kBpTile, (p.n + kBpTile - 1) / kBpTile, batch);
  if (!line_mode && (mode == CHIP_BP_SKIMAGE || mode == CHIP_BP_SKIMAGE_SYM)) {
    // whole sinograms and angle ranges: line windows staged in LDS per tile
    launch_bp_skimage_tiles(p, sino, maps, stride, vol, batch, mode == CHIP_BP_SKIMAGE_SYM ? 1 : 0, angle0, n_ang,
                            0, accumulate, st);
  } else if (mode == CHIP_BP_SKIMAGE) {
    hipLaunchKernelGGL(bp_skimage_kernel<0>, grid, dim3(256), 0, st, sino, maps, stride, vol, p.d_fp_geom, p.d_w,
                       p.n, p.n_angles, p.n_det, angle0, n_ang, line_mode, accumulate, 0);
  } else if (mode == CHIP_BP_SKIMAGE_SYM) {
    hipLaunchKernelGGL(bp_skimage_kernel<1>, grid, dim3(256), 0, st, sino, maps, stride, vol, p.d_fp_geom, p.d_w,
                       p.n, p.n_angles, p.n_det, angle0, n_ang, line_mode, accumulate, 0);
  } else if (!line_mode && angle0 == 0 && n_ang == p.n_angles && !accumulate) {
    launch_bp_adjoint_tiles(p, sino, maps, stride, vol, batch, 0, p.n_angles, 0, 0, st, scr);
  } else {
    hipLaunchKernelGGL(bp_adjoint_kernel, grid, dim3(256), 0, st, sino, maps, stride, vol, p.d_fp_geom, p.d_w, p.n,
                       p.n_angles, p.n_det, angle0, n_ang, line_mode, accumulate);
  }
}

void launch_bp_sinogram(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol, int batch,
                        int mode, hipStream_t st, const BpScratch& scr) {
  launch_bp(p, sino, maps, stride, vol, batch, mode, 0, p.n_angles, 0, 0, st, scr);
}

bool fp_range_tile(const Plan& p) {
  return p.fp_variant == 10 && p.n % 4 == 0 && fp_tile_count(p.n) <= kFpTileMaxTiles;
}

int fp_range_ready(const Plan& p, int chunk, hipStream_t st) {
  if (!fp_range_tile(p)) return CHIP_OK;
  const int rc = fp_tile_table(p, st);
  if (rc != CHIP_OK) return rc;
  const size_t need = fp_tile_part_bytes(p, 1, chunk);
  if (p.fp_tile_bytes >= need) return CHIP_OK;
  CHIP_CHECK_ARG(!stream_capturing(st), "the image-tile fp needs " + std::to_string(need) +
                                            " bytes of partial sums: run it once outside stream capture first");
  return ensure_buffer(p.d_fp_tile, p.fp_tile_bytes, need, st, "image-tile fp partial sums");
}

void launch_fp_range(const Plan& p, const float* vol, const float* maps, float* sino, int angle0, int n_ang,
                     int maps_a0, hipStream_t st) {
  const bool vec_ok = (p.n % 4 == 0) && (reinterpret_cast<uintptr_t>(vol) % 16 == 0) &&
                      (reinterpret_cast<uintptr_t>(maps) % 16 == 0);
  if (vec_ok && fp_range_tile(p) && p.fp_tab_ok && p.fp_tile_bytes >= fp_tile_part_bytes(p, 1, n_ang)) {
    launch_fp_tile(p, vol, maps, TileSlots{}, sino, 1, angle0, n_ang, maps_a0, p.d_fp_tile, p.fp_tile, st);
    return;
  }
  if (vec_ok && p.fp_variant >= 3) {
    for (int det0 = 0; det0 < p.n_det; det0 += 2) {
      const int nd = p.n_det - det0 >= 2 ? 2 : 1;
      if (p.fp_variant == 6) launch_fp_band(p, vol, maps, 0, sino, 1, det0, nd, angle0, n_ang, maps_a0, st);
      else launch_fp_stream(p, vol, maps, 0, sino, 1, det0, nd, angle0, n_ang, maps_a0, st);
    }
    return;
  }
  for (int det0 = 0; det0 < p.n_det; det0 += kMaxDetPerPass) {
    const int nd = std::min(kMaxDetPerPass, p.n_det - det0);
    switch (nd) {
      case 1: launch_fp_group<1>(p, vol, maps, 0, sino, 1, det0, angle0, n_ang, 0, st, maps_a0); break;
      case 2: launch_fp_group<2>(p, vol, maps, 0, sino, 1, det0, angle0, n_ang, 0, st, maps_a0); break;
      case 3: launch_fp_group<3>(p, vol, maps, 0, sino, 1, det0, angle0, n_ang, 0, st, maps_a0); break;
      default: launch_fp_group<4>(p, vol, maps, 0, sino, 1, det0, angle0, n_ang, 0, st, maps_a0); break;
    }
  }
}

void launch_bp_range(const Plan& p, const float* sino, const float* maps, float* vol, int mode, int angle0,
                     int n_ang, int maps_a0, int accumulate, hipStream_t st, const BpScratch& scr) {
  if (mode == CHIP_BP_ADJOINT) {
    launch_bp_adjoint_tiles(p, sino, maps, 0, vol, 1, angle0, n_ang, maps_a0, accumulate, st, scr);
    return;
  }
  launch_bp_skimage_tiles(p, sino, maps, 0, vol, 1, 1, angle0, n_ang, maps_a0, accumulate, st);
}

int tune_tile_shape(FpTileShape& sh, const std::string& k, int value) {
  if (k == "fp_tile_kind") {
    CHIP_CHECK_ARG(value >= -1 && value <= 3, "fp_tile_kind must be in -1..3 (-1: by detector count)");
    sh.kind = value;
  } else if (k == "fp_tile_group") {
    CHIP_CHECK_ARG(value >= 1 && value <= 4096, "fp_tile_group must be in 1..4096");
    sh.group = value;
  } else {
    CHIP_CHECK_ARG(value == 0 || value == 1, "fp_tile_pairs must be 0 or 1");
    sh.pairs = value;
  }
  return CHIP_OK;
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
  const hipStream_t st = as_stream(stream);
  const bool vec_ok = (p.n % 4 == 0) && (maps_slice_stride % 4 == 0) &&
                      (reinterpret_cast<uintptr_t>(vol) % 16 == 0) && (reinterpret_cast<uintptr_t>(maps) % 16 == 0);
  if (p.fp_variant == 10 && vec_ok && fp_tile_count(p.n) <= kFpTileMaxTiles) {
    ScratchUse use(*p.fp_state, st);
    int rc = fp_tile_table(p, st);
    if (rc != CHIP_OK) return rc;
    const size_t need = fp_tile_part_bytes(p, batch, p.n_angles);
    if (p.fp_tile_bytes < need) {
      CHIP_CHECK_ARG(!stream_capturing(st), "the image-tile fp needs " + std::to_string(need) +
                                                " bytes of partial sums: run it once outside stream capture first");
      rc = ensure_buffer(p.d_fp_tile, p.fp_tile_bytes, need, st, "image-tile fp partial sums");
      if (rc != CHIP_OK) return rc;
    }
    TileSlots sl;
    sl.vol_stride = static_cast<int64_t>(p.n) * p.n;
    sl.maps_stride = maps_slice_stride;
    sl.sino_stride = static_cast<int64_t>(p.n_det) * p.n_angles * p.n;
    launch_fp_tile(p, vol, maps, sl, sino, batch, 0, p.n_angles, 0, p.d_fp_tile, p.fp_tile, st);
    CHIP_LAUNCH_CHECK();
    return CHIP_OK;
  }
  launch_fp(p, vol, maps, maps_slice_stride, sino, batch, 0, p.n_angles, 0, st);
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
  const hipStream_t st = as_stream(stream);
  BpPartGuard guard(p, batch, mode, p.n_angles, st);
  launch_bp(p, sino, maps, maps_slice_stride, vol, batch, mode, 0, p.n_angles, 0, 0, st, guard.scratch());
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_plan_tune(chip_plan* plan, const char* key, int value) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && key != nullptr, "NULL argument");
  const std::string k(key);
  if (k == "fp_kernel") {
    CHIP_CHECK_ARG(value >= 0 && value <= 10, "fp_kernel must be in 0..10");
    plan->p.fp_variant = value;
    return CHIP_OK;
  }
  if (k == "fp_tile_kind" || k == "fp_tile_group" || k == "fp_tile_pairs")
    return tune_tile_shape(plan->p.fp_tile, k, value);
  if (k == "maps_exact") {
    CHIP_CHECK_ARG(value == 0 || value == 1, "maps_exact must be 0 or 1");
    plan->p.maps_exact = value;
    return CHIP_OK;
  }
  if (k == "bp_parts") {
    CHIP_CHECK_ARG(value == -1 || value == 0 || (value >= 2 && value <= kBpMaxParts),
                   "bp_parts must be -1 (from occupancy), 0 (never) or 2..8");
    plan->p.bp_parts_tune = value;
    return CHIP_OK;
  }
  if (k == "otf_budget_mib") {
    CHIP_CHECK_ARG(value >= 16, "otf_budget_mib must be >= 16");
    plan->p.otf_budget = static_cast<size_t>(value) << 20;
    return CHIP_OK;
  }
  set_error("chip_plan_tune: unknown key '" + k + "'");
  return CHIP_EINVAL;
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
