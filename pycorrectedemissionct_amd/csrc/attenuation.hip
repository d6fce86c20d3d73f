// Attenuation maps of libcorrct_hip (gfx950).
//
// Restates compute_attenuation (_projector_backends.py:43-99) for every beam
// direction of a plan and composes AttenuationVolume.compute_maps
// (physics/attenuation.py:122-180):
//   mu_p   = mu zero-padded by e on each side (side Np = N + 2e)
//   R      = rotate(mu_p, -deg)                 (order-1 warp, 0 outside)
//   C[r]   = sum_{k<=r} (R[k] + R[k-1 mod Np]) / 2   (float32, sequential in r)
//   E      = exp(-rotate(C, +deg))[e:-e, e:-e]
//   M[a,d] = E_in(theta_a) * E_out(theta_a + phi_d + pi)
//
// Kernel 1 (att_scan_kernel): per beam, 8-lane groups walk the columns of the
// rotated frame top to bottom in bands of 8 rows, sampling mu_p at the
// rotated positions and carrying the trapezoid prefix sum; the running sum is
// the same float32 sequence numpy's cumsum performs.
// Kernel 2 (att_compose_kernel): each thread takes two map pixels, samples the
// beams' prefix-sum images at the back-rotated positions, takes exp and
// multiplies in-beam by out-beam.
//
// The order-1 blend: by default in float32 (maps within ~1e-6 of scikit-image's,
// the bar is 1e-5); with chip_plan_tune "maps_exact" = 1 as scikit-image 0.18.3's
// compiled warp blends: the fractions r - floor(r) are float32, the products
// with the right-hand neighbours are float32, the rest of the blend is double,
// rounded to float32 -- so the maps match a float32 scikit-image evaluation to
// the last bit except where exp differs by an ulp.
#include <algorithm>

#include "common.hpp"
#include "solver_elem.hpp"

// skimage's separate float32 / float64 products and sums: no contraction of
// a * b + c into one rounding in this file (explicit fmaf / fma stay fused)
#pragma clang fp contract(off)

namespace chip {
namespace {
__device__ __forceinline__ float mul_rn(float a, float b) { return a * b; }
__device__ __forceinline__ float add_rn(float a, float b) { return a + b; }
__device__ __forceinline__ double dmul_rn(double a, double b) { return a * b; }
__device__ __forceinline__ double dadd_rn(double a, double b) { return a + b; }
}  // namespace
}  // namespace chip

namespace chip {

// EXACT = false: the same taps blended in float32 (t00 + dc (t01 - t00), then the
// rows): within a few float32 ulps of the exact blend, which keeps the maps within
// 1e-6 of scikit-image's (the 1e-5 bar) at a fraction of the FP64 cost.
template <bool EXACT = true>
__device__ __forceinline__ float warp_sample(const float* __restrict__ img, int side, int off, float xs, float ys) {
  // img is a side x side image embedded at offset `off` inside a zero frame
  // (the padded frame of compute_attenuation): pixel (r, c) of the frame is
  // img[(r - off) * side + (c - off)] when inside, else 0.
  const float c0f = floorf(xs);
  const float r0f = floorf(ys);
  const float dcf = xs - c0f;
  const float drf = ys - r0f;
  const int c0 = static_cast<int>(c0f) - off;
  const int r0 = static_cast<int>(r0f) - off;
  const int c1 = static_cast<int>(ceilf(xs)) - off;
  const int r1 = static_cast<int>(ceilf(ys)) - off;
  auto px = [&](int r, int c) -> float {
    return (r >= 0 && r < side && c >= 0 && c < side) ? img[static_cast<size_t>(r) * side + c] : 0.f;
  };
  const float tl = px(r0, c0), tr = px(r0, c1), bl = px(r1, c0), br = px(r1, c1);
  if (!EXACT) {
    const float top = fmaf(dcf, tr - tl, tl);
    const float bot = fmaf(dcf, br - bl, bl);
    return fmaf(drf, bot - top, top);
  }
  const double dc = dcf, dr = drf;
  // explicit roundings: no FMA contraction, as in the x86-64 build of scikit-image
  const double top = dadd_rn(dmul_rn(1.0 - dc, static_cast<double>(tl)), static_cast<double>(mul_rn(dcf, tr)));
  const double bot = dadd_rn(dmul_rn(1.0 - dc, static_cast<double>(bl)), static_cast<double>(mul_rn(dcf, br)));
  return static_cast<float>(dadd_rn(dmul_rn(1.0 - dr, top), dmul_rn(dr, bot)));
}

// The float32 blend with each row's two taps read as one 8-byte load (pixels c0 and
// c0 + 1; when xs is an integer the right tap's weight dc is 0, so reading c0 + 1
// instead of ceil(xs) changes nothing, and likewise for the rows).
__device__ __forceinline__ float warp_sample_pairs(const float* __restrict__ img, int side, int off, float xs,
                                                   float ys) {
  const float c0f = floorf(xs);
  const float r0f = floorf(ys);
  const float dc = xs - c0f;
  const float dr = ys - r0f;
  const int c0 = static_cast<int>(c0f) - off;
  const int r0 = static_cast<int>(r0f) - off;
  auto row = [&](int r, float& a, float& b) {
    a = b = 0.f;
    if (static_cast<unsigned>(r) >= static_cast<unsigned>(side)) return;
    const float* p = img + static_cast<size_t>(r) * side;
    if (c0 >= 0 && c0 + 1 < side) {
      float2 v;
      __builtin_memcpy(&v, p + c0, sizeof(float2));
      a = v.x;
      b = v.y;
    } else {
      if (static_cast<unsigned>(c0) < static_cast<unsigned>(side)) a = p[c0];
      if (static_cast<unsigned>(c0 + 1) < static_cast<unsigned>(side)) b = p[c0 + 1];
    }
  };
  float tl, tr, bl, br;
  row(r0, tl, tr);
  row(r0 + 1, bl, br);
  const float top = fmaf(dc, tr - tl, tl);
  const float bot = fmaf(dc, br - bl, bl);
  return fmaf(dr, bot - top, top);
}

template <bool EXACT>
__device__ __forceinline__ float warp_any(const float* __restrict__ img, int side, int off, float xs, float ys) {
  return EXACT ? warp_sample<true>(img, side, off, xs, ys) : warp_sample_pairs(img, side, off, xs, ys);
}

__device__ __forceinline__ void affine_apply(const Affine& m, float x, float y, float& xs, float& ys) {
  xs = add_rn(add_rn(mul_rn(m.m00, x), mul_rn(m.m01, y)), m.m02);
  ys = add_rn(add_rn(mul_rn(m.m10, x), mul_rn(m.m11, y)), m.m12);
}

// beams [0, n_in) use mu_in with rot[rot_in0 + beam], the rest mu_out with
// rot[rot_out0 + (beam - n_in)]; output C[beam][r][x], side np.
//
// A wave covers 8 columns x 8 rows of the rotated frame: lane = 8 * column
// + row. Its samples form a compact rotated 8 x 8 patch of mu, so each
// gather touches a few cache lines (a wave along one rotated row touched one
// line per lane). The prefix sum stays numpy's sequential float32 sum: within
// a column's 8 lanes the running value is passed from row to row in order
// (run_r = run_{r-1} + v_r), and lane 7 carries it to the next band of rows.
constexpr int kScanCols = 8;   // columns per wave
constexpr int kScanBand = 8;   // rows per band (lanes per column)
constexpr int kScanWaves = 4;  // waves per workgroup

// Samples whose taps all lie outside the unpadded slice are 0 (the rotated slice
// fills about half of the padded frame): they skip the gathers and the blend.
template <bool EXACT>
__global__ __launch_bounds__(256) void att_scan_kernel(const float* __restrict__ mu_in, const float* __restrict__ mu_out,
                                                       const Affine* __restrict__ rot_in,
                                                       const Affine* __restrict__ rot_out, int n_in, int n_beams,
                                                       int n, int pad, int np, float* __restrict__ cum) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int col = lane / kScanBand;
  const int row = lane - col * kScanBand;
  const int x = (blockIdx.x * kScanWaves + wave) * kScanCols + col;
  const int beam = blockIdx.y;
  if (beam >= n_beams) return;  // uniform over the workgroup
  // a band of the workgroup's 32 columns goes through LDS so that its rows are stored as
  // whole 128-byte segments (a wave's own 8 columns are 32 bytes per row)
  __shared__ float band[kScanBand][kScanWaves * kScanCols];
  __shared__ int rows_lo[kScanWaves], rows_hi[kScanWaves];
  const bool inb = beam < n_in;
  const float* mu = inb ? mu_in : mu_out;
  const Affine m = inb ? rot_in[beam] : rot_out[beam - n_in];
  const float xf = static_cast<float>(min(x, np - 1));
  float xs, ys;
  // a sample reads the slice only inside (pad - 1, pad + n) in both coordinates
  const float lo_f = static_cast<float>(pad - 1), hi_f = static_cast<float>(pad + n);
  auto sample = [&](float sx, float sy) -> float {
    if (sx > lo_f && sx < hi_f && sy > lo_f && sy < hi_f) return warp_any<EXACT>(mu, n, pad, sx, sy);
    return 0.f;
  };
  auto band_sample = [&](int r0) -> float {  // this lane's sample of the band starting at row r0
    float bx, by;
    affine_apply(m, xf, static_cast<float>(min(r0 + row, np - 1)), bx, by);
    return sample(bx, by);
  };
  // R[np - 1]: the roll wraps it onto row 0
  affine_apply(m, xf, static_cast<float>(np - 1), xs, ys);
  float last = sample(xs, ys);  // R of the row before the band, per column
  float carry = 0.f;                              // running sum at the row before the band
  // lane-to-lane moves within the 8-lane column groups run on DPP (VALU, no LDS round trip):
  // row_shr:1 hands each lane its predecessor's value (row r - 1 of the column; unused at row 0),
  // row_ror:9 hands row 0 the value of row 7 of its own group (a 16-lane DPP row holds two groups)
  auto from_prev = [](float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xF, 0xF, false));
  };
  auto from_row7 = [](float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x129, 0xF, 0xF, false));
  };
  // Rows of this wave's 8 columns that can be read later: the compose kernel reads a plane
  // only within 1.5 px of where the slice lies in the rotated frame, and R is 0 more than a
  // pixel outside it. Per column the rows whose samples fall within 3 px of the slice form an
  // interval; the wave scans the band-aligned union of its columns' intervals. Before it,
  // every R is 0, so the running sum is C[0] = (0 + R[np - 1]) / 2 (the roll) and R[r - 1]
  // is 0; rows after it are never read and are not written.
  int r_first = 0, r_last = np - 1;
  {
    const double lo_m = static_cast<double>(pad) - 4.0, hi_m = static_cast<double>(pad + n) + 3.0;
    double t = 0.0, b = static_cast<double>(np - 1);
    auto clamp_rows = [&](double a_r, double c0) {  // lo_m < a_r r + c0 < hi_m
      if (fabs(a_r) < 1e-12) {
        if (!(c0 > lo_m && c0 < hi_m)) b = -1.0;
        return;
      }
      double r1 = (lo_m - c0) / a_r, r2 = (hi_m - c0) / a_r;
      if (r1 > r2) { const double tmp = r1; r1 = r2; r2 = tmp; }
      t = fmax(t, r1);
      b = fmin(b, r2);
    };
    const double xd = static_cast<double>(min(x, np - 1));
    clamp_rows(m.m01, m.m00 * xd + m.m02);
    clamp_rows(m.m11, m.m10 * xd + m.m12);
    int lo_r = b < t ? np : static_cast<int>(floor(t)) - 1;
    int hi_r = b < t ? -1 : static_cast<int>(ceil(b)) + 1;
    // union over the wave's columns (lanes 8 c + row; all lanes of a column agree)
#pragma unroll
    for (int sh = 8; sh < 64; sh <<= 1) {
      lo_r = min(lo_r, __shfl_xor(lo_r, sh, 64));
      hi_r = max(hi_r, __shfl_xor(hi_r, sh, 64));
    }
    // union over the workgroup's waves: every wave walks the same bands (rows outside a
    // wave's own interval have R = 0 there and are computed like any other)
    if (lane == 0) {
      rows_lo[wave] = lo_r;
      rows_hi[wave] = hi_r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanWaves; ++k) {
      lo_r = min(lo_r, rows_lo[k]);
      hi_r = max(hi_r, rows_hi[k]);
    }
    r_first = max(lo_r, 0) & ~(kScanBand - 1);
    r_last = min(hi_r, np - 1);
  }
  if (r_last < r_first) return;  // no column of the workgroup meets the slice (uniform)
  if (r_first > 0) {
    carry = (0.f + last) * 0.5f;  // C[0] with R[0] = 0; rows 1 .. r_first - 1 add 0
    last = 0.f;                   // R[r_first - 1]
  }
  // the samples of the next kAhead bands are in flight while this band is summed: only the
  // running sum is carried from band to band, the gathers do not depend on it
  constexpr int kAhead = 4;
  float ahead[kAhead];
#pragma unroll
  for (int k = 0; k < kAhead; ++k) ahead[k] = band_sample(r_first + k * kScanBand);
  for (int r0 = r_first; r0 <= r_last; r0 += kScanBand) {
    const float cur = ahead[0];
#pragma unroll
    for (int k = 0; k + 1 < kAhead; ++k) ahead[k] = ahead[k + 1];
    ahead[kAhead - 1] = band_sample(r0 + kAhead * kScanBand);
    // R[r - 1]: the previous lane of the column, or the last row of the previous band
    const float up = from_prev(cur);
    const float prev = (row == 0) ? last : up;
    const float v = (cur + prev) * 0.5f;
    float run = (row == 0) ? ((r0 == 0) ? v : carry + v) : v;
#pragma unroll
    for (int k = 1; k < kScanBand; ++k) {
      const float before = from_prev(run);
      if (row == k) run = before + v;
    }
    carry = from_row7(run);  // used by row 0 only
    last = from_row7(cur);
    band[row][wave * kScanCols + col] = run;
    __syncthreads();
    {
      const int br = threadIdx.x / (kScanWaves * kScanCols), bc = threadIdx.x % (kScanWaves * kScanCols);
      const int xs0 = blockIdx.x * kScanWaves * kScanCols + bc;
      if (xs0 < np && r0 + br <= r_last) cum[static_cast<size_t>(beam) * np * np + static_cast<size_t>(r0 + br) * np + xs0] = band[br][bc];
    }
    __syncthreads();
  }
}

// maps[a - out_a0][d][i][j] for angles [a0, a0 + n_a) of the chunk whose
// prefix sums are in cum: beam index of angle a (chunk-local la): in = la
// (if has_in), out = n_in + (la * D + d) * n_sub + k (if has_out) for the
// detector's nsub[d] sub-angles k, whose exp(-.) are averaged in float32 in
// order (physics/attenuation.py:110-119).
template <bool EXACT>
__global__ __launch_bounds__(256) void att_compose_kernel(const float* __restrict__ cum,
                                                          const Affine* __restrict__ rot_in_back,
                                                          const Affine* __restrict__ rot_out_back, int has_in,
                                                          int has_out, int n_in, int a0, int n_det, int n, int pad,
                                                          int np, float* __restrict__ maps, int out_a0, int n_sub,
                                                          const int* __restrict__ nsub) {
  // a workgroup owns a 32 x 16 tile, each thread two pixels 8 rows apart (their gathers
  // are in flight together); each wave takes 8 columns of both 8-row halves, so its
  // back-rotated samples form compact patches of the prefix-sum planes. The tile's maps
  // go out through LDS as 128-byte rows (a wave's own columns are 32 bytes per row).
  __shared__ float tile_out[16][32];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j_raw = blockIdx.x * 32 + wave * 8 + (lane & 7);
  const int i0_raw = blockIdx.y * 16 + (lane >> 3);
  const int j = min(j_raw, n - 1), i0 = min(i0_raw, n - 1);
  const int ad = blockIdx.z;  // chunk-local (la, d)
  const int la = ad / n_det;
  const int d = ad - la * n_det;
  const int a = a0 + la;
  const size_t npp = static_cast<size_t>(np) * np;
  const float xf = static_cast<float>(j + pad);
  constexpr int kPix = 2;
  float e_in[kPix], e_out[kPix];
#pragma unroll
  for (int h = 0; h < kPix; ++h) e_in[h] = e_out[h] = 1.f;
  if (has_in) {
    const Affine m = rot_in_back[a];
    float v[kPix];
#pragma unroll
    for (int h = 0; h < kPix; ++h) {
      float xs, ys;
      affine_apply(m, xf, static_cast<float>(min(i0 + 8 * h, n - 1) + pad), xs, ys);
      v[h] = warp_any<EXACT>(cum + static_cast<size_t>(la) * npp, np, 0, xs, ys);
    }
#pragma unroll
    for (int h = 0; h < kPix; ++h) e_in[h] = expf(-v[h]);
  }
  if (has_out) {
    const int cnt = nsub ? nsub[d] : 1;
    float acc[kPix];
#pragma unroll
    for (int h = 0; h < kPix; ++h) acc[h] = 0.f;
    for (int k = 0; k < cnt; ++k) {
      const Affine m = rot_out_back[(static_cast<size_t>(a) * n_det + d) * n_sub + k];
      const float* plane = cum + static_cast<size_t>(n_in + (la * n_det + d) * n_sub + k) * npp;
      float v[kPix];
#pragma unroll
      for (int h = 0; h < kPix; ++h) {
        float xs, ys;
        affine_apply(m, xf, static_cast<float>(min(i0 + 8 * h, n - 1) + pad), xs, ys);
        v[h] = warp_any<EXACT>(plane, np, 0, xs, ys);
      }
#pragma unroll
      for (int h = 0; h < kPix; ++h) acc[h] += expf(-v[h]);
    }
#pragma unroll
    for (int h = 0; h < kPix; ++h) e_out[h] = acc[h] / static_cast<float>(cnt);
  }
#pragma unroll
  for (int h = 0; h < kPix; ++h) tile_out[(lane >> 3) + 8 * h][wave * 8 + (lane & 7)] = e_in[h] * e_out[h];
  __syncthreads();
  float* dst = maps + (static_cast<size_t>(a - out_a0) * n_det + d) * n * n;
  const int oc = blockIdx.x * 32 + (threadIdx.x & 31);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int tr = (threadIdx.x >> 5) + 8 * h;
    const int i = blockIdx.y * 16 + tr;
    if (i < n && oc < n) dst[static_cast<size_t>(i) * n + oc] = tile_out[tr][threadIdx.x & 31];
  }
}

// Grows a device buffer to `need` bytes. Work already queued on `st` may still
// read the old buffer, so the stream is drained first.
int ensure_buffer(float*& buf, size_t& bytes, size_t need, hipStream_t st, const char* what) {
  if (bytes >= need) return CHIP_OK;
  if (buf) {
    CHIP_HIP(hipStreamSynchronize(st));
    CHIP_HIP(hipFree(buf));
    buf = nullptr;
    bytes = 0;
  }
  hipError_t e = hipMalloc(&buf, need);
  if (e != hipSuccess) {
    set_error(std::string(what) + ": scratch hipMalloc of " + std::to_string(need) + " bytes -> " +
              hipGetErrorString(e));
    return CHIP_ENOMEM;
  }
  bytes = need;
  return CHIP_OK;
}

// Angles per chunk: prefix sums of `beams` images of Np^2 per angle (plus a
// D x N^2 maps block per angle when `with_block`) within `budget` bytes, and
// within the launch-grid limit of 65535 in y / z.
static int chunk_angles(const Plan& p, int beams, bool with_block, size_t budget) {
  // beams per angle include n_sub out-beams per detector
  const size_t npp = static_cast<size_t>(p.np) * p.np;
  const size_t per_angle = npp * sizeof(float) * beams +
                           (with_block ? static_cast<size_t>(p.n_det) * p.n * p.n * sizeof(float) : 0);
  int chunk = static_cast<int>(std::max<size_t>(1, budget / per_angle));
  chunk = std::min(chunk, p.n_angles);
  chunk = std::min(chunk, 65535 / std::max(beams, p.n_det));
  return chunk;
}

// Maps of angles [a0, a0 + na) of one slice into out[(a - out_a0)][d][n][n],
// through the prefix-sum scratch `scan` (>= na beams' images).
static int build_maps_chunk(const Plan& p, const float* mi, const float* mo, int a0, int na, float* out, int out_a0,
                            float* scan, hipStream_t st) {
  const int A = p.n_angles, D = p.n_det;
  const Affine* rin_f = p.d_rot;
  const Affine* rin_b = p.d_rot + A;
  const int S = p.n_sub;
  const Affine* rout_f = p.d_rot + 2 * A;
  const Affine* rout_b = p.d_rot + 2 * A + static_cast<size_t>(A) * D * S;
  const int n_in = mi ? na : 0;
  const int n_beams = n_in + (mo ? na * D * S : 0);
  dim3 g1((p.np + kScanWaves * kScanCols - 1) / (kScanWaves * kScanCols), n_beams);
  dim3 g2((p.n + 31) / 32, (p.n + 15) / 16, na * D);
  if (p.maps_exact) {
    hipLaunchKernelGGL(att_scan_kernel<true>, g1, dim3(256), 0, st, mi, mo, rin_f + a0,
                       rout_f + static_cast<size_t>(a0) * D * S, n_in, n_beams, p.n, p.pad, p.np, scan);
    CHIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(att_compose_kernel<true>, g2, dim3(256), 0, st, scan, rin_b, rout_b, mi ? 1 : 0,
                       mo ? 1 : 0, n_in, a0, D, p.n, p.pad, p.np, out, out_a0, S, p.d_nsub);
  } else {
    hipLaunchKernelGGL(att_scan_kernel<false>, g1, dim3(256), 0, st, mi, mo, rin_f + a0,
                       rout_f + static_cast<size_t>(a0) * D * S, n_in, n_beams, p.n, p.pad, p.np, scan);
    CHIP_LAUNCH_CHECK();
    hipLaunchKernelGGL(att_compose_kernel<false>, g2, dim3(256), 0, st, scan, rin_b, rout_b, mi ? 1 : 0,
                       mo ? 1 : 0, n_in, a0, D, p.n, p.pad, p.np, out, out_a0, S, p.d_nsub);
  }
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int att_maps_slices(const Plan& p, const float* mu_in, const float* mu_out, float* maps, int batch, int chunk,
                    float* scan, hipStream_t st) {
  const int A = p.n_angles, D = p.n_det, n = p.n;
  const size_t nn = static_cast<size_t>(n) * n;
  for (int b = 0; b < batch; ++b) {
    const float* mi = mu_in ? mu_in + b * nn : nullptr;
    const float* mo = mu_out ? mu_out + b * nn : nullptr;
    float* mb = maps + static_cast<size_t>(b) * A * D * nn;
    for (int a0 = 0; a0 < A; a0 += chunk) {
      const int rc = build_maps_chunk(p, mi, mo, a0, std::min(chunk, A - a0), mb, 0, scan, st);
      if (rc != CHIP_OK) return rc;
    }
  }
  return CHIP_OK;
}

int maps_block_angles(const Plan& p, bool in, bool out, bool with_block, size_t budget, size_t* scan_bytes,
                      size_t* block_bytes) {
  const int beams = (in ? 1 : 0) + (out ? p.n_det * p.n_sub : 0);
  const int chunk = chunk_angles(p, beams, with_block, budget);
  if (scan_bytes) *scan_bytes = static_cast<size_t>(chunk) * beams * p.np * p.np * sizeof(float);
  if (block_bytes) *block_bytes = with_block ? static_cast<size_t>(chunk) * p.n_det * p.n * p.n * sizeof(float) : 0;
  return chunk;
}

int otf_fp(const Plan& p, const float* vol, const float* mu_in, const float* mu_out, float* sino, int batch,
           int chunk, float* scan, float* block, hipStream_t st) {
  const size_t nn = static_cast<size_t>(p.n) * p.n;
  ScratchUse use(*p.fp_state, st);
  const int rc0 = fp_range_ready(p, chunk, st);
  if (rc0 != CHIP_OK) return rc0;
  for (int b = 0; b < batch; ++b) {
    const float* mi = mu_in ? mu_in + b * nn : nullptr;
    const float* mo = mu_out ? mu_out + b * nn : nullptr;
    const float* v = vol + b * nn;
    float* s = sino + static_cast<size_t>(b) * p.n_det * p.n_angles * p.n;
    for (int a0 = 0; a0 < p.n_angles; a0 += chunk) {
      const int na = std::min(chunk, p.n_angles - a0);
      const int rc = build_maps_chunk(p, mi, mo, a0, na, block, a0, scan, st);
      if (rc != CHIP_OK) return rc;
      launch_fp_range(p, v, block, s, a0, na, a0, st);
      CHIP_LAUNCH_CHECK();
    }
  }
  return CHIP_OK;
}

int otf_bp(const Plan& p, const float* sino, const float* mu_in, const float* mu_out, float* vol, int batch,
           int mode, int chunk, float* scan, float* block, hipStream_t st) {
  const size_t nn = static_cast<size_t>(p.n) * p.n;
  // only the first chunk of each slice writes a fresh output and may split into parts
  BpPartGuard guard(p, 1, mode, chunk, st);
  for (int b = 0; b < batch; ++b) {
    const float* mi = mu_in ? mu_in + b * nn : nullptr;
    const float* mo = mu_out ? mu_out + b * nn : nullptr;
    const float* s = sino + static_cast<size_t>(b) * p.n_det * p.n_angles * p.n;
    float* v = vol + b * nn;
    for (int a0 = 0; a0 < p.n_angles; a0 += chunk) {
      const int na = std::min(chunk, p.n_angles - a0);
      const int rc = build_maps_chunk(p, mi, mo, a0, na, block, a0, scan, st);
      if (rc != CHIP_OK) return rc;
      launch_bp_range(p, s, block, v, mode, a0, na, a0, a0 > 0 ? 1 : 0, st, guard.scratch());
      CHIP_LAUNCH_CHECK();
    }
  }
  return CHIP_OK;
}

// The data step of one angle block, between its fp and its bp: for slice `sb`'s rows
// [a0, a0 + na) of every detector (sinogram [D][A][N] per slice), p = the row op of ax.
__global__ __launch_bounds__(256) void row_op_kernel(float* __restrict__ p, const float* __restrict__ ax,
                                                     const float* __restrict__ sigma,
                                                     const float* __restrict__ b, const float* __restrict__ bg,
                                                     const float* __restrict__ mask, int op, int64_t base,
                                                     int n_det, int n_ang, int n, int a0, int na) {
  const int64_t per_det = static_cast<int64_t>(na) * n;
  const int64_t total = per_det * n_det;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int d = static_cast<int>(t / per_det);
    const int64_t i = base + (static_cast<int64_t>(d) * n_ang + a0) * n + (t - d * per_det);
    float a = ax[i];
    if (bg) a = a + bg[i];
    float v;
    if (op == CHIP_ROW_PDHG_L2) {
      v = elem_dual_l2(p[i], a, sigma[i], b[i]);
      if (mask) v = v * mask[i];
    } else if (op == CHIP_ROW_PDHG_KL) {
      v = elem_dual_kl(p[i], a, sigma[i], b[i]);
      if (mask) v = v * mask[i];
    } else if (op == CHIP_ROW_SIRT) {
      v = elem_residual(a, b[i]);
      if (mask) v = v * mask[i];
      if (sigma) v = v * sigma[i];
    } else {
      v = elem_mlem_ratio(a, b[i]);
    }
    p[i] = v;
  }
}

int otf_fp_row_bp(const Plan& p, const float* x, const float* mu_in, const float* mu_out, int batch, int op,
                  float* pd, const float* sigma, const float* b, const float* bg, const float* mask, float* ax,
                  float* vol, int mode, int chunk, float* scan, float* block, hipStream_t st) {
  const size_t nn = static_cast<size_t>(p.n) * p.n;
  const int64_t sino_n = static_cast<int64_t>(p.n_det) * p.n_angles * p.n;
  ScratchUse use(*p.fp_state, st);
  const int rc0 = fp_range_ready(p, chunk, st);
  if (rc0 != CHIP_OK) return rc0;
  BpPartGuard guard(p, 1, mode, chunk, st);
  for (int sb = 0; sb < batch; ++sb) {
    const float* mi = mu_in ? mu_in + sb * nn : nullptr;
    const float* mo = mu_out ? mu_out + sb * nn : nullptr;
    const int64_t base = sb * sino_n;
    float* v = vol + sb * nn;
    for (int a0 = 0; a0 < p.n_angles; a0 += chunk) {
      const int na = std::min(chunk, p.n_angles - a0);
      const int rc = build_maps_chunk(p, mi, mo, a0, na, block, a0, scan, st);
      if (rc != CHIP_OK) return rc;
      launch_fp_range(p, x + sb * nn, block, ax + base, a0, na, a0, st);
      CHIP_LAUNCH_CHECK();
      const int64_t total = static_cast<int64_t>(p.n_det) * na * p.n;
      const unsigned blocks = static_cast<unsigned>(std::min<int64_t>((total + 255) / 256, 65536));
      hipLaunchKernelGGL(row_op_kernel, dim3(blocks), dim3(256), 0, st, pd, ax, sigma, b, bg, mask, op, base,
                         p.n_det, p.n_angles, p.n, a0, na);
      CHIP_LAUNCH_CHECK();
      launch_bp_range(p, pd + base, block, v, mode, a0, na, a0, a0 > 0 ? 1 : 0, st, guard.scratch());
      CHIP_LAUNCH_CHECK();
    }
  }
  return CHIP_OK;
}

// out[i][z][y][x] = mean of the b^nd inputs of the block at (z, y, x) after
// cropping the volume axes at (o0, o1, o2): summed in row-major order in
// float32, divided by b^nd (skimage block_reduce with np.mean)
__global__ __launch_bounds__(256) void bin_blocks_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                         int64_t n_img, int d0, int d1, int d2, int o0, int o1,
                                                         int o2, int b0, int b, int e0, int e1, int e2) {
  const int64_t per = static_cast<int64_t>(e0) * e1 * e2;
  const int64_t total = n_img * per;
  const float cnt = static_cast<float>(b0 * b * b);
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t img = t / per;
    int64_t r = t - img * per;
    const int z = static_cast<int>(r / (static_cast<int64_t>(e1) * e2));
    r -= static_cast<int64_t>(z) * e1 * e2;
    const int y = static_cast<int>(r / e2);
    const int x = static_cast<int>(r - static_cast<int64_t>(y) * e2);
    const float* src = in + img * static_cast<int64_t>(d0) * d1 * d2;
    float acc = 0.f;
    for (int k = 0; k < b0; ++k)
      for (int i = 0; i < b; ++i) {
        const float* row = src + (static_cast<int64_t>(o0 + z * b0 + k) * d1 + (o1 + y * b + i)) * d2 + o2 + x * b;
        for (int j = 0; j < b; ++j) acc += row[j];
      }
    out[t] = acc / cnt;
  }
}

}  // namespace chip

using namespace chip;

extern "C" int chip_bin_images(const float* in, float* out, int64_t n_img, int nd, const int* dims, int binning,
                               void* stream) {
  clear_error();
  CHIP_CHECK_ARG(in != nullptr && out != nullptr && dims != nullptr, "NULL argument");
  CHIP_CHECK_ARG(n_img >= 1 && (nd == 2 || nd == 3) && binning >= 1, "need n_img >= 1, nd in {2, 3}, binning >= 1");
  int d[3] = {1, 1, 1};
  for (int k = 0; k < nd; ++k) {
    CHIP_CHECK_ARG(dims[k] >= binning, "every binned axis must hold at least one block");
    d[3 - nd + k] = dims[k];
  }
  const int b0 = nd == 3 ? binning : 1;
  int o[3], e[3];
  for (int k = 0; k < 3; ++k) {
    const int bk = k == 0 ? b0 : binning;
    const int excess = d[k] % bk;
    o[k] = excess - excess / 2;  // auto_crop: the larger half of the excess from the start
    e[k] = d[k] / bk;
  }
  const int64_t total = n_img * e[0] * e[1] * e[2];
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(bin_blocks_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, as_stream(stream), in, out,
                     n_img, d[0], d[1], d[2], o[0], o[1], o[2], b0, binning, e[0], e[1], e[2]);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

extern "C" int chip_att_maps(chip_plan* plan, const float* mu_in, const float* mu_out, float* maps, int batch,
                             void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(maps != nullptr, "maps is NULL");
  CHIP_CHECK_ARG(mu_in != nullptr || mu_out != nullptr, "no attenuation volumes were given");
  CHIP_CHECK_ARG(batch >= 1, "batch must be >= 1");
  Plan& p = plan->p;
  hipStream_t st = as_stream(stream);
  size_t need = 0;
  const int chunk = maps_block_angles(p, mu_in != nullptr, mu_out != nullptr, false, kMapsScanBudget, &need, nullptr);
  const int rc = ensure_buffer(p.d_scratch, p.scratch_bytes, need, st, "chip_att_maps");
  if (rc != CHIP_OK) return rc;
  return att_maps_slices(p, mu_in, mu_out, maps, batch, chunk, p.d_scratch, st);
}

// Shared set-up of the on-the-fly projectors: chunk size and both scratch buffers.
static int otf_prepare(Plan& p, const float* mu_in, const float* mu_out, hipStream_t st, int& chunk) {
  size_t need_c = 0, need_m = 0;
  chunk = maps_block_angles(p, mu_in != nullptr, mu_out != nullptr, true, p.otf_budget, &need_c, &need_m);
  int rc = ensure_buffer(p.d_scratch, p.scratch_bytes, need_c, st, "otf prefix sums");
  if (rc != CHIP_OK) return rc;
  return ensure_buffer(p.d_mblock, p.mblock_bytes, need_m, st, "otf maps block");
}

extern "C" int chip_fp_otf(chip_plan* plan, const float* vol, const float* mu_in, const float* mu_out, float* sino,
                           int batch, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(vol != nullptr && sino != nullptr, "vol/sino is NULL");
  CHIP_CHECK_ARG(mu_in != nullptr || mu_out != nullptr, "no attenuation volumes were given");
  CHIP_CHECK_ARG(batch >= 1, "batch must be >= 1");
  Plan& p = plan->p;
  hipStream_t st = as_stream(stream);
  int chunk = 0;
  const int rc = otf_prepare(p, mu_in, mu_out, st, chunk);
  if (rc != CHIP_OK) return rc;
  return otf_fp(p, vol, mu_in, mu_out, sino, batch, chunk, p.d_scratch, p.d_mblock, st);
}

extern "C" int chip_bp_otf(chip_plan* plan, const float* sino, const float* mu_in, const float* mu_out, float* vol,
                           int batch, int mode, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(vol != nullptr && sino != nullptr, "vol/sino is NULL");
  CHIP_CHECK_ARG(mu_in != nullptr || mu_out != nullptr, "no attenuation volumes were given");
  CHIP_CHECK_ARG(batch >= 1, "batch must be >= 1");
  CHIP_CHECK_ARG(mode == CHIP_BP_SKIMAGE_SYM || mode == CHIP_BP_ADJOINT,
                 "mode must be CHIP_BP_SKIMAGE_SYM or CHIP_BP_ADJOINT (CHIP_BP_SKIMAGE uses no maps: call chip_bp)");
  Plan& p = plan->p;
  hipStream_t st = as_stream(stream);
  int chunk = 0;
  const int rc = otf_prepare(p, mu_in, mu_out, st, chunk);
  if (rc != CHIP_OK) return rc;
  return otf_bp(p, sino, mu_in, mu_out, vol, batch, mode, chunk, p.d_scratch, p.d_mblock, st);
}

extern "C" int chip_otf_fp_row_bp(chip_plan* plan, const float* x, const float* mu_in, const float* mu_out, int batch,
                                  int row_op, float* p, const float* sigma, const float* b, const float* bg,
                                  const float* b_mask, float* ax, float* vol, int bp_mode, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  CHIP_CHECK_ARG(x != nullptr && p != nullptr && b != nullptr && ax != nullptr && vol != nullptr, "NULL argument");
  CHIP_CHECK_ARG(mu_in != nullptr || mu_out != nullptr, "no attenuation volumes were given");
  CHIP_CHECK_ARG(batch >= 1, "batch must be >= 1");
  CHIP_CHECK_ARG(row_op >= CHIP_ROW_PDHG_L2 && row_op <= CHIP_ROW_MLEM, "unknown row_op");
  CHIP_CHECK_ARG(sigma != nullptr || row_op == CHIP_ROW_SIRT || row_op == CHIP_ROW_MLEM,
                 "the PDHG row ops need sigma");
  CHIP_CHECK_ARG(bp_mode == CHIP_BP_SKIMAGE_SYM || bp_mode == CHIP_BP_ADJOINT,
                 "bp_mode must be CHIP_BP_SKIMAGE_SYM or CHIP_BP_ADJOINT");
  CHIP_CHECK_ARG(p != ax, "p and ax must be different buffers");
  Plan& pl = plan->p;
  hipStream_t st = as_stream(stream);
  int chunk = 0;
  const int rc = otf_prepare(pl, mu_in, mu_out, st, chunk);
  if (rc != CHIP_OK) return rc;
  return otf_fp_row_bp(pl, x, mu_in, mu_out, batch, row_op, p, sigma, b, bg, b_mask, ax, vol, bp_mode, chunk,
                       pl.d_scratch, pl.d_mblock, st);
}
