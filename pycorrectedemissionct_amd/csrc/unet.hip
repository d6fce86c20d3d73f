// Building blocks of the warm-start U-Net (gfx950): BF16 activations and
// weights, FP32 accumulation on the matrix cores.
//
// Layouts: activations NHWC bf16 (channels contiguous), conv weights
// [9 taps][Cout][Cin] bf16 (tap = 3 * (dh + 1) + (dw + 1)), biases and master
// weights fp32. A 3x3 convolution ("same" zero padding) is an implicit GEMM
// with M = pixels, N = output channels, K = 9 taps x input channels; the
// input may be the channel concatenation of two tensors (the U-Net's skip
// connections) and the output may be split in two (the gradient of such a
// concatenation).
//
// Matrix cores: v_mfma_f32_16x16x32_bf16 for the convolutions whose channel
// counts are multiples of 32 (conv3x3_mfma_kernel: LDS-staged input patches),
// v_mfma_f32_16x16x16_bf16 for the rest (the 1-channel input layer) and the
// weight gradient. 16x16x16 operand layout (lane l, element i):
// A[m][k] at m = l % 16, k = 4 (l / 16) + i; B[k][n] at n = l % 16,
// k = 4 (l / 16) + i; D[m][n] at n = l % 16, m = 4 (l / 16) + i.
//
// Every reduction (weight / bias gradients, loss) writes per-chunk partial
// sums and adds them in a fixed order, so results are reproducible.
#include <algorithm>

#include "common.hpp"

namespace chip {
namespace {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

// float -> bf16, round to nearest even (NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);
  return static_cast<uint16_t>((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

struct Img {
  int n, h, w;
  __device__ __forceinline__ int64_t pixels() const { return static_cast<int64_t>(n) * h * w; }
};

// ---------------------------------------------------------------------------
// forward 3x3 convolution
// ---------------------------------------------------------------------------
constexpr int kPxWave = 32;            // pixels per wave (2 MFMA row blocks)
constexpr int kPxBlock = 4 * kPxWave;  // pixels per workgroup
constexpr int kCoBlock = 32;           // output channels per workgroup (2 MFMA column blocks)

// 4 consecutive input channels [ci, ci + 4) of pixel (n, hh, ww) of concat(x1, x2)
__device__ __forceinline__ s4v load_in4(const uint16_t* __restrict__ x1, int c1, const uint16_t* __restrict__ x2,
                                        int c2, int64_t pix, int ci, bool vec) {
  s4v r = {0, 0, 0, 0};
  const int cin = c1 + c2;
  if (vec) {  // c1, c2 multiples of 4: the four channels lie in one tensor
    if (ci < c1) {
      const uint2 v = *reinterpret_cast<const uint2*>(x1 + pix * c1 + ci);
      r = __builtin_bit_cast(s4v, v);
    } else if (ci < cin) {
      const uint2 v = *reinterpret_cast<const uint2*>(x2 + pix * c2 + (ci - c1));
      r = __builtin_bit_cast(s4v, v);
    }
    return r;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int c = ci + i;
    uint16_t v = 0;
    if (c < c1) v = x1[pix * c1 + c];
    else if (c < cin) v = x2[pix * c2 + (c - c1)];
    r[i] = static_cast<short>(v);
  }
  return r;
}

__global__ __launch_bounds__(256) void conv3x3_kernel(const uint16_t* __restrict__ x1, int c1,
                                                      const uint16_t* __restrict__ x2, int c2,
                                                      const uint16_t* __restrict__ wt, const float* __restrict__ bias,
                                                      uint16_t* __restrict__ y1, int o1, uint16_t* __restrict__ y2,
                                                      Img im, int cout, int relu) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int64_t P = im.pixels();
  const int64_t pbase = static_cast<int64_t>(blockIdx.x) * kPxBlock + wave * kPxWave;
  const int cbase = blockIdx.y * kCoBlock;
  const int cin = c1 + c2;
  const bool vec_in = (c1 % 4 == 0) && (c2 % 4 == 0);
  const bool vec_w = (cin % 4 == 0);
  const int kq = 4 * (lane >> 4);  // this lane's k offset within a 16-wide k step

  // the lane's A-operand pixel of each 16-pixel block
  int ph[2], pw[2], pn[2];
  bool pv[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int64_t p = pbase + 16 * b + (lane & 15);
    pv[b] = p < P;
    const int64_t pc = pv[b] ? p : 0;
    pw[b] = static_cast<int>(pc % im.w);
    const int64_t t = pc / im.w;
    ph[b] = static_cast<int>(t % im.h);
    pn[b] = static_cast<int>(t / im.h);
  }
  int co_l[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) co_l[c] = cbase + 16 * c + (lane & 15);

  f4v acc[2][2];
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[b][c] = f4v{0.f, 0.f, 0.f, 0.f};

  for (int tap = 0; tap < 9; ++tap) {
    const int dh = tap / 3 - 1, dw = tap % 3 - 1;
    int64_t src[2];
    bool sv[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int hh = ph[b] + dh, ww = pw[b] + dw;
      sv[b] = pv[b] && hh >= 0 && hh < im.h && ww >= 0 && ww < im.w;
      src[b] = (static_cast<int64_t>(pn[b]) * im.h + (sv[b] ? hh : 0)) * im.w + (sv[b] ? ww : 0);
    }
    const uint16_t* wtap = wt + static_cast<size_t>(tap) * cout * cin;
    for (int ci0 = 0; ci0 < cin; ci0 += 16) {
      const int ci = ci0 + kq;
      s4v a[2], bm[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        a[b] = sv[b] ? load_in4(x1, c1, x2, c2, src[b], ci, vec_in) : s4v{0, 0, 0, 0};
      }
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        s4v r = {0, 0, 0, 0};
        if (co_l[c] < cout) {
          const uint16_t* wr = wtap + static_cast<size_t>(co_l[c]) * cin;
          if (vec_w) {
            if (ci < cin) r = __builtin_bit_cast(s4v, *reinterpret_cast<const uint2*>(wr + ci));
          } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) r[i] = (ci + i < cin) ? static_cast<short>(wr[ci + i]) : 0;
          }
        }
        bm[c] = r;
      }
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[b][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[b], bm[c], acc[b][c], 0, 0, 0);
    }
  }

  // epilogue: D[m][n] at n = lane % 16, m = 4 (lane / 16) + i
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int co = co_l[c];
    if (co >= cout) continue;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t p = pbase + 16 * b + kq + i;
        if (p >= P) continue;
        float v = acc[b][c][i] + bb;
        if (relu) v = fmaxf(v, 0.f);
        if (co < o1) y1[p * o1 + co] = f2bf(v);
        else y2[p * (cout - o1) + (co - o1)] = f2bf(v);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// forward 3x3 convolution on gfx950's v_mfma_f32_16x16x32_bf16 (channel counts
// that are multiples of 32: every layer of the U-Net but the first)
//
// A workgroup owns a spatial tile of 4 RB rows x 16 columns of one image and
// CO_T output channels; wave w computes tile rows RB w .. RB w + RB - 1 (RB
// 16-pixel MFMA row blocks, each weight operand used RB times) for all CO_T
// channels. Input channels run in chunks of 32: the chunk's (4 RB + 2) x 18
// input patch (the tile plus its one-pixel halo, zero outside the image, from
// x1 or x2 of a concatenation) is staged in LDS, pixel pitch 40 bf16 (80 bytes:
// the 16 A-operand lanes of a read fall in distinct banks), and so are the
// chunk's weights; the 9 taps then read both operands from LDS.
// Operand layout (lane l, element i): A[m][k] at m = l % 16, k = 8 (l / 16) + i;
// B[k][n] at n = l % 16, k = 8 (l / 16) + i; D[m][n] at n = l % 16, m = 4 (l / 16) + i.
// ---------------------------------------------------------------------------
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

constexpr int kMtRows = 8, kMtCols = 16;                 // spatial tile
constexpr int kMtPitch = 40;                             // bf16 per staged pixel (32 channels + 8 pad)
constexpr int kMtPatch = (kMtRows + 2) * (kMtCols + 2);  // staged pixels

// The chunk's 9 x CO_T x 32 weights are staged in LDS too ([tap][co][ci], pitch 40 bf16: the
// 16 B-operand lanes of a read fall in distinct banks) after the patch, so the four waves read
// their B operands from LDS instead of each from L2 (85 against 117 us per 16-row, 64-channel
// launch at 256^2); the epilogue reuses the whole area.
template <int CO_T, int RB>
__global__ __launch_bounds__(256) void conv3x3_mfma_kernel(const uint16_t* __restrict__ x1, int c1,
                                                           const uint16_t* __restrict__ x2, int c2,
                                                           const uint16_t* __restrict__ wt,
                                                           const float* __restrict__ bias, uint16_t* __restrict__ y1,
                                                           int o1, uint16_t* __restrict__ y2, Img im, int cout,
                                                           int relu) {
  constexpr int kCb = CO_T / 16;  // MFMA column blocks
  constexpr int kRows = 4 * RB;    // tile rows: RB row blocks per wave
  constexpr int kPatch = (kRows + 2) * (kMtCols + 2);
  constexpr int kOutPitch = CO_T + 8;  // bf16 per pixel of the staged output tile
  constexpr int kWPitch = 40;          // bf16 per staged (tap, co) weight row of 32 ci
  constexpr int kInHalfs = kPatch * kMtPitch + 9 * CO_T * kWPitch;
  constexpr int kLdsHalfs = kInHalfs > kRows * kMtCols * kOutPitch ? kInHalfs : kRows * kMtCols * kOutPitch;
  __shared__ __align__(16) uint16_t patch[kLdsHalfs];
  uint16_t* const wl = patch + kPatch * kMtPitch;  // staged weights
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_w = (im.w + kMtCols - 1) / kMtCols, tiles_h = (im.h + kRows - 1) / kRows;
  int t = blockIdx.x;
  const int tx = t % tiles_w;
  t /= tiles_w;
  const int ty = t % tiles_h;
  const int img = t / tiles_h;
  const int h0 = ty * kRows, w0 = tx * kMtCols;
  const int cbase = blockIdx.y * CO_T;
  const int cin = c1 + c2;
  const int m = lane & 15, q = lane >> 4;

  f4v acc[RB][kCb];
#pragma unroll
  for (int b = 0; b < RB; ++b)
#pragma unroll
    for (int c = 0; c < kCb; ++c) acc[b][c] = f4v{0.f, 0.f, 0.f, 0.f};

  for (int ci0 = 0; ci0 < cin; ci0 += 32) {
    // the chunk lies wholly in x1 or in x2 (c1 % 32 == 0)
    const uint16_t* src = ci0 < c1 ? x1 : x2;
    const int cs = ci0 < c1 ? c1 : c2;
    const int co0 = ci0 < c1 ? ci0 : ci0 - c1;
    __syncthreads();  // the previous chunk's reads are done
    for (int k = threadIdx.x; k < kPatch * 4; k += 256) {
      const int px = k >> 2, part = k & 3;  // 4 x 16 bytes per pixel
      const int pr = px / (kMtCols + 2), pc = px - pr * (kMtCols + 2);
      const int hh = h0 + pr - 1, ww = w0 + pc - 1;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (hh >= 0 && hh < im.h && ww >= 0 && ww < im.w) {
        const int64_t pix = (static_cast<int64_t>(img) * im.h + hh) * im.w + ww;
        v = *reinterpret_cast<const uint4*>(src + pix * cs + co0 + 8 * part);
      }
      *reinterpret_cast<uint4*>(patch + px * kMtPitch + 8 * part) = v;
    }
    for (int k = threadIdx.x; k < 9 * CO_T * 4; k += 256) {
      const int row = k >> 2, part = k & 3;  // row = tap * CO_T + co, 4 x 16 bytes of 32 ci
      const int tap = row / CO_T, co = row - tap * CO_T;
      const uint16_t* wr = wt + (static_cast<size_t>(tap) * cout + cbase + co) * cin + ci0 + 8 * part;
      *reinterpret_cast<uint4*>(wl + row * kWPitch + 8 * part) = *reinterpret_cast<const uint4*>(wr);
    }
    __syncthreads();
#pragma unroll 1
    for (int tap = 0; tap < 9; ++tap) {
      const int dh = tap / 3, dw = tap % 3;  // offsets into the halo-padded patch
      bf8v bm[kCb];
#pragma unroll
      for (int c = 0; c < kCb; ++c) {
        bm[c] = __builtin_bit_cast(
            bf8v, *reinterpret_cast<const uint4*>(wl + (tap * CO_T + 16 * c + m) * kWPitch + 8 * q));
      }
#pragma unroll
      for (int b = 0; b < RB; ++b) {
        const int pr = RB * wave + b + dh, pc = m + dw;
        const bf8v a = __builtin_bit_cast(
            bf8v, *reinterpret_cast<const uint4*>(patch + (pr * (kMtCols + 2) + pc) * kMtPitch + 8 * q));
#pragma unroll
        for (int c = 0; c < kCb; ++c) acc[b][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bm[c], acc[b][c], 0, 0, 0);
      }
    }
  }

  // epilogue: D[m][n] with n = lane % 16 (channel), m = 4 (lane / 16) + i (tile column); the
  // tile goes through LDS as [pixel][channel] so that each pixel's channels leave as 16-byte stores
  __syncthreads();  // every wave is done with the patch
#pragma unroll
  for (int c = 0; c < kCb; ++c) {
    const int col = 16 * c + m;
    const float bb = bias ? bias[cbase + col] : 0.f;
#pragma unroll
    for (int b = 0; b < RB; ++b)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        float v = acc[b][c][i] + bb;
        if (relu) v = fmaxf(v, 0.f);
        patch[((RB * wave + b) * kMtCols + 4 * q + i) * kOutPitch + col] = f2bf(v);
      }
  }
  __syncthreads();
  constexpr int kGroups = CO_T / 8;
  for (int k = threadIdx.x; k < kRows * kMtCols * kGroups; k += 256) {
    const int px = k / kGroups, g = k - px * kGroups;
    const int hh = h0 + px / kMtCols, ww = w0 + px % kMtCols;
    if (hh >= im.h || ww >= im.w) continue;
    const int64_t p = (static_cast<int64_t>(img) * im.h + hh) * im.w + ww;
    const uint4 v = *reinterpret_cast<const uint4*>(patch + px * kOutPitch + 8 * g);
    const int co = cbase + 8 * g;  // the 8 channels lie on one side of o1 (o1 % 16 == 0)
    if (co < o1) *reinterpret_cast<uint4*>(y1 + p * o1 + co) = v;
    else *reinterpret_cast<uint4*>(y2 + p * (cout - o1) + (co - o1)) = v;
  }
}

// the 1-channel input layer (K = 9): a thread per pixel, 8 output channels per pass, the
// weights and biases in LDS, 16-byte stores
__global__ __launch_bounds__(256) void conv3x3_c1_kernel(const uint16_t* __restrict__ x,
                                                         const uint16_t* __restrict__ wt,
                                                         const float* __restrict__ bias, uint16_t* __restrict__ y,
                                                         Img im, int cout, int relu) {
  extern __shared__ float wsh[];  // [9][cout] weights, then [cout] biases
  for (int k = threadIdx.x; k < 9 * cout; k += blockDim.x) wsh[k] = bf2f(wt[k]);  // wt[tap][co][0]
  for (int k = threadIdx.x; k < cout; k += blockDim.x) wsh[9 * cout + k] = bias ? bias[k] : 0.f;
  __syncthreads();
  const int64_t P = im.pixels();
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const int ww = static_cast<int>(p % im.w);
  const int64_t t = p / im.w;
  const int hh = static_cast<int>(t % im.h);
  const int64_t img = t / im.h;
  float xv[9];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap) {
    const int h2 = hh + tap / 3 - 1, w2 = ww + tap % 3 - 1;
    xv[tap] = (h2 >= 0 && h2 < im.h && w2 >= 0 && w2 < im.w) ? bf2f(x[(img * im.h + h2) * im.w + w2]) : 0.f;
  }
  for (int co0 = 0; co0 < cout; co0 += 8) {
    u16x8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      float v = 0.f;
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) v = fmaf(xv[tap], wsh[tap * cout + co0 + e], v);
      v += wsh[9 * cout + co0 + e];
      if (relu) v = fmaxf(v, 0.f);
      o[e] = f2bf(v);
    }
    *reinterpret_cast<u16x8*>(y + p * cout + co0) = o;
  }
}

// ---------------------------------------------------------------------------
// weight gradient: part[chunk][tap][co][ci] = sum over the chunk's pixels p of
// dy[p][co] * x[p + offset(tap)][ci]   (M = co, N = ci, K = pixels)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint16_t in_at(const uint16_t* __restrict__ x1, int c1, const uint16_t* __restrict__ x2,
                                          int c2, int64_t pix, int c) {
  if (c < c1) return x1[pix * c1 + c];
  if (c < c1 + c2) return x2[pix * c2 + (c - c1)];
  return 0;
}

__global__ __launch_bounds__(256) void conv3x3_wgrad_kernel(const uint16_t* __restrict__ dy,
                                                            const uint16_t* __restrict__ x1, int c1,
                                                            const uint16_t* __restrict__ x2, int c2, Img im, int cout,
                                                            int64_t chunk_px, float* __restrict__ part) {
  __shared__ float red[4][64][16];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int cin = c1 + c2;
  const int n_cob = (cout + 31) / 32, n_cib = (cin + 31) / 32;
  int bid = blockIdx.x;
  const int cib = bid % n_cib;
  bid /= n_cib;
  const int cob = bid % n_cob;
  const int tap = bid / n_cob;
  const int dh = tap / 3 - 1, dw = tap % 3 - 1;
  const int chunk = blockIdx.y;
  const int64_t P = im.pixels();
  const int64_t p_lo = static_cast<int64_t>(chunk) * chunk_px;
  const int64_t p_hi = min(P, p_lo + chunk_px);
  const int kq = 4 * (lane >> 4);
  int co_l[2], ci_l[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    co_l[b] = cob * 32 + 16 * b + (lane & 15);
    ci_l[b] = cib * 32 + 16 * b + (lane & 15);
  }
  f4v acc[2][2];
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[b][c] = f4v{0.f, 0.f, 0.f, 0.f};

  for (int64_t k0 = p_lo + 16 * wave; k0 < p_hi; k0 += 64) {
    s4v a[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}}, bm[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t p = k0 + kq + i;
      if (p >= p_hi) continue;
      const int pw = static_cast<int>(p % im.w);
      const int64_t t = p / im.w;
      const int ph = static_cast<int>(t % im.h);
      const int pn = static_cast<int>(t / im.h);
      const int hh = ph + dh, ww = pw + dw;
      const bool sv = hh >= 0 && hh < im.h && ww >= 0 && ww < im.w;
      const int64_t src = (static_cast<int64_t>(pn) * im.h + hh) * im.w + ww;
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        if (co_l[b] < cout) a[b][i] = static_cast<short>(dy[p * cout + co_l[b]]);
        if (sv && ci_l[b] < cin) bm[b][i] = static_cast<short>(in_at(x1, c1, x2, c2, src, ci_l[b]));
      }
    }
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) acc[b][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[b], bm[c], acc[b][c], 0, 0, 0);
  }
  // sum the four waves' tiles in a fixed order
#pragma unroll
  for (int b = 0; b < 2; ++b)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int i = 0; i < 4; ++i) red[wave][lane][(b * 2 + c) * 4 + i] = acc[b][c][i];
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int ci = ci_l[c];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int co = cob * 32 + 16 * b + kq + i;
          const int r = (b * 2 + c) * 4 + i;
          const float s = ((red[0][lane][r] + red[1][lane][r]) + red[2][lane][r]) + red[3][lane][r];
          if (co < cout && ci < cin)
            part[((static_cast<size_t>(chunk) * 9 + tap) * cout + co) * cin + ci] = s;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// weight gradient on the matrix cores (channel counts multiples of 32):
// dW[tap][co][ci] = sum_p dy[p][co] x[p + tap][ci], a GEMM per tap with
// M = output channels, N = input channels, K = pixels. A workgroup owns CO_B
// output and 32 input channels and a run of 8 x 16 spatial tiles. Per tile it
// stages dy transposed to [co][pixel] and the 10 x 18 input halo transposed to
// [ci][row][column] in three copies, copy s shifted by s - 1 columns, so that
// every operand -- 8 consecutive pixels of one channel -- is one aligned
// 16-byte LDS read for every tap. Wave w computes the 16 output channels
// w % (CO_B / 16) against CO_B / 32 blocks of 16 input channels, all 9 taps.
// ---------------------------------------------------------------------------
constexpr int kWgPx = kMtRows * kMtCols;                 // pixels per tile (128)
constexpr int kWgDyPitch = kWgPx + 8;                    // bf16 per staged dy channel
constexpr int kWgXPitch = (kMtRows + 2) * kMtCols + 8;   // bf16 per staged input channel of one copy

template <int CO_B>
__global__ __launch_bounds__(256) void conv3x3_wgrad_mfma_kernel(const uint16_t* __restrict__ dy,
                                                                 const uint16_t* __restrict__ x1, int c1,
                                                                 const uint16_t* __restrict__ x2, int c2, Img im,
                                                                 int cout, int tiles_per_chunk,
                                                                 float* __restrict__ part) {
  constexpr int kCoSub = CO_B / 16;  // 16-channel output blocks of the workgroup
  constexpr int kNci = CO_B / 32;    // 16-channel input blocks per wave (4 waves cover kCoSub x 2 blocks)
  constexpr int kG = CO_B / 8;       // 16-byte groups of a dy pixel
  __shared__ __align__(16) uint16_t dyT[CO_B * kWgDyPitch];
  __shared__ __align__(16) uint16_t xT[3 * 32 * kWgXPitch];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, q = lane >> 4;
  const int cin = c1 + c2;
  const int n_cob = cout / CO_B;
  const int cob = blockIdx.x % n_cob, cib = blockIdx.x / n_cob;
  const int co0 = cob * CO_B, ci0 = cib * 32;
  // the 32 input channels lie wholly in x1 or in x2 (c1 % 32 == 0)
  const uint16_t* src = ci0 < c1 ? x1 : x2;
  const int cs = ci0 < c1 ? c1 : c2;
  const int cc0 = ci0 < c1 ? ci0 : ci0 - c1;
  const int co_sub = wave % kCoSub;
  const int ci_sub0 = (wave / kCoSub) * kNci;
  const int tiles_w = (im.w + kMtCols - 1) / kMtCols, tiles_h = (im.h + kMtRows - 1) / kMtRows;
  const int64_t n_tiles = static_cast<int64_t>(im.n) * tiles_h * tiles_w;
  const int64_t t_lo = static_cast<int64_t>(blockIdx.y) * tiles_per_chunk;
  const int64_t t_hi = min(n_tiles, t_lo + tiles_per_chunk);

  f4v acc[9][kNci];
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int j = 0; j < kNci; ++j) acc[tap][j] = f4v{0.f, 0.f, 0.f, 0.f};

  for (int64_t t = t_lo; t < t_hi; ++t) {
    const int tx = static_cast<int>(t % tiles_w);
    const int64_t r = t / tiles_w;
    const int ty = static_cast<int>(r % tiles_h);
    const int img = static_cast<int>(r / tiles_h);
    const int h0 = ty * kMtRows, w0 = tx * kMtCols;
    __syncthreads();  // the previous tile's operand reads are done
    // dy -> dyT[co][pixel]; consecutive lanes take consecutive pixels (consecutive LDS halves)
    for (int k = threadIdx.x; k < kWgPx * kG; k += 256) {
      const int px = k % kWgPx, g = k / kWgPx;
      const int hh = h0 + px / kMtCols, ww = w0 + px % kMtCols;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (hh < im.h && ww < im.w)
        v = *reinterpret_cast<const uint4*>(dy + ((static_cast<int64_t>(img) * im.h + hh) * im.w + ww) * cout + co0 +
                                            8 * g);
      const u16x8 e = __builtin_bit_cast(u16x8, v);
#pragma unroll
      for (int j = 0; j < 8; ++j) dyT[(8 * g + j) * kWgDyPitch + px] = e[j];
    }
    // input halo -> xT[s][ci][row][col] = x[h0 + row - 1][w0 + col + s - 1][ci0 + ci], 0 outside the image
    for (int k = threadIdx.x; k < kMtPatch * 4; k += 256) {
      const int hp = k % kMtPatch, g = k / kMtPatch;
      const int hr = hp / (kMtCols + 2), hc = hp - hr * (kMtCols + 2);
      const int hh = h0 + hr - 1, ww = w0 + hc - 1;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (hh >= 0 && hh < im.h && ww >= 0 && ww < im.w)
        v = *reinterpret_cast<const uint4*>(src + ((static_cast<int64_t>(img) * im.h + hh) * im.w + ww) * cs + cc0 +
                                            8 * g);
      const u16x8 e = __builtin_bit_cast(u16x8, v);
#pragma unroll
      for (int sft = 0; sft < 3; ++sft) {
        const int col = hc - sft;
        if (col < 0 || col >= kMtCols) continue;
#pragma unroll
        for (int j = 0; j < 8; ++j) xT[(sft * 32 + 8 * g + j) * kWgXPitch + hr * kMtCols + col] = e[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int kc = 0; kc < kWgPx / 32; ++kc) {
      // K = 32 pixels: tile rows 2 kc, 2 kc + 1; lane group q holds row 2 kc + q / 2, columns 8 (q % 2) + i
      const bf8v a = __builtin_bit_cast(
          bf8v, *reinterpret_cast<const uint4*>(dyT + (co_sub * 16 + m) * kWgDyPitch + 32 * kc + 8 * q));
      const int tr = 2 * kc + (q >> 1), c0 = 8 * (q & 1);
#pragma unroll
      for (int tap = 0; tap < 9; ++tap) {
        const int dh = tap / 3, dw = tap % 3;  // halo offsets
#pragma unroll
        for (int j = 0; j < kNci; ++j) {
          const bf8v b = __builtin_bit_cast(
              bf8v, *reinterpret_cast<const uint4*>(xT + (dw * 32 + (ci_sub0 + j) * 16 + m) * kWgXPitch +
                                                    (tr + dh) * kMtCols + c0));
          acc[tap][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[tap][j], 0, 0, 0);
        }
      }
    }
  }
  // D[m'][n]: n = lane % 16 (input channel), m' = 4 (lane / 16) + i (output channel)
  const size_t chunk_base = static_cast<size_t>(blockIdx.y) * 9;
#pragma unroll
  for (int tap = 0; tap < 9; ++tap)
#pragma unroll
    for (int j = 0; j < kNci; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = co0 + co_sub * 16 + 4 * q + i;
        const int ci = ci0 + (ci_sub0 + j) * 16 + m;
        part[((chunk_base + tap) * cout + co) * cin + ci] = acc[tap][j][i];
      }
}

// out[i] = sum_c part[c][i]: 32 elements x 8 part groups per workgroup, group g summing parts
// g, g + 8, ... in order, the 8 group sums then added in order (the bits depend only on n_parts)
__global__ __launch_bounds__(256) void sum_parts_wide_kernel(const float* __restrict__ part, int n_parts, int64_t len,
                                                              float* __restrict__ out) {
  __shared__ float red[8][32];
  const int e = threadIdx.x & 31, g = threadIdx.x >> 5;
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 32 + e;
  float s = 0.f;
  if (i < len)
    for (int c = g; c < n_parts; c += 8) s += part[static_cast<size_t>(c) * len + i];
  red[g][e] = s;
  __syncthreads();
  if (g == 0 && i < len) {
    float r = red[0][e];
#pragma unroll
    for (int k = 1; k < 8; ++k) r += red[k][e];
    out[i] = r;
  }
}

// part[chunk][ci] = sum over the chunk's pixels of wgt(p) x[p][ci] (wgt = dout[p], or 1 when
// dout is NULL), and with `extra` part[chunk][c] = sum of wgt(p); with dfeat, dfeat[p][ci] =
// bf16(dout[p] wb[ci]). c = 8 G channels, G a power of two <= 256: thread (g = tid % G,
// pl = tid / G) takes channels [8 g, 8 g + 8) of pixels pl, pl + 256 / G, ...; the 256 / G
// pixel lanes are added in order.
__global__ __launch_bounds__(256) void channel_sum_vec_kernel(const uint16_t* __restrict__ x, int64_t P, int c,
                                                               int64_t chunk_px, const float* __restrict__ dout,
                                                               const float* __restrict__ wb, uint16_t* __restrict__ dfeat,
                                                               int extra, float* __restrict__ part) {
  __shared__ float red[256 * 8 + 256];
  const int G = c / 8, PP = 256 / G;
  const int g = threadIdx.x % G, pl = threadIdx.x / G;
  const int64_t p_lo = static_cast<int64_t>(blockIdx.x) * chunk_px;
  const int64_t p_hi = min(P, p_lo + chunk_px);
  float acc[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float accw = 0.f;
  float wbv[8];
  if (dfeat) {
#pragma unroll
    for (int j = 0; j < 8; ++j) wbv[j] = wb[8 * g + j];
  }
  for (int64_t p = p_lo + pl; p < p_hi; p += PP) {
    const u16x8 e = __builtin_bit_cast(u16x8, *reinterpret_cast<const uint4*>(x + p * c + 8 * g));
    const float wgt = dout ? dout[p] : 1.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] += wgt * bf2f(e[j]);
    accw += wgt;
    if (dfeat) {
      u16x8 d;
#pragma unroll
      for (int j = 0; j < 8; ++j) d[j] = f2bf(dout[p] * wbv[j]);
      *reinterpret_cast<uint4*>(dfeat + p * c + 8 * g) = __builtin_bit_cast(uint4, d);
    }
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) red[pl * c + 8 * g + j] = acc[j];
  if (g == 0) red[256 * 8 + pl] = accw;
  __syncthreads();
  const int n_out = c + (extra ? 1 : 0);
  for (int ci = threadIdx.x; ci < n_out; ci += 256) {
    float s = 0.f;
    if (ci < c) {
      for (int k = 0; k < PP; ++k) s += red[k * c + ci];
    } else {
      for (int k = 0; k < PP; ++k) s += red[256 * 8 + k];
    }
    part[static_cast<size_t>(blockIdx.x) * n_out + ci] = s;
  }
}

// part[chunk][co] = sum over the chunk's pixels of dy[p][co]
__global__ void channel_sum_kernel(const uint16_t* __restrict__ dy, int64_t P, int c, int64_t chunk_px,
                                   float* __restrict__ part) {
  const int chunk = blockIdx.x;
  const int64_t p_lo = static_cast<int64_t>(chunk) * chunk_px;
  const int64_t p_hi = min(P, p_lo + chunk_px);
  for (int ch = threadIdx.x; ch < c; ch += blockDim.x) {
    float s = 0.f;
    for (int64_t p = p_lo; p < p_hi; ++p) s += bf2f(dy[p * c + ch]);
    part[static_cast<size_t>(chunk) * c + ch] = s;
  }
}

// wt_t[tap][ci][co] = wt[8 - tap][co][ci]: the weights of the data-gradient convolution
__global__ void flip_weights_kernel(const uint16_t* __restrict__ wt, int cout, int cin, uint16_t* __restrict__ wt_t) {
  const int64_t total = 9LL * cout * cin;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int tap = static_cast<int>(k / (static_cast<int64_t>(cout) * cin));
    const int rem = static_cast<int>(k - static_cast<int64_t>(tap) * cout * cin);
    const int ci = rem / cout, co = rem - ci * cout;
    wt_t[k] = wt[(static_cast<size_t>(8 - tap) * cout + co) * cin + ci];
  }
}

// relu_mask_kernel / add_bf16_kernel on 8 elements (16 bytes) per step
__global__ void relu_mask_vec_kernel(uint16_t* __restrict__ d, const uint16_t* __restrict__ y, int64_t count8) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < count8;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const u16x8 yv = __builtin_bit_cast(u16x8, reinterpret_cast<const uint4*>(y)[k]);
    u16x8 dv = __builtin_bit_cast(u16x8, reinterpret_cast<const uint4*>(d)[k]);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (!(bf2f(yv[j]) > 0.f)) dv[j] = 0;
    reinterpret_cast<uint4*>(d)[k] = __builtin_bit_cast(uint4, dv);
  }
}

__global__ void add_bf16_vec_kernel(uint16_t* __restrict__ y, const uint16_t* __restrict__ x, int64_t n8) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n8;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const u16x8 xv = __builtin_bit_cast(u16x8, reinterpret_cast<const uint4*>(x)[k]);
    u16x8 yv = __builtin_bit_cast(u16x8, reinterpret_cast<uint4*>(y)[k]);
#pragma unroll
    for (int j = 0; j < 8; ++j) yv[j] = f2bf(bf2f(yv[j]) + bf2f(xv[j]));
    reinterpret_cast<uint4*>(y)[k] = __builtin_bit_cast(uint4, yv);
  }
}

// weight gradient of the 1-channel input layer: part[chunk][tap][co] = sum_p dy[p][co] x[p + tap].
// cout = 8 G (G a power of two <= 256): thread (g = tid % G, pl = tid / G) takes output channels
// [8 g, 8 g + 8) of pixels pl, pl + 256 / G, ... of the chunk, all 9 taps; the pixel lanes are
// then added in order, one tap at a time through LDS.
__global__ __launch_bounds__(256) void conv3x3_wgrad_c1_kernel(const uint16_t* __restrict__ dy,
                                                               const uint16_t* __restrict__ x, Img im, int cout,
                                                               int64_t chunk_px, float* __restrict__ part) {
  __shared__ float red[256 * 8];
  const int G = cout / 8, PP = 256 / G;
  const int g = threadIdx.x % G, pl = threadIdx.x / G;
  const int64_t P = im.pixels();
  const int64_t p_lo = static_cast<int64_t>(blockIdx.x) * chunk_px;
  const int64_t p_hi = min(P, p_lo + chunk_px);
  float acc[9][8];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[t][j] = 0.f;
  for (int64_t p = p_lo + pl; p < p_hi; p += PP) {
    const u16x8 e = __builtin_bit_cast(u16x8, *reinterpret_cast<const uint4*>(dy + p * cout + 8 * g));
    const int pw = static_cast<int>(p % im.w);
    const int64_t r = p / im.w;
    const int ph = static_cast<int>(r % im.h);
    const int64_t base = p - pw - static_cast<int64_t>(ph) * im.w;  // the image's first pixel
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int hh = ph + t / 3 - 1, ww = pw + t % 3 - 1;
      const float xv = (hh >= 0 && hh < im.h && ww >= 0 && ww < im.w)
                           ? bf2f(x[base + static_cast<int64_t>(hh) * im.w + ww])
                           : 0.f;
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[t][j] += bf2f(e[j]) * xv;
    }
  }
  for (int t = 0; t < 9; ++t) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; ++j) red[pl * cout + 8 * g + j] = acc[t][j];
    __syncthreads();
    for (int co = threadIdx.x; co < cout; co += 256) {
      float sum = 0.f;
      for (int k = 0; k < PP; ++k) sum += red[k * cout + co];
      part[(static_cast<size_t>(blockIdx.x) * 9 + t) * cout + co] = sum;
    }
  }
}

__global__ void relu_mask_kernel(uint16_t* __restrict__ d, const uint16_t* __restrict__ y, int64_t count) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < count;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    if (!(bf2f(y[k]) > 0.f)) d[k] = 0;
  }
}

// y[n][h][w][c] = (x[2h][2w] + x[2h][2w+1] + x[2h+1][2w] + x[2h+1][2w+1]) / 4; h, w of the input
// 8 channels per thread (c % 8 == 0): one 16-byte access per pixel and tensor, the pixel
// index decomposed once per 8 channels; the same float32 sums in the same order as the
// per-element kernels below.
__device__ __forceinline__ u16x8 ld8(const uint16_t* p) { return *reinterpret_cast<const u16x8*>(p); }
__device__ __forceinline__ void st8(uint16_t* p, u16x8 v) { *reinterpret_cast<u16x8*>(p) = v; }

// mode 0: avgpool2 forward (x: n x 2ho x 2wo, y: n x ho x wo); 1: upsample2 backward (the same
// 2 x 2 sum without the 1/4)
__global__ void pool_sum_vec_kernel(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int n, int ho, int wo,
                                    int c, int mode) {
  const int cg = c / 8;
  const int64_t total = static_cast<int64_t>(n) * ho * wo * cg;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int g = static_cast<int>(k % cg);
    const int64_t pix = k / cg;
    const int j = static_cast<int>(pix % wo);
    const int64_t t = pix / wo;
    const int i = static_cast<int>(t % ho);
    const int64_t b = t / ho;
    const int64_t r0 = ((b * 2 * ho + 2 * i) * (2 * wo) + 2 * j) * c + 8 * g;
    const int64_t r1 = r0 + static_cast<int64_t>(2 * wo) * c;
    const u16x8 a = ld8(x + r0), bb = ld8(x + r0 + c), cc = ld8(x + r1), d = ld8(x + r1 + c);
    u16x8 o;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float sum = (bf2f(a[e]) + bf2f(bb[e])) + (bf2f(cc[e]) + bf2f(d[e]));
      o[e] = f2bf(mode == 0 ? 0.25f * sum : sum);
    }
    st8(y + pix * c + 8 * g, o);
  }
}

// mode 0: upsample2 forward (x: n x h x w, y: n x 2h x 2w, y[2i + a][2j + b] = x[i][j]);
// 1: avgpool2 backward (the same copy times 1/4)
__global__ void spread_vec_kernel(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int n, int h, int w, int c,
                                  int mode) {
  const int cg = c / 8;
  const int64_t total = static_cast<int64_t>(n) * h * w * cg;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int g = static_cast<int>(k % cg);
    const int64_t pix = k / cg;
    const int j = static_cast<int>(pix % w);
    const int64_t t = pix / w;
    const int i = static_cast<int>(t % h);
    const int64_t b = t / h;
    u16x8 v = ld8(x + pix * c + 8 * g);
    if (mode == 1) {
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = f2bf(0.25f * bf2f(v[e]));
    }
    const int64_t r0 = ((b * 2 * h + 2 * i) * (2 * w) + 2 * j) * c + 8 * g;
    const int64_t r1 = r0 + static_cast<int64_t>(2 * w) * c;
    st8(y + r0, v);
    st8(y + r0 + c, v);
    st8(y + r1, v);
    st8(y + r1 + c, v);
  }
}

__global__ void avgpool2_kernel(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int n, int h, int w, int c) {
  const int ho = h / 2, wo = w / 2;
  const int64_t total = static_cast<int64_t>(n) * ho * wo * c;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int ch = static_cast<int>(k % c);
    int64_t t = k / c;
    const int j = static_cast<int>(t % wo);
    t /= wo;
    const int i = static_cast<int>(t % ho);
    const int64_t b = t / ho;
    const int64_t r0 = ((b * h + 2 * i) * w + 2 * j) * c + ch;
    const int64_t r1 = r0 + static_cast<int64_t>(w) * c;
    const float s = (bf2f(x[r0]) + bf2f(x[r0 + c])) + (bf2f(x[r1]) + bf2f(x[r1 + c]));
    y[k] = f2bf(0.25f * s);
  }
}

// dx[2i + a][2j + b] = dy[i][j] / 4; h, w of the input (dx)
__global__ void avgpool2_bwd_kernel(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx, int n, int h, int w,
                                    int c) {
  const int64_t total = static_cast<int64_t>(n) * h * w * c;
  const int wo = w / 2, ho = h / 2;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int ch = static_cast<int>(k % c);
    int64_t t = k / c;
    const int j = static_cast<int>(t % w);
    t /= w;
    const int i = static_cast<int>(t % h);
    const int64_t b = t / h;
    dx[k] = f2bf(0.25f * bf2f(dy[((b * ho + i / 2) * wo + j / 2) * c + ch]));
  }
}

// y[2i + a][2j + b] = x[i][j]; h, w of the input (x)
__global__ void upsample2_kernel(const uint16_t* __restrict__ x, uint16_t* __restrict__ y, int n, int h, int w, int c) {
  const int ho = 2 * h, wo = 2 * w;
  const int64_t total = static_cast<int64_t>(n) * ho * wo * c;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int ch = static_cast<int>(k % c);
    int64_t t = k / c;
    const int j = static_cast<int>(t % wo);
    t /= wo;
    const int i = static_cast<int>(t % ho);
    const int64_t b = t / ho;
    y[k] = x[((b * h + i / 2) * w + j / 2) * c + ch];
  }
}

// dx[i][j] = sum of dy over the 2 x 2 block it was copied to; h, w of dx
__global__ void upsample2_bwd_kernel(const uint16_t* __restrict__ dy, uint16_t* __restrict__ dx, int n, int h, int w,
                                     int c) {
  const int64_t total = static_cast<int64_t>(n) * h * w * c;
  const int wo = 2 * w;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int ch = static_cast<int>(k % c);
    int64_t t = k / c;
    const int j = static_cast<int>(t % w);
    t /= w;
    const int i = static_cast<int>(t % h);
    const int64_t b = t / h;
    const int64_t r0 = ((b * 2 * h + 2 * i) * wo + 2 * j) * c + ch;
    const int64_t r1 = r0 + static_cast<int64_t>(wo) * c;
    const float s = (bf2f(dy[r0]) + bf2f(dy[r0 + c])) + (bf2f(dy[r1]) + bf2f(dy[r1 + c]));
    dx[k] = f2bf(s);
  }
}

// out[p] = xin[p] + wb[c] + sum_ci wb[ci] * feat[p][ci]   (1 x 1 head, residual on the input; fp32)
__global__ void head_fwd_kernel(const uint16_t* __restrict__ feat, int c, const float* __restrict__ wb,
                                const float* __restrict__ xin, float* __restrict__ out, int64_t P) {
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < P;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    float s = wb[c];
    for (int ci = 0; ci < c; ++ci) s = fmaf(wb[ci], bf2f(feat[p * c + ci]), s);
    out[p] = xin[p] + s;
  }
}

// head_fwd_kernel with 16-byte loads of each pixel's features (c % 8 == 0): the same sums in the same order
__global__ void head_fwd_vec_kernel(const uint16_t* __restrict__ feat, int c, const float* __restrict__ wb,
                                    const float* __restrict__ xin, float* __restrict__ out, int64_t P) {
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < P;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    float s = wb[c];
    for (int c8 = 0; c8 < c; c8 += 8) {
      const u16x8 e = __builtin_bit_cast(u16x8, *reinterpret_cast<const uint4*>(feat + p * c + c8));
#pragma unroll
      for (int j = 0; j < 8; ++j) s = fmaf(wb[c8 + j], bf2f(e[j]), s);
    }
    out[p] = xin[p] + s;
  }
}

// dfeat[p][ci] = dout[p] * wb[ci]; part[chunk][ci] = sum_p dout[p] feat[p][ci], part[chunk][c] = sum_p dout[p]
__global__ void head_bwd_kernel(const float* __restrict__ dout, const uint16_t* __restrict__ feat, int c,
                                const float* __restrict__ wb, uint16_t* __restrict__ dfeat, int64_t P,
                                int64_t chunk_px, float* __restrict__ part) {
  const int chunk = blockIdx.x;
  const int64_t p_lo = static_cast<int64_t>(chunk) * chunk_px;
  const int64_t p_hi = min(P, p_lo + chunk_px);
  for (int64_t p = p_lo + threadIdx.x; p < p_hi; p += blockDim.x) {
    for (int ci = 0; ci < c; ++ci) dfeat[p * c + ci] = f2bf(dout[p] * wb[ci]);
  }
  for (int ci = threadIdx.x; ci <= c; ci += blockDim.x) {
    float s = 0.f;
    for (int64_t p = p_lo; p < p_hi; ++p) s += dout[p] * (ci < c ? bf2f(feat[p * c + ci]) : 1.f);
    part[static_cast<size_t>(chunk) * (c + 1) + ci] = s;
  }
}

// dout[p] = 2 (out[p] - tgt[p]) / P; part[block] = sum of squared errors of the block's pixels
__global__ void mse_kernel(const float* __restrict__ out, const float* __restrict__ tgt, int64_t P,
                           float* __restrict__ dout, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; p < P;
       p += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float e = out[p] - tgt[p];
    dout[p] = 2.f * e / static_cast<float>(P);
    s += static_cast<double>(e) * e;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out = scale * sum part[0 .. n): thread t sums part[t], part[t + 256], ... in order, then a
// fixed pairwise tree over the 256 thread sums (one workgroup of 256 threads)
__global__ __launch_bounds__(256) void sum_double_kernel(const double* __restrict__ part, int n, double scale,
                                                         double* __restrict__ out) {
  __shared__ double red[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * scale;
}

// Adam on fp32 master weights; optional bf16 copy of the result
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                            float c2, uint16_t* __restrict__ w_bf16) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float wk = w[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    w[k] = wk;
    if (w_bf16) w_bf16[k] = f2bf(wk);
  }
}

__global__ void add_bf16_kernel(uint16_t* __restrict__ y, const uint16_t* __restrict__ x, int64_t n) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[k] = f2bf(bf2f(y[k]) + bf2f(x[k]));
}

__global__ void to_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[k] = f2bf(x[k]);
}

__global__ void to_f32_kernel(const uint16_t* __restrict__ x, float* __restrict__ y, int64_t n) {
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x)
    y[k] = bf2f(x[k]);
}

bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

// channel counts channel_sum_vec_kernel takes: 8 G with G a power of two <= 256
bool vec_channels(int c) {
  const int g = c / 8;
  return c % 8 == 0 && g >= 1 && g <= 256 && (g & (g - 1)) == 0;
}

dim3 grid_for(int64_t n, int block = 256) {
  return dim3(static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>((n + block - 1) / block, 1), 65536)));
}

// chunks of pixels for a split reduction whose partial rows hold `row` floats,
// within `part_elems` floats of scratch
int64_t chunk_pixels(int64_t P, int64_t row, int64_t part_elems, int64_t min_px, int& n_chunks) {
  int64_t chunks = std::max<int64_t>(1, std::min<int64_t>(part_elems / std::max<int64_t>(row, 1), (P + min_px - 1) / min_px));
  chunks = std::min<int64_t>(chunks, 65535);
  int64_t px = (P + chunks - 1) / chunks;
  n_chunks = static_cast<int>((P + px - 1) / px);
  return px;
}

}  // namespace
}  // namespace chip

using namespace chip;

extern "C" {

int chip_nn_conv3x3(const uint16_t* x1, int c1, const uint16_t* x2, int c2, const uint16_t* wt, const float* bias,
                    uint16_t* y1, int o1, uint16_t* y2, int n, int h, int w, int cout, int relu, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x1 && wt && y1, "NULL argument");
  CHIP_CHECK_ARG(c1 >= 1 && c2 >= 0 && (c2 == 0 || x2), "bad input channels");
  CHIP_CHECK_ARG(cout >= 1 && o1 >= 1 && o1 <= cout && (o1 == cout || y2), "bad output channels");
  CHIP_CHECK_ARG(n >= 1 && h >= 1 && w >= 1, "bad image shape");
  CHIP_CHECK_ARG((c1 % 4 != 0 || reinterpret_cast<uintptr_t>(x1) % 8 == 0) &&
                     (c2 % 4 != 0 || reinterpret_cast<uintptr_t>(x2) % 8 == 0) &&
                     reinterpret_cast<uintptr_t>(wt) % 8 == 0,
                 "x1 / x2 / wt must be 8-byte aligned");
  const int64_t P = static_cast<int64_t>(n) * h * w;
  const bool al16 = reinterpret_cast<uintptr_t>(y1) % 16 == 0 && reinterpret_cast<uintptr_t>(y2) % 16 == 0;
  const bool mfma32 = c1 % 32 == 0 && c2 % 32 == 0 && cout % 32 == 0 && (o1 == cout || o1 % 16 == 0) &&
                      reinterpret_cast<uintptr_t>(x1) % 16 == 0 && reinterpret_cast<uintptr_t>(x2) % 16 == 0 &&
                      reinterpret_cast<uintptr_t>(wt) % 16 == 0 && al16;
  if (c1 == 1 && c2 == 0 && o1 == cout && cout % 8 == 0 && cout <= 1024 && al16) {
    dim3 grid(static_cast<unsigned>((P + 255) / 256));
    CHIP_CHECK_ARG(P <= 2147483647LL * 256, "too many pixels");
    hipLaunchKernelGGL(conv3x3_c1_kernel, grid, dim3(256), sizeof(float) * 10 * cout, as_stream(stream), x1, wt, bias,
                       y1, Img{n, h, w}, cout, relu);
    CHIP_LAUNCH_CHECK();
    return CHIP_OK;
  }
  if (mfma32) {
    const bool co64 = cout % 64 == 0;
    const int n_cob = cout / (co64 ? 64 : 32);
    // 16-row tiles (4 row blocks per wave) where they still give 1024 workgroups, else 8-row tiles
    const int64_t tiles16 = static_cast<int64_t>(n) * ((h + 15) / 16) * ((w + kMtCols - 1) / kMtCols);
    const bool rb4 = tiles16 * n_cob >= 1024;
    const int64_t tiles = rb4 ? tiles16 : static_cast<int64_t>(n) * ((h + 7) / 8) * ((w + kMtCols - 1) / kMtCols);
    CHIP_CHECK_ARG(tiles <= 2147483647, "too many pixels");
    dim3 grid(static_cast<unsigned>(tiles), n_cob);
    hipStream_t st = as_stream(stream);
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, dim3(256), 0, st, x1, c1, x2, c2, wt, bias, y1, o1, y2, Img{n, h, w}, cout, relu);
    };
    if (co64 && rb4) go(conv3x3_mfma_kernel<64, 4>);
    else if (co64) go(conv3x3_mfma_kernel<64, 2>);
    else if (rb4) go(conv3x3_mfma_kernel<32, 4>);
    else go(conv3x3_mfma_kernel<32, 2>);
    CHIP_LAUNCH_CHECK();
    return CHIP_OK;
  }
  dim3 grid(static_cast<unsigned>((P + kPxBlock - 1) / kPxBlock), (cout + kCoBlock - 1) / kCoBlock);
  CHIP_CHECK_ARG(grid.x <= 2147483647u, "too many pixels");
  hipLaunchKernelGGL(conv3x3_kernel, grid, dim3(256), 0, as_stream(stream), x1, c1, x2, c2, wt, bias, y1, o1, y2,
                     Img{n, h, w}, cout, relu);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_conv3x3_wgrad(const uint16_t* dy, const uint16_t* x1, int c1, const uint16_t* x2, int c2, int n, int h,
                          int w, int cout, float* part, int64_t part_elems, float* dw, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(dy && x1 && part && dw, "NULL argument");
  CHIP_CHECK_ARG(c1 >= 1 && c2 >= 0 && (c2 == 0 || x2) && cout >= 1, "bad channels");
  const int cin = c1 + c2;
  const int64_t P = static_cast<int64_t>(n) * h * w;
  const int64_t row = 9LL * cout * cin;
  CHIP_CHECK_ARG(part_elems >= row, "partial-sum scratch smaller than one weight set");
  hipStream_t st = as_stream(stream);
  const bool mfma = cout % 32 == 0 && c1 % 32 == 0 && c2 % 32 == 0 && aligned16(dy) && aligned16(x1) &&
                    (c2 == 0 || aligned16(x2));
  int n_chunks = 1;
  if (mfma) {
    const int co_b = cout % 64 == 0 ? 64 : 32;
    const int n_blocks = (cout / co_b) * (cin / 32);
    const int64_t n_tiles = static_cast<int64_t>(n) * ((h + kMtRows - 1) / kMtRows) * ((w + kMtCols - 1) / kMtCols);
    // about 1024 workgroups, at most 512 partial sets, within the scratch
    int64_t chunks = std::max<int64_t>(1, (1024 + n_blocks - 1) / n_blocks);
    chunks = std::min<int64_t>({chunks, 512, n_tiles, std::max<int64_t>(1, part_elems / row), 65535});
    const int64_t tpc = (n_tiles + chunks - 1) / chunks;
    n_chunks = static_cast<int>((n_tiles + tpc - 1) / tpc);
    const dim3 grid(n_blocks, n_chunks);
    if (co_b == 64)
      hipLaunchKernelGGL(conv3x3_wgrad_mfma_kernel<64>, grid, dim3(256), 0, st, dy, x1, c1, x2, c2, Img{n, h, w}, cout,
                         static_cast<int>(tpc), part);
    else
      hipLaunchKernelGGL(conv3x3_wgrad_mfma_kernel<32>, grid, dim3(256), 0, st, dy, x1, c1, x2, c2, Img{n, h, w}, cout,
                         static_cast<int>(tpc), part);
  } else if (c1 == 1 && c2 == 0 && vec_channels(cout) && aligned16(dy)) {
    const int64_t px = chunk_pixels(P, row, part_elems, 1024, n_chunks);
    hipLaunchKernelGGL(conv3x3_wgrad_c1_kernel, dim3(n_chunks), dim3(256), 0, st, dy, x1, Img{n, h, w}, cout, px,
                       part);
  } else {
    const int64_t px = chunk_pixels(P, row, part_elems, 256, n_chunks);
    const int n_tiles = 9 * ((cout + 31) / 32) * ((cin + 31) / 32);
    hipLaunchKernelGGL(conv3x3_wgrad_kernel, dim3(n_tiles, n_chunks), dim3(256), 0, st, dy, x1, c1, x2, c2,
                       Img{n, h, w}, cout, px, part);
  }
  CHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_parts_wide_kernel, dim3(static_cast<unsigned>((row + 31) / 32)), dim3(256), 0, st, part,
                     n_chunks, row, dw);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_bias_grad(const uint16_t* dy, int64_t n_pix, int c, float* part, int64_t part_elems, float* db,
                      void* stream) {
  clear_error();
  CHIP_CHECK_ARG(dy && part && db && c >= 1 && n_pix >= 1, "bad argument");
  CHIP_CHECK_ARG(part_elems >= c, "partial-sum scratch too small");
  int n_chunks = 1;
  hipStream_t st = as_stream(stream);
  if (vec_channels(c) && aligned16(dy)) {
    const int64_t px = chunk_pixels(n_pix, c, part_elems, 1024, n_chunks);
    hipLaunchKernelGGL(channel_sum_vec_kernel, dim3(n_chunks), dim3(256), 0, st, dy, n_pix, c, px, nullptr, nullptr,
                       nullptr, 0, part);
  } else {
    const int64_t px = chunk_pixels(n_pix, c, part_elems, 1024, n_chunks);
    hipLaunchKernelGGL(channel_sum_kernel, dim3(n_chunks), dim3(256), 0, st, dy, n_pix, c, px, part);
  }
  CHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_parts_wide_kernel, dim3(static_cast<unsigned>((c + 31) / 32)), dim3(256), 0, st, part,
                     n_chunks, static_cast<int64_t>(c), db);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_flip_weights(const uint16_t* wt, int cout, int cin, uint16_t* wt_t, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(wt && wt_t && cout >= 1 && cin >= 1, "bad argument");
  hipLaunchKernelGGL(flip_weights_kernel, grid_for(9LL * cout * cin), dim3(256), 0, as_stream(stream), wt, cout, cin,
                     wt_t);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_relu_mask(uint16_t* d, const uint16_t* y, int64_t count, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(d && y && count >= 0, "bad argument");
  if (count % 8 == 0 && aligned16(d) && aligned16(y))
    hipLaunchKernelGGL(relu_mask_vec_kernel, grid_for(count / 8), dim3(256), 0, as_stream(stream), d, y, count / 8);
  else
    hipLaunchKernelGGL(relu_mask_kernel, grid_for(count), dim3(256), 0, as_stream(stream), d, y, count);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_avgpool2(const uint16_t* x, uint16_t* y, int n, int h, int w, int c, int backward, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && y && n >= 1 && c >= 1 && h >= 2 && w >= 2 && h % 2 == 0 && w % 2 == 0, "bad argument");
  const int64_t count = static_cast<int64_t>(n) * h * w * c;
  const bool vec = c % 8 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0;
  if (vec) {
    if (backward)  // x = dy of the pooled shape (h / 2 x w / 2), y = dx of h x w
      hipLaunchKernelGGL(spread_vec_kernel, grid_for(count / 32), dim3(256), 0, as_stream(stream), x, y, n, h / 2,
                         w / 2, c, 1);
    else
      hipLaunchKernelGGL(pool_sum_vec_kernel, grid_for(count / 32), dim3(256), 0, as_stream(stream), x, y, n, h / 2,
                         w / 2, c, 0);
  } else if (backward) {
    hipLaunchKernelGGL(avgpool2_bwd_kernel, grid_for(count), dim3(256), 0, as_stream(stream), x, y, n, h, w, c);
  } else {
    hipLaunchKernelGGL(avgpool2_kernel, grid_for(count / 4), dim3(256), 0, as_stream(stream), x, y, n, h, w, c);
  }
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_upsample2(const uint16_t* x, uint16_t* y, int n, int h, int w, int c, int backward, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && y && n >= 1 && c >= 1 && h >= 1 && w >= 1, "bad argument");
  const int64_t count = static_cast<int64_t>(n) * h * w * c;
  const bool vec = c % 8 == 0 && reinterpret_cast<uintptr_t>(x) % 16 == 0 && reinterpret_cast<uintptr_t>(y) % 16 == 0;
  if (vec) {
    if (backward)  // x = dy of 2h x 2w, y = dx of h x w
      hipLaunchKernelGGL(pool_sum_vec_kernel, grid_for(count / 8), dim3(256), 0, as_stream(stream), x, y, n, h, w, c,
                         1);
    else
      hipLaunchKernelGGL(spread_vec_kernel, grid_for(count / 8), dim3(256), 0, as_stream(stream), x, y, n, h, w, c, 0);
  } else if (backward) {
    hipLaunchKernelGGL(upsample2_bwd_kernel, grid_for(count), dim3(256), 0, as_stream(stream), x, y, n, h, w, c);
  } else {
    hipLaunchKernelGGL(upsample2_kernel, grid_for(4 * count), dim3(256), 0, as_stream(stream), x, y, n, h, w, c);
  }
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_head_fwd(const uint16_t* feat, int c, const float* wb, const float* xin, float* out, int64_t n_pix,
                     void* stream) {
  clear_error();
  CHIP_CHECK_ARG(feat && wb && xin && out && c >= 1 && n_pix >= 1, "bad argument");
  if (c % 8 == 0 && aligned16(feat))
    hipLaunchKernelGGL(head_fwd_vec_kernel, grid_for(n_pix), dim3(256), 0, as_stream(stream), feat, c, wb, xin, out,
                       n_pix);
  else
    hipLaunchKernelGGL(head_fwd_kernel, grid_for(n_pix), dim3(256), 0, as_stream(stream), feat, c, wb, xin, out,
                       n_pix);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_head_bwd(const float* dout, const uint16_t* feat, int c, const float* wb, uint16_t* dfeat,
                     float* part, int64_t part_elems, float* dwb, int64_t n_pix, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(dout && feat && wb && dfeat && part && dwb && c >= 1 && n_pix >= 1, "bad argument");
  CHIP_CHECK_ARG(part_elems >= c + 1, "partial-sum scratch too small");
  int n_chunks = 1;
  const bool vec = vec_channels(c) && aligned16(feat) && aligned16(dfeat);
  const int64_t px = chunk_pixels(n_pix, c + 1, part_elems, vec ? 1024 : 4096, n_chunks);
  hipStream_t st = as_stream(stream);
  if (vec)
    hipLaunchKernelGGL(channel_sum_vec_kernel, dim3(n_chunks), dim3(256), 0, st, feat, n_pix, c, px, dout, wb, dfeat,
                       1, part);
  else
    hipLaunchKernelGGL(head_bwd_kernel, dim3(n_chunks), dim3(256), 0, st, dout, feat, c, wb, dfeat, n_pix, px, part);
  CHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_parts_wide_kernel, dim3(static_cast<unsigned>((c + 1 + 31) / 32)), dim3(256), 0, st, part,
                     n_chunks, static_cast<int64_t>(c + 1), dwb);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_mse(const float* out, const float* tgt, int64_t n_pix, float* dout, double* part, int part_elems,
                double* loss, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(out && tgt && dout && part && loss && n_pix >= 1 && part_elems >= 1, "bad argument");
  const int blocks = static_cast<int>(std::min<int64_t>({static_cast<int64_t>(part_elems), (n_pix + 255) / 256, 1024}));
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(mse_kernel, dim3(blocks), dim3(256), 0, st, out, tgt, n_pix, dout, part);
  CHIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(sum_double_kernel, dim3(1), dim3(256), 0, st, part, blocks, 1.0 / static_cast<double>(n_pix),
                     loss);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
                 float eps, int step, uint16_t* w_bf16, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(w && g && m && v && n >= 1 && step >= 1, "bad argument");
  const float c1 = 1.f - powf(beta1, static_cast<float>(step));
  const float c2 = 1.f - powf(beta2, static_cast<float>(step));
  hipLaunchKernelGGL(adam_kernel, grid_for(n), dim3(256), 0, as_stream(stream), w, g, m, v, n, lr, beta1, beta2, eps,
                     c1, c2, w_bf16);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_add_bf16(uint16_t* y, const uint16_t* x, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && y && n >= 0, "bad argument");
  if (n % 8 == 0 && aligned16(x) && aligned16(y))
    hipLaunchKernelGGL(add_bf16_vec_kernel, grid_for(n / 8), dim3(256), 0, as_stream(stream), y, x, n / 8);
  else
    hipLaunchKernelGGL(add_bf16_kernel, grid_for(n), dim3(256), 0, as_stream(stream), y, x, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_to_bf16(const float* x, uint16_t* y, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && y && n >= 0, "bad argument");
  hipLaunchKernelGGL(to_bf16_kernel, grid_for(n), dim3(256), 0, as_stream(stream), x, y, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_nn_to_f32(const uint16_t* x, float* y, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && y && n >= 0, "bad argument");
  hipLaunchKernelGGL(to_f32_kernel, grid_for(n), dim3(256), 0, as_stream(stream), x, y, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

}  // extern "C"
