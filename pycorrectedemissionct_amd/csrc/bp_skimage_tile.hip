// Whole-sinogram skimage back-projection with staged line windows (gfx950).
//
// Same arithmetic as bp_skimage_pixel (projector.hip): per pixel and angle
// t = (j - k) cos - (i - k) sin, np.interp of the line at t with 0 outside
// [-k, n - 1 - k] (membership decided in float64 near the ends), summed over
// angles then detectors in that order, times w[a, d] M[a, d][i, j] for the
// symmetric form (_projector_backends.py:335-359, projectors.py:552-568);
// iradon's x 2 / pi is folded into the plan's weights by the caller as before.
//
// A workgroup owns a 16 x 16 pixel tile and walks the angles in blocks of
// 32: the bins the tile can touch (<= 15 (|c| + |s|) + 4 per angle and
// detector) are staged in LDS once per block instead of being gathered
// through the caches by every pixel, and the symmetric form's map values are
// loaded 8 angles at a time, the next 8 while the current are summed.
#include <algorithm>

#include "common.hpp"

namespace chip {
namespace {

constexpr int kSkTile = 16;
constexpr int kSkAngBlk = 32;
constexpr int kSkGrp = 8;
constexpr int kSkWin = 28;  // >= 15 (|c| + |s|) + 4, with one bin of slack each side
// symmetric single slices: 32 x 8 tiles, so that each wave's map reads are whole 128-byte
// rows (16-wide rows read half lines, which the caches then fetch twice); the window spans
// 31 |c| + 7 |s| + 4 <= 36 bins (indices 1..34 checked on every tile and angle of N = 33,
// 64, 70, 130 and 512 at 0.5 degree steps)
constexpr int kSymTileW = 32, kSymTileH = 8, kSymWin = 36;

// One TW x TH tile (tx, ty) of one slice: sb = the slice's sinogram [n_det][n_angles][n],
// mb = its maps or NULL, out = the slice's image; W window bins per angle and detector.
template <int SYM, int D, int TW, int TH, int W>
__device__ __forceinline__ void bp_skimage_tile_block(const float* __restrict__ sb, const float* __restrict__ mb,
                                                      float* __restrict__ out, const float4* __restrict__ geom,
                                                      const float* __restrict__ w, int n, int n_angles, int n_det,
                                                      int det0, int angle0, int n_ang, int maps_a0, int accumulate,
                                                      int tx, int ty, float (*win)[D][W], int* wlo_s) {
  static_assert(TW * TH == 256, "one pixel per thread");
  const int tid = threadIdx.x;
  const int j = tx * TW + (tid % TW);
  const int i = ty * TH + (tid / TW);
  const int k = n / 2;
  const size_t nn = static_cast<size_t>(n) * n;
  const bool inside = i < n && j < n;
  const size_t pix = static_cast<size_t>(min(i, n - 1)) * n + min(j, n - 1);
  const float jc = static_cast<float>(j - k);
  const float ic = static_cast<float>(i - k);
  const float tlo = static_cast<float>(-k);
  const float thi = static_cast<float>(n - 1 - k);
  // tile corners, centred
  const float jc0 = static_cast<float>(tx * TW - k), jc1 = jc0 + (TW - 1);
  const float ic0 = static_cast<float>(ty * TH - k), ic1 = ic0 + (TH - 1);

  float acc = 0.f;
  for (int a0 = angle0; a0 < angle0 + n_ang; a0 += kSkAngBlk) {
    const int na = min(kSkAngBlk, angle0 + n_ang - a0);
    __syncthreads();  // the previous block's windows are no longer read
    if (tid < na) {
      const float4 g = geom[a0 + tid];
      const float t00 = fmaf(jc0, g.x, -(ic0 * g.y)), t01 = fmaf(jc1, g.x, -(ic0 * g.y));
      const float t10 = fmaf(jc0, g.x, -(ic1 * g.y)), t11 = fmaf(jc1, g.x, -(ic1 * g.y));
      const float tmin = fminf(fminf(t00, t01), fminf(t10, t11));
      wlo_s[tid] = static_cast<int>(floorf(tmin)) + k - 1;
    }
    __syncthreads();
    for (int q = tid; q < na * D * W; q += 256) {
      const int al = q / (D * W);
      const int rem = q - al * (D * W);
      const int d = rem / W;
      const int e = rem - d * W;
      const int bin = wlo_s[al] + e;
      win[al][d][e] =
          (bin >= 0 && bin < n) ? sb[(static_cast<size_t>(det0 + d) * n_angles + a0 + al) * n + bin] : 0.f;
    }
    __syncthreads();
    if (inside) {
      float mcur[kSkGrp][D], mnxt[kSkGrp][D];
      auto load_maps = [&](float (*m)[D], int g0) {
#pragma unroll
        for (int q = 0; q < kSkGrp; ++q)
#pragma unroll
          for (int d = 0; d < D; ++d)
            m[q][d] = (mb != nullptr && g0 + q < na)
                          ? mb[(static_cast<size_t>(a0 + g0 + q - maps_a0) * n_det + det0 + d) * nn + pix]
                          : 1.f;
      };
      if (SYM) load_maps(mcur, 0);
      for (int g0 = 0; g0 < na; g0 += kSkGrp) {
        if (SYM && g0 + kSkGrp < na) load_maps(mnxt, g0 + kSkGrp);
#pragma unroll
        for (int q = 0; q < kSkGrp; ++q) {
          const int al = g0 + q;
          if (al >= na) break;
          const int a = a0 + al;
          const float4 g = geom[a];
          float t = fmaf(jc, g.x, -(ic * g.y));
          // np.interp is discontinuous at the ends of the line (0 outside): decide
          // membership with the reference's double-precision t whenever the
          // float32 t is within its error bound of an end
          bool member;
          if (fabsf(t - tlo) < 1e-3f || fabsf(t - thi) < 1e-3f) {
            const double td = fma(static_cast<double>(jc), static_cast<double>(g.x),
                                  -(static_cast<double>(ic) * static_cast<double>(g.y)));
            member = td >= static_cast<double>(tlo) && td <= static_cast<double>(thi);
            t = fminf(fmaxf(t, tlo), thi);
          } else {
            member = t >= tlo && t <= thi;
          }
          if (!member) continue;
          const float tf = floorf(t);
          const float fr = t - tf;
          const int e = static_cast<int>(tf) + k - wlo_s[al];  // bins past n - 1 read the zero padding
#pragma unroll
          for (int d = 0; d < D; ++d) {
            const float s0 = win[al][d][e];
            const float s1 = win[al][d][e + 1];
            float val = fmaf(fr, s1 - s0, s0);
            if (SYM) {
              float f = w[static_cast<size_t>(a) * n_det + det0 + d];
              if (mb != nullptr) f *= mcur[q][d];
              val *= f;
            }
            acc += val;
          }
        }
        if (SYM) {
#pragma unroll
          for (int q = 0; q < kSkGrp; ++q)
#pragma unroll
            for (int d = 0; d < D; ++d) mcur[q][d] = mnxt[q][d];
        }
      }
    }
  }
  if (inside) {
    float* o = out + pix;
    if (accumulate > 1) unsafeAtomicAdd(o, acc);  // angle halves of a zeroed slice (see launch_tiles)
    else *o = accumulate ? *o + acc : acc;
  }
}

// split = 1: one slice, blockIdx.z = half of the angle range; both halves add their
// sums to the zeroed slice (two addends onto 0 give the same bits in either order)
template <int SYM, int D, int TW, int TH, int W>
__global__ __launch_bounds__(256) void bp_skimage_tile_kernel(const float* __restrict__ sino,
                                                              const float* __restrict__ maps,
                                                              int64_t maps_slice_stride, float* __restrict__ vol,
                                                              const float4* __restrict__ geom,
                                                              const float* __restrict__ w, int n, int n_angles,
                                                              int n_det, int det0, int angle0, int n_ang,
                                                              int maps_a0, int accumulate, int split) {
  __shared__ float win[kSkAngBlk][D][W];
  __shared__ int wlo_s[kSkAngBlk];
  int b = blockIdx.z;
  if (split) {
    const int first = (n_ang + 1) / 2;
    b = 0;
    if (blockIdx.z == 0) {
      n_ang = first;
    } else {
      angle0 += first;
      n_ang -= first;
    }
    accumulate = 2;
  }
  const float* sb = sino + static_cast<size_t>(b) * n_det * n_angles * n;
  const float* mb = (SYM && maps) ? maps + static_cast<size_t>(b) * maps_slice_stride : nullptr;
  bp_skimage_tile_block<SYM, D, TW, TH, W>(sb, mb, vol + static_cast<size_t>(b) * n * n, geom, w, n, n_angles, n_det, det0,
                                angle0, n_ang, maps_a0, accumulate, static_cast<int>(blockIdx.x),
                                static_cast<int>(blockIdx.y), win, wlo_s);
}

// Persistent form over a batch of slices of different sizes: item k is the
// 16 x 16 tile of slice s found in the tile prefix sums.
template <int SYM, int D>
__global__ __launch_bounds__(256) void bp_skimage_tile_batch_kernel(const BatchSlice* __restrict__ slices,
                                                                    int n_slices,
                                                                    const int64_t* __restrict__ item_start,
                                                                    const float* __restrict__ sino,
                                                                    const float* __restrict__ maps,
                                                                    float* __restrict__ vol,
                                                                    const float4* __restrict__ geom,
                                                                    const float* __restrict__ w, int n_angles,
                                                                    int n_det, int det0, int accumulate) {
  __shared__ float win[kSkAngBlk][D][kSkWin];
  __shared__ int wlo_s[kSkAngBlk];
  const int64_t total = item_start[n_slices];
  for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
    const int s = find_slice(item_start, n_slices, it);
    const BatchSlice sl = slices[s];
    const int local = static_cast<int>(it - item_start[s]);
    const int ntx = (sl.n + kSkTile - 1) / kSkTile;
    const int ty = local / ntx, tx = local - ty * ntx;
    bp_skimage_tile_block<SYM, D, kSkTile, kSkTile, kSkWin>(sino + sl.sino_off, (SYM && maps) ? maps + sl.maps_off : nullptr,
                                  vol + sl.vol_off, geom + static_cast<size_t>(sl.size_idx) * n_angles, w, sl.n,
                                  n_angles, n_det, det0, 0, n_angles, 0, accumulate, tx, ty, win, wlo_s);
  }
}

__global__ void zero_kernel(float* __restrict__ v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.f;
}

template <int SYM, int D>
void launch_tiles(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol, int batch,
                  int det0, int angle0, int n_ang, int maps_a0, int accumulate, hipStream_t st) {
  // one slice fills only n^2 / 64 waves (4 per SIMD at 512^2): with a fresh output and every
  // detector in this pass, the non-symmetric form (VALU-bound) splits the angle range over
  // twice the workgroups; the symmetric form is bound by its map loads and gains nothing
  const int split =
      (!SYM && batch == 1 && !accumulate && det0 == 0 && p.n_det <= 2 && n_ang >= 2 * kSkAngBlk) ? 1 : 0;
  if (split) {
    const int64_t nn = static_cast<int64_t>(p.n) * p.n;
    hipLaunchKernelGGL(zero_kernel, dim3(static_cast<unsigned>((nn + 255) / 256)), dim3(256), 0, st, vol, nn);
  }
  constexpr int TW = SYM ? kSymTileW : kSkTile, TH = SYM ? kSymTileH : kSkTile, W = SYM ? kSymWin : kSkWin;
  dim3 grid((p.n + TW - 1) / TW, (p.n + TH - 1) / TH, split ? 2 : batch);
  hipLaunchKernelGGL((bp_skimage_tile_kernel<SYM, D, TW, TH, W>), grid, dim3(256), 0, st, sino, maps, stride, vol,
                     p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, angle0, n_ang, maps_a0, accumulate, split);
}

}  // namespace

// Detectors in passes of two, the passes after the first accumulating; with
// one or two detectors the sums run in the same order as bp_skimage_kernel,
// except that a single non-symmetric slice with 64 or more angles adds its two
// angle halves.
void launch_bp_skimage_tiles(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol,
                             int batch, int sym, int angle0, int n_ang, int maps_a0, int accumulate,
                             hipStream_t st) {
  for (int det0 = 0; det0 < p.n_det; det0 += 2) {
    const bool two = p.n_det - det0 >= 2;
    const int acc = (det0 > 0) ? 1 : accumulate;
    if (sym) {
      if (two) launch_tiles<1, 2>(p, sino, maps, stride, vol, batch, det0, angle0, n_ang, maps_a0, acc, st);
      else launch_tiles<1, 1>(p, sino, maps, stride, vol, batch, det0, angle0, n_ang, maps_a0, acc, st);
    } else {
      if (two) launch_tiles<0, 2>(p, sino, nullptr, 0, vol, batch, det0, angle0, n_ang, 0, acc, st);
      else launch_tiles<0, 1>(p, sino, nullptr, 0, vol, batch, det0, angle0, n_ang, 0, acc, st);
    }
  }
}

// Batched form: one persistent launch per pass of two detectors over the 16 x 16
// tiles of every slice.
void launch_bp_skimage_tiles_batch(const BatchPlan& bp, const float* sino, const float* maps, float* vol, int sym,
                                   hipStream_t st) {
  const int64_t total = bp.h_item_start_bps.back();
  // bp_wg_per_cu = 0: one workgroup per tile (a flat grid), else a persistent grid
  const int64_t want = bp.bp_wg_per_cu > 0 ? static_cast<int64_t>(bp.bp_wg_per_cu) * bp.n_cu : total;
  const int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(total, want)));
  for (int det0 = 0; det0 < bp.n_det; det0 += 2) {
    const bool two = bp.n_det - det0 >= 2;
    const int acc = det0 > 0 ? 1 : 0;
    const float* m = sym ? maps : nullptr;
#define CHIP_SK_BATCH(S, DD)                                                                                      \
  hipLaunchKernelGGL((bp_skimage_tile_batch_kernel<S, DD>), dim3(grid), dim3(256), 0, st, bp.d_slices, bp.n_slices, \
                     bp.d_item_start_bps, sino, m, vol, bp.d_geom, bp.d_w, bp.n_angles, bp.n_det, det0, acc)
    if (sym) {
      if (two) CHIP_SK_BATCH(1, 2);
      else CHIP_SK_BATCH(1, 1);
    } else {
      if (two) CHIP_SK_BATCH(0, 2);
      else CHIP_SK_BATCH(0, 1);
    }
#undef CHIP_SK_BATCH
  }
}

}  // namespace chip
