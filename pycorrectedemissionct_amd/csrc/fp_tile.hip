// Whole-sinogram forward projection by image tiles (gfx950).
//
// Discretisation: scikit-image 0.18.3 radon as called by corrct's skimage
// backend (_projector_backends.py:312-333) on vol * M[a, d]
// (projectors.py:519-525): sample (y, u) of angle a reads
//   x' = (c u + s y) + m02,  y' = (-s u + c y) + m12       (float32, no FMA)
// with an order-1 blend, 0 outside the slice; the line is the sum over y,
// times w[a, d].
//
// Decomposition. A sample contributes only when its cell (floor x', floor y')
// lies in [-1, N-1]^2; those cells are cut into tiles of 64 columns x 32 rows
// (the first tile of a row or column also owns cell -1, the last one ends at
// N - 1). A workgroup owns one tile of one or two slices and a group of
// consecutive angles:
//  * the slices' tile plus a one-pixel halo row and column stays in
//    registers for the whole group;
//  * per angle, the tile of each map plane is read once, in whole 256-byte
//    rows (neighbouring workgroups overlap only in the halo), multiplied by
//    the slice and stored in LDS with the C = slices x detectors channels of
//    a pixel side by side, while the next angle's map tiles are already being
//    fetched into registers;
//  * for a fixed bin u, x' and y' are monotone in y (also in float32: every
//    rounding is monotone), so the samples whose cell lies in the tile form one
//    interval of rows. The intervals depend on the geometry only: a set-up
//    kernel bounds them in double and refines them with the samples' own
//    float32 arithmetic, once per plan, into a table (so every sample is
//    counted by exactly one tile);
//  * eight lanes walk a bin's interval (rows lo + lane % 8 + 8 k); the
//    sample's position, floors and blend weights serve all C channels; the
//    eight lanes' sums are added in a fixed butterfly order (DPP lane moves)
//    and each (slice,
//    tile, angle) writes the partial sums of its <= 80 bins. The bin groups
//    rotate over the waves from angle to angle, so no SIMD always gets the
//    longest rays;
//  * fp_tile_reduce_kernel adds the partial sums of a line in a fixed order
//    (the tiles sorted by first bin) and applies w[a, d].
// The result is bitwise reproducible; it differs from the sequential sum over
// y by the order of the additions and by the blend, written as four weighted
// taps accumulated with fused multiply-adds.
#include "common.hpp"

// the sample coordinates are skimage's separate float32 products and sums:
// no contraction of a * b + c into one rounding (explicit fmaf stays fused)
#pragma clang fp contract(off)

namespace chip {
namespace {

__device__ __forceinline__ float mul_rn(float a, float b) { return a * b; }
__device__ __forceinline__ float add_rn(float a, float b) { return a + b; }

constexpr int kTX = kFpTileCols;        // cells per tile row
constexpr int kTY = kFpTileRows;        // cells per tile column
constexpr int kNb = kFpTileBins;        // bin slots per (tile, angle)
constexpr int kGroups = kNb / 8;        // lane groups of eight (one bin each)
constexpr int kTab = kFpTileTab;        // ints per (angle, tile) table record: ub, nb, 2 spare, kNb rows
constexpr int kRowsL = kTY + 2;         // LDS rows: tile rows -1 .. kTY
// LDS pixels per row (tile columns -4 .. kTX + 4 and up) for C channels per pixel:
// consecutive rows start 20 (C = 4), 12 (C = 2) or 8 (C = 1) banks apart, so the
// eight consecutive samples of a bin spread over the 64 banks; the pitch keeps the
// staging's float4 stores aligned
template <int C>
constexpr int pitch_of() { return C == 4 ? kTX + 5 : C == 2 ? kTX + 6 : kTX + 8; }
static_assert(pitch_of<4>() % 16 == 5 && pitch_of<2>() % 4 == 2 && pitch_of<1>() % 4 == 0, "bank spread");
constexpr int kQ = kTX / 4;             // float4 slots per tile row
constexpr int kMain = kTX * kTY / 4 / 256;  // float4 slots of the tile's own rows per thread
static_assert(kTX * kTY / 4 == 256 * kMain, "the tile's float4 slots split evenly over 256 threads");
static_assert(kQ + kTY + 1 <= 256, "one extra slot per thread holds the halo row and column");
static_assert(kNb % 8 == 0 && kTab == kNb + 4, "bin slots in groups of eight; table record layout");

// Workgroup p runs on XCD p % 8: logical ids are permuted so that each XCD
// walks a contiguous run of them (the tiles of one angle group, whose halos
// then meet in its L2).
__device__ __forceinline__ int xcd_logical(int p, int g) {
  const int q = g >> 3, r = g & 7;
  const int x = p & 7, k = p >> 3;
  return x * q + min(x, r) + k;
}

// the cells a tile owns and its origin
struct TileGeo {
  int x0, y0;              // first column / row of the tile's pixels
  int cx0, cx1, cy0, cy1;  // owned cells [cx0, cx1) x [cy0, cy1)
};

__device__ __forceinline__ TileGeo tile_geo(int t, int tiles_x, int n) {
  TileGeo g;
  const int ti = t / tiles_x, tj = t - ti * tiles_x;
  g.x0 = tj * kTX;
  g.y0 = ti * kTY;
  g.cx0 = tj == 0 ? -1 : g.x0;
  g.cy0 = ti == 0 ? -1 : g.y0;
  g.cx1 = min(g.x0 + kTX, n);
  g.cy1 = min(g.y0 + kTY, n);
  return g;
}

// Bins whose samples may lie in the tile: u of the four corners of the owned
// region under the inverse map, widened by 0.01 (float32 sample positions lie
// within 2e-3 of the exact ones for N <= 4096). At most
// sqrt(65^2 + 33^2) + 2.02 < 75 bins.
__device__ __forceinline__ int first_bin(const TileGeo& g, float cs, float sn, float m02, float m12, int n,
                                         int* nb) {
  const double c = cs, s = sn;
  const double xa = g.cx0 - static_cast<double>(m02), xb = g.cx1 - static_cast<double>(m02);
  const double ya = g.cy0 - static_cast<double>(m12), yb = g.cy1 - static_cast<double>(m12);
  const double u0 = c * xa - s * ya, u1 = c * xb - s * ya, u2 = c * xa - s * yb, u3 = c * xb - s * yb;
  const double umin = fmin(fmin(u0, u1), fmin(u2, u3)) - 0.01;
  const double umax = fmax(fmax(u0, u1), fmax(u2, u3)) + 0.01;
  const int ub = max(0, static_cast<int>(floor(fmax(umin, -1.0))));
  const int ue = min(n - 1, static_cast<int>(ceil(fmin(umax, static_cast<double>(n)))));
  *nb = ue - ub + 1;
  return ub;
}

// Is the cell of sample (u, y) in the tile? The same float32 arithmetic as the gather.
__device__ __forceinline__ bool member(const TileGeo& g, float xu, float yu, float cs, float sn, float m02, float m12,
                                       int y) {
  const float yf = static_cast<float>(y);
  const float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
  const float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
  return xs >= static_cast<float>(g.cx0) && xs < static_cast<float>(g.cx1) && ys >= static_cast<float>(g.cy0) &&
         ys < static_cast<float>(g.cy1);
}

// Rows [lo, hi] of bin u whose cells lie in the tile (lo > hi: none).
__device__ __forceinline__ int2 bin_rows(const TileGeo& g, int u, float cs, float sn, float m02, float m12, int n) {
  // superset from the exact map, widened by e (>= the float32 position error)
  const double e = 1e-6 * n + 1e-3;
  const double c = cs, s = sn;
  const double xu0 = c * u + m02, yu0 = -s * u + m12;
  double lo = 0.0, hi = n - 1.0;
  // x' = xu0 + s y in [cx0 - e, cx1 + e]
  if (s > 0.0) {
    lo = fmax(lo, (g.cx0 - e - xu0) / s);
    hi = fmin(hi, (g.cx1 + e - xu0) / s);
  } else if (s < 0.0) {
    lo = fmax(lo, (g.cx1 + e - xu0) / s);
    hi = fmin(hi, (g.cx0 - e - xu0) / s);
  } else if (xu0 < g.cx0 - e || xu0 > g.cx1 + e) {
    hi = -1.0;
  }
  // y' = yu0 + c y in [cy0 - e, cy1 + e]
  if (c > 0.0) {
    lo = fmax(lo, (g.cy0 - e - yu0) / c);
    hi = fmin(hi, (g.cy1 + e - yu0) / c);
  } else if (c < 0.0) {
    lo = fmax(lo, (g.cy1 + e - yu0) / c);
    hi = fmin(hi, (g.cy0 - e - yu0) / c);
  } else if (yu0 < g.cy0 - e || yu0 > g.cy1 + e) {
    hi = -1.0;
  }
  // (clamped before the conversion: near-axis angles divide by a tiny s or c)
  int ylo = static_cast<int>(ceil(fmin(fmax(lo, -1.0), static_cast<double>(n))));
  int yhi = static_cast<int>(floor(fmax(fmin(hi, static_cast<double>(n)), -1.0)));
  // the exact rows form an interval inside [ylo, yhi]: walk in to its ends
  const float uf = static_cast<float>(u);
  const float xu = mul_rn(cs, uf), yu = mul_rn(-sn, uf);
  while (ylo <= yhi && !member(g, xu, yu, cs, sn, m02, m12, ylo)) ++ylo;
  while (yhi >= ylo && !member(g, xu, yu, cs, sn, m02, m12, yhi)) --yhi;
  return make_int2(ylo, yhi);
}

// Table record of (angle a, tile t): [ub, nb, 0, 0, rows of bins ub .. ub + kNb - 1],
// each row entry packing the first row (low 16 bits) and the row count (high 16 bits).
__global__ __launch_bounds__(128) void fp_tile_table_kernel(int* __restrict__ tab, const float4* __restrict__ geom,
                                                            int n, int tiles_x, int n_tiles) {
  const int a = blockIdx.x, t = blockIdx.y;
  const TileGeo g = tile_geo(t, tiles_x, n);
  const float4 gm = geom[a];
  int nb;
  const int ub = first_bin(g, gm.x, gm.y, gm.z, gm.w, n, &nb);
  int* rec = tab + (static_cast<size_t>(a) * n_tiles + t) * kTab;
  const int j = threadIdx.x;
  if (j == 0) {
    rec[0] = ub;
    rec[1] = nb;
    rec[2] = 0;
    rec[3] = 0;
  }
  if (j < kNb) {
    int packed = 0;
    if (j < nb) {
      const int2 r = bin_rows(g, ub + j, gm.x, gm.y, gm.z, gm.w, n);
      if (r.x <= r.y) packed = r.x | ((r.y - r.x + 1) << 16);
    }
    rec[4 + j] = packed;
  }
}

// Per angle, the tiles sorted by (first bin, tile): keys (ub << 16) | t in
// ascending order (bitonic sort in LDS of the next power of two >= T keys), so
// that the tiles whose bins can include u form one run of the sorted list.
__global__ __launch_bounds__(256) void fp_tile_sort_kernel(const int* __restrict__ tab,
                                                           unsigned* __restrict__ sorted, int n_tiles, int p2) {
  extern __shared__ unsigned keys[];
  const int a = blockIdx.x;
  for (int i = threadIdx.x; i < p2; i += 256)
    keys[i] = i < n_tiles ? (static_cast<unsigned>(tab[(static_cast<size_t>(a) * n_tiles + i) * kTab]) << 16) |
                                static_cast<unsigned>(i)
                          : 0xffffffffu;
  __syncthreads();
  for (int k = 2; k <= p2; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < p2; i += 256) {
        const int ij = i ^ j;
        if (ij > i) {
          const unsigned x = keys[i], y = keys[ij];
          if ((x > y) == ((i & k) == 0)) {
            keys[i] = y;
            keys[ij] = x;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < n_tiles; i += 256) sorted[static_cast<size_t>(a) * n_tiles + i] = keys[i];
}

// v of the lane paired with this one inside its group of eight, by DPP lane moves
// (no LDS traffic): xor 1 and xor 2 within a quad, then the mirror of the
// eight-lane half row (lane i <-> 7 - i)
__device__ __forceinline__ float dpp_xor1(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_xor2(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_half_mirror(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
}

template <int C>
struct Pix;
template <>
struct Pix<1> {
  using T = float;
};
template <>
struct Pix<2> {
  using T = float2;
};
template <>
struct Pix<4> {
  using T = float4;
};
__device__ __forceinline__ float ch(float v, int) { return v; }
__device__ __forceinline__ float ch(float2 v, int c) { return c == 0 ? v.x : v.y; }
__device__ __forceinline__ float ch(float4 v, int c) { return c == 0 ? v.x : c == 1 ? v.y : c == 2 ? v.z : v.w; }

// One workgroup: tile t of slices [S b, S b + S), angles [la_begin, la_end) of
// the launch's range. D detectors per pass with maps (channels ordered slice
// major), or the bare slices (MAPS false, D = 1). PF: the next angle's map
// tiles are fetched into registers during the current angle's gathers (else
// loaded at staging, other workgroups covering the latency). U: samples per
// lane and loop trip.
template <int D, int S, bool MAPS, bool PF, int U>
__global__ __launch_bounds__(256) void fp_tile_kernel(const float* __restrict__ vol, const float* __restrict__ maps,
                                                      TileSlots slots, float* __restrict__ part,
                                                      const int* __restrict__ tab, const float4* __restrict__ geom,
                                                      int n, int n_det_total, int det0, int angle0, int n_ang,
                                                      int maps_a0, int group, int tiles_x, int n_tiles, int b_off) {
  constexpr int C = D * S;
  constexpr int kPitch = pitch_of<C>();
  using PixT = typename Pix<C>::T;
  __shared__ float tile[C * kRowsL * kPitch];
  // U = 1: the partial sums of the last two angles (by parity) are gathered in LDS and
  // stored to `part` after the next angle's staging, so that the staging's wait for the
  // prefetched map tiles does not also wait for the stores just issued (-2 % at 2 x 512^2
  // x 720 x 2, -4 % at 32 x 256^2 x 360); with U = 2 the extra registers cost occupancy
  constexpr bool kDefer = U == 1;
  __shared__ float psum[kDefer ? 2 : 1][kDefer ? S * kNb * D : 1];
  const int tid = threadIdx.x;
  const int lb = xcd_logical(blockIdx.x, gridDim.x);
  const int gi = lb / n_tiles;
  const int t = lb - gi * n_tiles;
  const int b0 = b_off + S * static_cast<int>(blockIdx.y);  // first slice of the workgroup
  const int la_begin = gi * group;
  const int la_end = min(la_begin + group, n_ang);
  const size_t nn = static_cast<size_t>(n) * n;
  const TileGeo g = tile_geo(t, tiles_x, n);

  // zero padding: row -1 and column -1 (read by cells -1 of the first tile
  // row / column) are never staged
  for (int i = tid; i < kPitch * C; i += 256) tile[i] = 0.f;
  for (int r = tid; r < kRowsL; r += 256)
#pragma unroll
    for (int c = 0; c < C; ++c) tile[(r * kPitch + 3) * C + c] = 0.f;

  // staging slots: kMain float4 of rows 0 .. kTY-1 (slot i: row tid / kQ +
  // (256 / kQ) i, columns 4 (tid % kQ) ..), then one extra slot (threads
  // 0 .. kQ-1: a float4 of halo row kTY; the next kTY + 1 threads: the halo
  // column kTX of row tid - kQ, a single pixel)
  const int r0 = tid / kQ, q0 = tid - r0 * kQ;
  const int off0 = (g.y0 + r0) * n + g.x0 + 4 * q0;  // slot i: off0 + (256 / kQ) i n
  const int lpix0 = (r0 + 1) * kPitch + 4 * q0 + 4;  // slot i: lpix0 + (256 / kQ) i kPitch
  const bool col_ok = g.x0 + 4 * q0 < n;
  unsigned okm = 0u;
#pragma unroll
  for (int i = 0; i < kMain; ++i)
    if (col_ok && g.y0 + r0 + (256 / kQ) * i < n) okm |= 1u << i;
  const bool ext_row = tid < kQ;
  const bool ext_col = tid >= kQ && tid < kQ + kTY + 1;
  int off_x, lpix_x;
  {
    const int gr = ext_row ? g.y0 + kTY : g.y0 + (tid - kQ);
    const int gc = ext_row ? g.x0 + 4 * tid : g.x0 + kTX;
    if ((ext_row || ext_col) && gr < n && gc < n) okm |= 1u << kMain;
    off_x = gr * n + gc;
    lpix_x = ext_row ? (kTY + 1) * kPitch + 4 * tid + 4 : (tid - kQ + 1) * kPitch + kTX + 4;
  }
  auto slot_off = [&](int i) { return i < kMain ? off0 + (256 / kQ) * i * n : off_x; };
  auto slot_lpix = [&](int i) { return i < kMain ? lpix0 + (256 / kQ) * i * kPitch : lpix_x; };

  float4 fr[S][kMain + 1];
#pragma unroll
  for (int sl = 0; sl < S; ++sl) {
    const float* v = vol + slots.vol(b0 + sl);
#pragma unroll
    for (int i = 0; i < kMain + 1; ++i) {
      fr[sl][i] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (okm & (1u << i)) {
        if (i < kMain || ext_row) fr[sl][i] = *reinterpret_cast<const float4*>(v + slot_off(i));
        else fr[sl][i].x = v[slot_off(i)];
      }
    }
  }

  constexpr int kM = MAPS ? S * D : 1;
  float4 mr[kM][kMain + 1];
  auto issue = [&](int la) {
    const int a = angle0 + la;
#pragma unroll
    for (int sl = 0; sl < S; ++sl)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const float* mb = maps + slots.maps(b0 + sl) + (static_cast<size_t>(a - maps_a0) * n_det_total + det0 + d) * nn;
#pragma unroll
        for (int i = 0; i < kMain + 1; ++i) {
          float4& m = mr[MAPS ? sl * D + d : 0][i];
          m = make_float4(0.f, 0.f, 0.f, 0.f);
          if (okm & (1u << i)) {
            if (i < kMain || ext_row) m = *reinterpret_cast<const float4*>(mb + slot_off(i));
            else m.x = mb[slot_off(i)];
          }
        }
      }
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < kMain + 1; ++i) {
      if (i == kMain && !ext_row && !ext_col) continue;
      const bool single = i == kMain && !ext_row;
      float p[C][4];  // channel c = sl * D + d
#pragma unroll
      for (int sl = 0; sl < S; ++sl)
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const float4 f = fr[sl][i];
          const float4 m = mr[MAPS ? sl * D + d : 0][i];
          p[sl * D + d][0] = MAPS ? f.x * m.x : f.x;
          p[sl * D + d][1] = MAPS ? f.y * m.y : f.y;
          p[sl * D + d][2] = MAPS ? f.z * m.z : f.z;
          p[sl * D + d][3] = MAPS ? f.w * m.w : f.w;
        }
      float* dst = tile + slot_lpix(i) * C;
      if (single) {
#pragma unroll
        for (int c = 0; c < C; ++c) dst[c] = p[c][0];
      } else {
        // 4 pixels x C channels, pixel major
#pragma unroll
        for (int h = 0; h < C; ++h) {
          float4 v4;
          v4.x = p[(4 * h + 0) % C][(4 * h + 0) / C];
          v4.y = p[(4 * h + 1) % C][(4 * h + 1) / C];
          v4.z = p[(4 * h + 2) % C][(4 * h + 2) / C];
          v4.w = p[(4 * h + 3) % C][(4 * h + 3) / C];
          *reinterpret_cast<float4*>(dst + 4 * h) = v4;
        }
      }
    }
  };

  const int wave = tid >> 6, lane = tid & 63;
  const int sub = lane & 7;  // the lane's phase within its bin
  const int lbase = (1 - g.y0) * kPitch + 4 - g.x0;  // LDS pixel of cell (r, c) = r kPitch + c + lbase
  // bin groups of this wave at angle la: g = w, w + 4, w + 8 with w = (wave + la) % 4
  constexpr int kGPerWave = (kGroups + 3) / 4;
  auto rows_of = [&](int la, int (&rw)[kGPerWave]) {
    const int* rec = tab + (static_cast<size_t>(angle0 + la) * n_tiles + t) * kTab + 4;
    const int w = (wave + la) & 3;
#pragma unroll
    for (int k = 0; k < kGPerWave; ++k) {
      const int gg = w + 4 * k;
      rw[k] = gg < kGroups ? rec[8 * gg + (lane >> 3)] : 0;
    }
  };

  int rw[kGPerWave];
  rows_of(la_begin, rw);
  if (MAPS) {
    if (PF) issue(la_begin);
  } else {
    stage();
  }

  // angle l's partial sums from LDS to `part` (after a barrier that follows their writes)
  auto store_part = [&](int l) {
    const float* src = psum[l & 1];
    for (int i = tid; i < S * kNb * D; i += 256) {
      const int sl = i / (kNb * D), r = i - sl * (kNb * D);
      part[((static_cast<size_t>(b0 + sl) * n_ang + l) * n_tiles + t) * (kNb * D) + r] = src[i];
    }
  };
  for (int la = la_begin; la < la_end; ++la) {
    const int a = angle0 + la;
    const float4 gm = geom[a];
    const float cs = gm.x, sn = gm.y, m02 = gm.z, m12 = gm.w;
    const int ub = tab[(static_cast<size_t>(a) * n_tiles + t) * kTab];
    const int nb = tab[(static_cast<size_t>(a) * n_tiles + t) * kTab + 1];
    if (MAPS && !PF) issue(la);
    __syncthreads();  // the previous angle's gathers are done with the tile
    // the staging waves outrank the other workgroups' gathers on their SIMDs: the tile's
    // four waves then leave the barrier sooner (2 % at 2 x 512^2 x 720 x 2)
    __builtin_amdgcn_s_setprio(3);
    if (MAPS) stage();
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    if (kDefer && la > la_begin) store_part(la - 1);
    int rcur[kGPerWave];
#pragma unroll
    for (int k = 0; k < kGPerWave; ++k) rcur[k] = rw[k];
    if (la + 1 < la_end) {
      rows_of(la + 1, rw);
      if (MAPS && PF) issue(la + 1);
    }
    const int w = (wave + la) & 3;
#pragma unroll 1
    for (int k = 0; k < kGPerWave; ++k) {
      const int gg = w + 4 * k;
      if (gg >= kGroups) break;
      const int j = 8 * gg + (lane >> 3);  // the lane group's bin (slot)
      if (8 * gg >= nb) {
        // no bin of the group meets the tile: its partial sums are 0
        if (sub == 0) {
#pragma unroll
          for (int sl = 0; sl < S; ++sl) {
            float* pr = kDefer ? psum[la & 1] + sl * (kNb * D)
                               : part + ((static_cast<size_t>(b0 + sl) * n_ang + la) * n_tiles + t) * (kNb * D);
            if (D == 2) reinterpret_cast<float2*>(pr)[j] = make_float2(0.f, 0.f);
            else pr[j] = 0.f;
          }
        }
        continue;
      }
      const int ylo = rcur[k] & 0xffff, cnt = rcur[k] >> 16;
      const float uf = static_cast<float>(ub + j);
      const float xu = mul_rn(cs, uf), yu = mul_rn(-sn, uf);
      float acc[C];
#pragma unroll
      for (int c = 0; c < C; ++c) acc[c] = 0.f;
      // rows counted in float (exact below 2^24)
      const float ylast = static_cast<float>(ylo + cnt - 1);
      for (float yf0 = static_cast<float>(ylo + sub); yf0 <= ylast; yf0 += 8.f * U) {
#pragma unroll
        for (int q = 0; q < U; ++q) {
          // samples past the interval repeat its last row with zero weights
          const bool in = q == 0 || yf0 + 8.f * q <= ylast;
          const float yf = q == 0 ? yf0 : fminf(yf0 + 8.f * q, ylast);
          const float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
          const float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
          const float c0f = floorf(xs);
          const float r0f = floorf(ys);
          const float dc = xs - c0f;
          const float dr = in ? ys - r0f : 0.f;
          const float dcw = in ? dc : 0.f;
          const float one = in ? 1.f : 0.f;
          const int li = static_cast<int>(fmaf(r0f, static_cast<float>(kPitch), c0f)) + lbase;
          const PixT* tp = reinterpret_cast<const PixT*>(tile) + li;
          const PixT t00 = tp[0], t01 = tp[1], t10 = tp[kPitch], t11 = tp[kPitch + 1];
          // order-1 blend as weights: w11 = dc dr, w01 = dc - w11, w10 = dr - w11, w00 = 1 - dc - w10
          const float w11 = dcw * dr;
          const float w01 = dcw - w11;
          const float w10 = dr - w11;
          const float w00 = (one - dcw) - w10;
#pragma unroll
          for (int c = 0; c < C; ++c) {
            acc[c] = fmaf(w00, ch(t00, c), acc[c]);
            acc[c] = fmaf(w01, ch(t01, c), acc[c]);
            acc[c] = fmaf(w10, ch(t10, c), acc[c]);
            acc[c] = fmaf(w11, ch(t11, c), acc[c]);
          }
        }
      }
      // fixed-order sum over the eight lanes of the bin (pairs, quads, then the two quads)
#pragma unroll
      for (int c = 0; c < C; ++c) {
        acc[c] += dpp_xor1(acc[c]);
        acc[c] += dpp_xor2(acc[c]);
        acc[c] += dpp_half_mirror(acc[c]);
      }
      if (sub == 0) {
#pragma unroll
        for (int sl = 0; sl < S; ++sl) {
          float* pr = kDefer ? psum[la & 1] + sl * (kNb * D)
                             : part + ((static_cast<size_t>(b0 + sl) * n_ang + la) * n_tiles + t) * (kNb * D);
          if (D == 2) reinterpret_cast<float2*>(pr)[j] = make_float2(acc[sl * D], acc[sl * D + D - 1]);
          else pr[j] = acc[sl * D];
        }
      }
    }
  }
  if (kDefer && la_end > la_begin) {
    __syncthreads();
    store_part(la_end - 1);
  }
}

// sino of slot b, detector det0 + d, angle a, bin u = w[a][det0 + d] * (sum of the
// tiles' partial sums of bin u, in the order of the angle's sorted tile list: by
// first bin, then tile); with D = 1 partial sums (no maps) every detector of
// [det0, det0 + nd) gets the same sum. A workgroup serves 256 bins [u0, u0 + 256):
// the tiles that can hold them (u0 - kNb < ub < u0 + 256) form one run of the
// sorted list; the workgroup walks it in order, each tile's partial sums read by
// the lanes whose bin it covers (one contiguous row of partial sums).
template <int D>
__global__ __launch_bounds__(256) void fp_tile_reduce_kernel(const float* __restrict__ part,
                                                             const unsigned* __restrict__ sorted,
                                                             float* __restrict__ sino, TileSlots slots,
                                                             const float* __restrict__ w, int n, int n_angles,
                                                             int n_det_total, int det0, int nd, int angle0, int n_ang,
                                                             int n_tiles) {
  extern __shared__ unsigned keys[];
  const int la = blockIdx.x, b = blockIdx.y;
  const int a = angle0 + la;
  for (int i = threadIdx.x; i < n_tiles; i += 256) keys[i] = sorted[static_cast<size_t>(a) * n_tiles + i];
  __syncthreads();
  const float* pr = part + (static_cast<size_t>(b) * n_ang + la) * n_tiles * (kNb * D);
  const int u0 = static_cast<int>(blockIdx.z) * 256;
  const int u = u0 + threadIdx.x;
  // first sorted position with key >= kk
  auto lower = [&](unsigned kk) {
    int lo = 0, hi = n_tiles;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (keys[mid] < kk) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int i0 = lower(u0 - kNb + 1 > 0 ? static_cast<unsigned>(u0 - kNb + 1) << 16 : 0u);
  const int i1 = lower(static_cast<unsigned>(u0 + 256) << 16);
  float s[D];
#pragma unroll
  for (int d = 0; d < D; ++d) s[d] = 0.f;
  // branch-free: tiles whose window misses u read their own first slot and add 0, so
  // eight loads per trip are in flight together
#pragma unroll 8
  for (int i = i0; i < i1; ++i) {
    const unsigned k = keys[i];
    const int j = u - static_cast<int>(k >> 16);
    const bool in = static_cast<unsigned>(j) < static_cast<unsigned>(kNb);
    const float* q = pr + (static_cast<size_t>(k & 0xffffu) * kNb + (in ? j : 0)) * D;
#pragma unroll
    for (int d = 0; d < D; ++d) s[d] += in ? q[d] : 0.f;
  }
  if (u >= n) return;
  float* out = sino + slots.sino(b);
  for (int d = 0; d < nd; ++d) {
    const float sd = s[D == 1 ? 0 : d];
    out[(static_cast<size_t>(det0 + d) * n_angles + a) * n + u] = sd * w[static_cast<size_t>(a) * n_det_total + det0 + d];
  }
}

template <int D, int S, bool MAPS, class... Args>
void launch_kind(int kind, dim3 grid, hipStream_t st, Args... args) {
  switch (kind) {
    case 1: hipLaunchKernelGGL((fp_tile_kernel<D, S, MAPS, true, 1>), grid, dim3(256), 0, st, args...); break;
    case 2: hipLaunchKernelGGL((fp_tile_kernel<D, S, MAPS, false, 2>), grid, dim3(256), 0, st, args...); break;
    case 3: hipLaunchKernelGGL((fp_tile_kernel<D, S, MAPS, false, 1>), grid, dim3(256), 0, st, args...); break;
    default: hipLaunchKernelGGL((fp_tile_kernel<D, S, MAPS, true, 2>), grid, dim3(256), 0, st, args...); break;
  }
}

}  // namespace

int fp_tile_count(int n) { return ((n + kTX - 1) / kTX) * ((n + kTY - 1) / kTY); }

static unsigned* fp_tile_sorted(const Plan& p) {
  return reinterpret_cast<unsigned*>(p.d_fp_tab + static_cast<size_t>(p.n_angles) * fp_tile_count(p.n) * kTab);
}

static int pow2_at_least(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

int fp_tile_table(const Plan& p, hipStream_t st) {
  const int n_tiles = fp_tile_count(p.n);
  if (n_tiles > kFpTileMaxTiles) {
    set_error("the image-tile fp handles at most " + std::to_string(kFpTileMaxTiles) + " tiles per slice");
    return CHIP_EINVAL;
  }
  // records [A][T][kTab], then the sorted tile keys [A][T]
  const size_t need = static_cast<size_t>(p.n_angles) * n_tiles * (kTab + 1) * sizeof(int);
  if (p.fp_tab_ok && p.fp_tab_bytes >= need) return CHIP_OK;
  if (stream_capturing(st)) {
    set_error("the image-tile fp builds its bin table (" + std::to_string(need) +
              " bytes) on first use: run it once outside stream capture first");
    return CHIP_EINVAL;
  }
  const int rc = ensure_buffer(reinterpret_cast<float*&>(p.d_fp_tab), p.fp_tab_bytes, need, st, "image-tile fp table");
  if (rc != CHIP_OK) return rc;
  const int tiles_x = (p.n + kTX - 1) / kTX;
  hipLaunchKernelGGL(fp_tile_table_kernel, dim3(p.n_angles, n_tiles), dim3(128), 0, st, p.d_fp_tab, p.d_fp_geom, p.n,
                     tiles_x, n_tiles);
  const int p2 = pow2_at_least(n_tiles);
  hipLaunchKernelGGL(fp_tile_sort_kernel, dim3(p.n_angles), dim3(256), sizeof(unsigned) * p2, st, p.d_fp_tab,
                     fp_tile_sorted(p), n_tiles, p2);
  CHIP_LAUNCH_CHECK();
  p.fp_tab_ok = true;
  return CHIP_OK;
}

size_t fp_tile_part_bytes(const Plan& p, int slots, int n_ang) {
  const int dp = std::min(2, std::max(1, p.n_det));
  return static_cast<size_t>(slots) * n_ang * fp_tile_count(p.n) * kNb * dp * sizeof(float);
}

void launch_fp_tile(const Plan& p, const float* vol, const float* maps, const TileSlots& sl, float* sino, int slots,
                    int angle0, int n_ang, int maps_a0, float* part, const FpTileShape& shape, hipStream_t st) {
  const int tiles_x = (p.n + kTX - 1) / kTX;
  const int n_tiles = fp_tile_count(p.n);
  const int group = std::max(1, std::min(shape.group, n_ang));
  const int n_groups = (n_ang + group - 1) / group;
  const int* tab = p.d_fp_tab;
  const unsigned* sorted = fp_tile_sorted(p);
  const size_t red_lds = sizeof(unsigned) * n_tiles;
  // slots in pairs (one workgroup serves both), the odd one alone
  const int pairs = shape.pairs ? slots / 2 : 0;
  const int singles = slots - 2 * pairs;
  const dim3 grid2(n_groups * n_tiles, pairs), grid1(n_groups * n_tiles, singles);
  const dim3 rgrid(n_ang, slots, (p.n + 255) / 256);
  // kind -1 (default): map tiles fetched one angle ahead with two detectors (four
  // channels); loaded at staging in one-detector passes, where the registers of the
  // prefetch cost more than it hides (32 x 256^2 x 360: -2 %)
  const int kind = shape.kind >= 0 ? shape.kind : 1;
  const int kind1 = shape.kind >= 0 ? shape.kind : 3;
  if (!maps) {
    if (pairs)
      launch_kind<1, 2, false>(kind, grid2, st, vol, maps, sl, part, tab, p.d_fp_geom, p.n, p.n_det, 0, angle0, n_ang,
                               maps_a0, group, tiles_x, n_tiles, 0);
    if (singles)
      launch_kind<1, 1, false>(kind, grid1, st, vol, maps, sl, part, tab, p.d_fp_geom, p.n, p.n_det, 0, angle0, n_ang,
                               maps_a0, group, tiles_x, n_tiles, 2 * pairs);
    hipLaunchKernelGGL((fp_tile_reduce_kernel<1>), rgrid, dim3(256), red_lds, st, part, sorted, sino, sl, p.d_w, p.n,
                       p.n_angles, p.n_det, 0, p.n_det, angle0, n_ang, n_tiles);
    return;
  }
  for (int det0 = 0; det0 < p.n_det; det0 += 2) {
    const int nd = p.n_det - det0 >= 2 ? 2 : 1;
    if (nd == 2) {
      if (pairs)
        launch_kind<2, 2, true>(kind, grid2, st, vol, maps, sl, part, tab, p.d_fp_geom, p.n, p.n_det, det0, angle0,
                                n_ang, maps_a0, group, tiles_x, n_tiles, 0);
      if (singles)
        launch_kind<2, 1, true>(kind, grid1, st, vol, maps, sl, part, tab, p.d_fp_geom, p.n, p.n_det, det0, angle0,
                                n_ang, maps_a0, group, tiles_x, n_tiles, 2 * pairs);
      hipLaunchKernelGGL((fp_tile_reduce_kernel<2>), rgrid, dim3(256), red_lds, st, part, sorted, sino, sl, p.d_w, p.n,
                         p.n_angles, p.n_det, det0, 2, angle0, n_ang, n_tiles);
    } else {
      if (pairs)
        launch_kind<1, 2, true>(kind1, grid2, st, vol, maps, sl, part, tab, p.d_fp_geom, p.n, p.n_det, det0, angle0,
                                n_ang, maps_a0, group, tiles_x, n_tiles, 0);
      if (singles)
        launch_kind<1, 1, true>(kind1, grid1, st, vol, maps, sl, part, tab, p.d_fp_geom, p.n, p.n_det, det0, angle0,
                                n_ang, maps_a0, group, tiles_x, n_tiles, 2 * pairs);
      hipLaunchKernelGGL((fp_tile_reduce_kernel<1>), rgrid, dim3(256), red_lds, st, part, sorted, sino, sl, p.d_w, p.n,
                         p.n_angles, p.n_det, det0, 1, angle0, n_ang, n_tiles);
    }
  }
}

ScratchUse::ScratchUse(BpPartState& s, hipStream_t st) : s_(s), st_(st), lk_(s.mu) {
  if (stream_capturing(st)) return;
  if (s.used && s.last != st && s.done) (void)hipStreamWaitEvent(st, s.done, 0);
  if (!s.done) (void)hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
  record_ = s.done != nullptr;
}

ScratchUse::~ScratchUse() {
  if (!record_) return;
  if (hipEventRecord(s_.done, st_) == hipSuccess) {
    s_.last = st_;
    s_.used = true;
  }
}

}  // namespace chip
