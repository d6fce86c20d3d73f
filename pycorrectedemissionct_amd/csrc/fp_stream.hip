// Whole-sinogram forward projection streaming the attenuation maps (gfx950).
//
// Discretisation: scikit-image 0.18.3 radon as called by corrct's skimage
// backend (_projector_backends.py:312-333) on vol * M[a, d]
// (projectors.py:519-525): sample (y, u) of angle a reads
//   x' = (c u + s y) + m02,  y' = (-s u + c y) + m12       (float32, no FMA)
// with an order-1 blend, 0 outside the slice; the line is the sum over y,
// times w[a, d].
//
// Decomposition: a workgroup owns one angle and 64 consecutive bins and walks
// the rays in chunks of 32 sample rows. For each chunk the bounding box of
// the chunk's bilinear footprint (monotone float32 map => spanned by the four
// corner samples, +1 for the right/bottom taps) is staged in LDS as
// T_d = vol * M[a, d] for the two detectors of the pass; the 2048 samples then
// gather their taps from LDS. Staging is software-pipelined: the next
// chunk's box is fetched into registers as float4 (box columns start on a
// multiple of 4; with N % 4 == 0 each float4 lies wholly inside or outside
// the slice) while the current chunk is gathered. Each slot of the box has a
// per-thread constant offset, so interior chunks cost one add per slot;
// chunks touching the slice border take a checked path that also writes
// the zero padding.
//
// Thread (bin = tid % 64, group = tid / 64) accumulates 8 consecutive rows
// of each chunk in order and the four groups are added in a fixed order, so
// the result is bitwise reproducible.
//
// Workgroup p runs on XCD p % 8; logical block ids are permuted so that the
// bin-blocks of one angle run back to back on one XCD and share its L2.
#include "common.hpp"

// the sample coordinates are skimage's separate float32 products and sums:
// no contraction of a * b + c into one rounding (explicit fmaf stays fused)
#pragma clang fp contract(off)

namespace chip {
namespace {

// single-rounding float32 product and sum (no contraction under this file's pragma)
__device__ __forceinline__ float mul_rn(float a, float b) { return a * b; }
__device__ __forceinline__ float add_rn(float a, float b) { return a + b; }

// Tile shapes: BINS detector bins x ROWS sample rows per chunk, 256 threads in
// 256 / BINS row groups. The staged box of a chunk spans at most
// (BINS - 1)|s| + (ROWS - 1)|c| + 3 rows (+1 on the border, where clamped
// samples tap N and N + 1) and as many columns plus 3 for the float4
// alignment of its first column; the bounds below hold for every angle
// (max of a|s| + b|c| is sqrt(a^2 + b^2)).
template <int BINS, int ROWS, int BOX_ROWS, int SLOTS, int PITCH, int WAVES_PER_SIMD>
struct FpTile {
  static constexpr int kWaves = WAVES_PER_SIMD;  // occupancy the register budget is sized for
  static constexpr int kBins = BINS;
  static constexpr int kRows = ROWS;
  static constexpr int kGroups = 256 / BINS;
  static constexpr int kRowsPerGroup = ROWS / kGroups;
  static constexpr int kBoxRows = BOX_ROWS;
  static constexpr int kSlots = SLOTS;  // float4 slots per box row
  static constexpr int kPitch = PITCH;  // floats per tile row and detector, a multiple of 4
  static constexpr int kTile = BOX_ROWS * PITCH;  // floats per detector; a pass holds D * kTile
  static constexpr int kPerThread = (BOX_ROWS * SLOTS + 255) / 256;
  static_assert(ROWS % kGroups == 0, "rows split evenly over the groups");
  static_assert(PITCH % 4 == 0 && PITCH >= 4 * SLOTS, "float4-aligned rows holding every slot");
};
// 64 x 32: sqrt(63^2 + 31^2) = 70.3 -> 74.3 rows, 77.3 columns
using Tile64 = FpTile<64, 32, 76, 20, 84, 1>;
// 32 x 32: sqrt(31^2 + 31^2) = 43.9 -> 47.9 rows, 50.9 columns; a third of the LDS
using Tile32 = FpTile<32, 32, 48, 13, 52, 1>;
// the persistent batch kernel's tile; the batch plan counts its bin blocks with kFpBatchBins
using BatchTile = Tile32;
static_assert(BatchTile::kBins == kFpBatchBins, "batch plan and kernel agree on the bin block");

typedef float f4v __attribute__((ext_vector_type(4)));

struct Box {
  int clo, chi, rlo, rhi;  // inclusive; clo is a multiple of 4 (or -4)
  int ylo, yhi;            // sample rows of the chunk
  int all_valid;           // every sample of the chunk lies in (-1, N)^2
  int empty;               // the chunk misses the slice
  int border;              // the box reaches outside the slice
};

__device__ __forceinline__ int logical_block(int p, int g) {
  const int q = g >> 3, r = g & 7;
  const int x = p & 7, k = p >> 3;
  return x * q + min(x, r) + k;
}

template <class T>
__device__ __forceinline__ Box chunk_box(int yc, int n, float fn, float xu_a, float yu_a, float xu_b, float yu_b,
                                         float cs, float sn, float m02, float m12) {
  const int y_last = min(yc + T::kRows, n) - 1;
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
  Box bx;
  bx.ylo = yc;
  bx.yhi = y_last;
  bx.all_valid = (xmin > -1.f && xmax < fn && ymin > -1.f && ymax < fn) ? 1 : 0;
  const int clo = max(static_cast<int>(floorf(fmaxf(xmin, -2.f))), -1);
  bx.chi = min(static_cast<int>(floorf(fminf(xmax, fn + 1.f))) + 1, n);
  bx.rlo = max(static_cast<int>(floorf(fmaxf(ymin, -2.f))), -1);
  bx.rhi = min(static_cast<int>(floorf(fminf(ymax, fn + 1.f))) + 1, n);
  bx.clo = clo >= 0 ? (clo & ~3) : -4;
  bx.empty = (bx.chi < 0 || clo > n - 1 || bx.rhi < 0 || bx.rlo > n - 1) ? 1 : 0;
  bx.border = (bx.rlo < 0 || bx.rhi > n - 1 || bx.clo < 0 || bx.chi > n - 1) ? 1 : 0;
  // border chunks clamp sample coordinates to [-1, N]; a sample clamped to N
  // reads taps N and N + 1, so the box then carries one more zero row/column
  if (bx.chi == n) bx.chi = n + 1;
  if (bx.rhi == n) bx.rhi = n + 1;
  return bx;
}

// A box packed into two ints for LDS: columns / rows offset by 8 in 13 bits
// each (N <= 4096), the empty and border flags in bits 26 and 27.
constexpr int kMaxChunks = 128;  // chunk boxes precomputed per workgroup (N <= 128 * rows)

__device__ __forceinline__ int2 pack_box(const Box& b) {
  int2 p;
  p.x = (b.clo + 8) | ((b.chi + 8) << 13) | (b.empty << 26) | (b.border << 27);
  p.y = (b.rlo + 8) | ((b.rhi + 8) << 13);
  return p;
}

__device__ __forceinline__ Box unpack_box(int2 p, int yc, int rows, int n) {
  Box b;
  b.clo = (p.x & 8191) - 8;
  b.chi = ((p.x >> 13) & 8191) - 8;
  b.empty = (p.x >> 26) & 1;
  b.border = (p.x >> 27) & 1;
  b.rlo = (p.y & 8191) - 8;
  b.rhi = ((p.y >> 13) & 8191) - 8;
  b.ylo = yc;
  b.yhi = min(yc + rows, n) - 1;
  b.all_valid = 0;
  return b;
}

// One workgroup's work: bins [u0, u0 + 64) of one angle of one slice, for the
// detectors [det0, det0 + D) of a pass.
struct FpItem {
  const float* v;         // slice [n][n]
  const float* mb;        // maps of (angle, det0): D planes of n*n, or NULL
  float* out;             // sinogram line of (det0, angle); detector d at out + d * det_stride
  size_t det_stride;      // elements between detectors' lines (A * n)
  const float* wa;        // weights of (angle, det0 .. det0 + D)
  float4 gm;              // (cos, sin, m02, m12) of the angle
  int n;
  int u0;
};

// TRIM: interior chunks stage only the slots whose centre lies within 3.5
// pixels (in the rotated frame of the samples) of the chunk's sample
// rectangle -- a tapped pixel lies within sqrt(2) of a sample and a slot
// centre within 1.5 of its pixels -- instead of the whole axis-aligned box,
// which at 45 degrees is twice the footprint. Border chunks (clamped samples
// read the zero padding) keep the whole box.
//
// PATCH (32-bin tiles): wave w gathers bins [8 w, 8 w + 8), lane l sample rows
// l / 8 + 8 k of each chunk, so one gather instruction of a wave reads an 8 x 8
// patch of samples (a compact region of the staged tile, fewer LDS bank
// conflicts than 32 bins of one row at steep angles) instead of 32 bins x 2 rows.
template <class T, int D, bool MAPS, bool TRIM, bool PATCH = false>
__device__ __forceinline__ void fp_stream_block(const FpItem& it, float* lds) {
  static_assert(D == 1 || D == 2, "one or two detectors per pass");
  static_assert(!PATCH || (T::kBins == 32 && T::kRows == 32), "8 x 8 patches tile a 32 x 32 chunk");
  constexpr int kBins = T::kBins, kRows = T::kRows, kGroups = T::kGroups, kRowsPerGroup = T::kRowsPerGroup;
  constexpr int kSlots = T::kSlots, kPitch = T::kPitch, kPerThread = T::kPerThread;
  const int n = it.n;
  const int u0 = it.u0;
  const int tid = threadIdx.x;
  const int lane = PATCH ? 8 * (tid >> 6) + (tid & 7) : tid % kBins;  // the thread's bin in the block
  const int grp = PATCH ? (tid & 63) >> 3 : tid / kBins;               // its row group
  const int u = u0 + lane;

  const float4 gm = it.gm;
  const float cs = gm.x, sn = gm.y, m02 = gm.z, m12 = gm.w;
  const size_t nn = static_cast<size_t>(n) * n;
  const float* v = it.v;
  const float* mb = it.mb;
  const float fn = static_cast<float>(n);
  const float uf = static_cast<float>(u);
  const float xu = mul_rn(cs, uf);
  const float yu = mul_rn(-sn, uf);
  const int u_last = min(u0 + kBins, n) - 1;
  const float xu_a = mul_rn(cs, static_cast<float>(u0)), yu_a = mul_rn(-sn, static_cast<float>(u0));
  const float xu_b = mul_rn(cs, static_cast<float>(u_last)), yu_b = mul_rn(-sn, static_cast<float>(u_last));
  const float uf_lo = static_cast<float>(u0) - 3.5f, uf_hi = static_cast<float>(u_last) + 3.5f;

  // staging slots of this thread: box row, first column, offset from the box origin
  int s_row[kPerThread], s_col[kPerThread], s_off[kPerThread];
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    const int s = tid + 256 * i;
    s_row[i] = s / kSlots;
    s_col[i] = 4 * (s - s_row[i] * kSlots);
    s_off[i] = s_row[i] * n + s_col[i];
  }

  float4 fr[kPerThread];
  float4 mr[D][kPerThread];
  unsigned inmask = 0u;

  auto issue = [&](const Box& bx) {
    const int th = bx.rhi - bx.rlo + 1;
    const int cw = bx.chi - bx.clo + 1;
    const int base = bx.rlo * n + bx.clo;
    inmask = 0u;
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      bool in = s_row[i] < th && s_col[i] < cw;
      if (bx.border) {
        const int r = bx.rlo + s_row[i];
        const int c = bx.clo + s_col[i];
        in = in && static_cast<unsigned>(r) < static_cast<unsigned>(n) &&
             static_cast<unsigned>(c) < static_cast<unsigned>(n);
      } else if (TRIM) {
        const float px = static_cast<float>(bx.clo + s_col[i]) + (1.5f - m02);
        const float py = static_cast<float>(bx.rlo + s_row[i]) - m12;
        const float us = cs * px - sn * py;  // the slot centre in sample coordinates
        const float ysm = sn * px + cs * py;
        in = in && us >= uf_lo && us <= uf_hi && ysm >= static_cast<float>(bx.ylo) - 3.5f &&
             ysm <= static_cast<float>(bx.yhi) + 3.5f;
      }
      if (in) {
        inmask |= 1u << i;
        // uniform base + 32-bit byte offset: scalar-base addressing, no 64-bit temporaries
        const unsigned off = static_cast<unsigned>(base + s_off[i]) * 4u;
        fr[i] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(v) + off);
        if (MAPS) {
          // regular (L2-cached) loads: neighbouring bin-blocks re-read the overlap of their boxes
#pragma unroll
          for (int d = 0; d < D; ++d)
            mr[d][i] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(mb + d * nn) + off);
        }
      }
    }
  };
  auto stage = [&](const Box& bx) {
    const int th = bx.rhi - bx.rlo + 1;
    const int cw = bx.chi - bx.clo + 1;
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      // tile layout [row][col][D]: the D detectors' products of one pixel are adjacent
      const int o = (s_row[i] * kPitch + s_col[i]) * D;
      if (inmask & (1u << i)) {
        float p[D][4];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          p[d][0] = MAPS ? fr[i].x * mr[d][i].x : fr[i].x;
          p[d][1] = MAPS ? fr[i].y * mr[d][i].y : fr[i].y;
          p[d][2] = MAPS ? fr[i].z * mr[d][i].z : fr[i].z;
          p[d][3] = MAPS ? fr[i].w * mr[d][i].w : fr[i].w;
        }
#pragma unroll
        for (int h = 0; h < D; ++h) {
          // float4 h holds columns 4h / D .. of the slot, detectors interleaved
          float4 v4;
          v4.x = p[(4 * h + 0) % D][(4 * h + 0) / D];
          v4.y = p[(4 * h + 1) % D][(4 * h + 1) / D];
          v4.z = p[(4 * h + 2) % D][(4 * h + 2) / D];
          v4.w = p[(4 * h + 3) % D][(4 * h + 3) / D];
          *reinterpret_cast<float4*>(lds + o + 4 * h) = v4;
        }
      } else if (bx.border && s_row[i] < th && s_col[i] < cw) {  // zero padding outside the slice
#pragma unroll
        for (int h = 0; h < D; ++h) *reinterpret_cast<float4*>(lds + o + 4 * h) = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  };

  float acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0.f;

  // the chunk boxes are uniform over the workgroup: computed once, one chunk per thread,
  // into LDS behind the tile (slices of more than kMaxChunks chunks compute them per chunk)
  const int n_chunks = (n + kRows - 1) / kRows;
  const bool lds_boxes = n_chunks <= kMaxChunks;
  int2* boxes = reinterpret_cast<int2*>(lds + D * T::kTile);
  if (lds_boxes) {
    for (int c = tid; c < n_chunks; c += 256)
      boxes[c] = pack_box(chunk_box<T>(c * kRows, n, fn, xu_a, yu_a, xu_b, yu_b, cs, sn, m02, m12));
    __syncthreads();
  }
  auto box_of = [&](int y) -> Box {
    return lds_boxes ? unpack_box(boxes[y / kRows], y, kRows, n)
                     : chunk_box<T>(y, n, fn, xu_a, yu_a, xu_b, yu_b, cs, sn, m02, m12);
  };

  int yc = 0;
  Box cur;
  for (;; yc += kRows) {
    if (yc >= n) break;
    cur = box_of(yc);
    if (!cur.empty) break;
  }
  if (yc < n) issue(cur);
  while (yc < n) {
    __syncthreads();  // the previous chunk's gathers are done with the tile
    stage(cur);
    __syncthreads();
    int yn = yc + kRows;  // next chunk meeting the slice; its loads overlap this chunk's gathers
    Box nxt;
    for (; yn < n; yn += kRows) {
      nxt = box_of(yn);
      if (!nxt.empty) break;
    }
    if (yn < n) issue(nxt);

    if (u < n) {
      // rows yb0 + k * ystep, k < kRowsPerGroup (PATCH: grp + 8 k; else 4 consecutive rows, wave-uniform)
      const int yb0 = PATCH ? yc + grp : yc + grp * kRowsPerGroup;
      constexpr int ystep = PATCH ? 8 : 1;
      const int base = cur.rlo * kPitch + cur.clo;
      // one sample: skimage inverse map, taps from LDS, order-1 blend
      auto sample = [&](int y, bool clamp) {
        const float yf = static_cast<float>(y);
        float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
        float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
        if (clamp) {
          // outside (-1, N) every tap is padding: clamping keeps the taps on
          // the zero rows/columns of the box and the sample at exactly 0
          xs = fminf(fmaxf(xs, -1.f), fn);
          ys = fminf(fmaxf(ys, -1.f), fn);
        }
        const float c0f = floorf(xs);
        const float r0f = floorf(ys);
        const float dc = xs - c0f;
        const float dr = ys - r0f;
        const int li = static_cast<int>(fmaf(r0f, static_cast<float>(kPitch), c0f)) - base;
        float t00[D], t01[D], t10[D], t11[D];
        if (D == 2) {
          const float2* t = reinterpret_cast<const float2*>(lds) + li;
          const float2 a0 = t[0], a1 = t[1], b0 = t[kPitch], b1 = t[kPitch + 1];
          t00[0] = a0.x, t00[D - 1] = a0.y, t01[0] = a1.x, t01[D - 1] = a1.y;
          t10[0] = b0.x, t10[D - 1] = b0.y, t11[0] = b1.x, t11[D - 1] = b1.y;
        } else {
          const float* t = lds + li;
          t00[0] = t[0], t01[0] = t[1], t10[0] = t[kPitch], t11[0] = t[kPitch + 1];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const float top = fmaf(dc, t01[d] - t00[d], t00[d]);
          const float bot = fmaf(dc, t11[d] - t10[d], t10[d]);
          acc[d] += fmaf(dr, bot - top, top);
        }
      };
      // full row groups run straight-line, so the LDS reads of several samples overlap
      const bool full = yb0 + (kRowsPerGroup - 1) * ystep < n;
      if (!cur.border) {
        if (full) {
#pragma unroll
          for (int k = 0; k < kRowsPerGroup; ++k) sample(yb0 + k * ystep, false);
        } else {
#pragma unroll
          for (int k = 0; k < kRowsPerGroup; ++k)
            if (yb0 + k * ystep < n) sample(yb0 + k * ystep, false);
        }
      } else {
        if (full) {
#pragma unroll
          for (int k = 0; k < kRowsPerGroup; ++k) sample(yb0 + k * ystep, true);
        } else {
#pragma unroll
          for (int k = 0; k < kRowsPerGroup; ++k)
            if (yb0 + k * ystep < n) sample(yb0 + k * ystep, true);
        }
      }
    }
    yc = yn;
    cur = nxt;
  }

  __syncthreads();
#pragma unroll
  for (int d = 0; d < D; ++d) lds[(d * kGroups + grp) * kBins + lane] = acc[d];
  __syncthreads();
  if (grp == 0 && u < n) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float s = lds[(d * kGroups) * kBins + lane];
#pragma unroll
      for (int g = 1; g < kGroups; ++g) s = s + lds[(d * kGroups + g) * kBins + lane];
      it.out[d * it.det_stride + u] = s * it.wa[d];
    }
  }
}

// ---------------------------------------------------------------------------
// Direct-to-LDS staging (fp_kernel 7): the chunk boxes of the slice and of the
// D map planes go from HBM to LDS with gfx950's buffer_load_dwordx4 ... lds,
// no VGPR round trip and no multiply at staging; the sample gathers read the
// slice and map taps from their own planes and form the products fl(v * M)
// there (the same values, so the same bits as fp_kernel 5). Slots outside the
// slice, outside the chunk's footprint (interior chunks) or past the box get an
// out-of-range buffer offset: the hardware writes zeros and reads nothing,
// which is also the zero padding border chunks need. One buffer per
// workgroup; the other workgroups of the CU overlap a chunk's load latency.
// ---------------------------------------------------------------------------
constexpr unsigned kOob = 0x80000000u;  // byte offset past any plane: loads return 0
constexpr int kDirectPlane = BatchTile::kBoxRows * 4 * BatchTile::kSlots;  // floats per staged plane

typedef __attribute__((address_space(3))) void lds_void_t;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const float* base, size_t nn) {
  // raw buffer over one n x n plane (gfx9 word 3: 32-bit data, no swizzle)
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(base), 0, static_cast<int>(nn * sizeof(float)),
                                           0x00020000);
}

template <int D, bool TRIM>
__device__ __forceinline__ void fp_direct_block(const FpItem& it, float* lds) {
  using T = BatchTile;
  constexpr int kBins = T::kBins, kRows = T::kRows, kGroups = T::kGroups, kRowsPerGroup = T::kRowsPerGroup;
  constexpr int kSlots = T::kSlots, kPitch = 4 * T::kSlots;
  constexpr int kIter = (T::kBoxRows * kSlots + 255) / 256;
  const int n = it.n;
  const int u0 = it.u0;
  const int tid = threadIdx.x;
  const int lane = tid % kBins;
  const int grp = tid / kBins;
  const int u = u0 + lane;
  const float4 gm = it.gm;
  const float cs = gm.x, sn = gm.y, m02 = gm.z, m12 = gm.w;
  const size_t nn = static_cast<size_t>(n) * n;
  const float fn = static_cast<float>(n);
  const float uf = static_cast<float>(u);
  const float xu = mul_rn(cs, uf);
  const float yu = mul_rn(-sn, uf);
  const int u_last = min(u0 + kBins, n) - 1;
  const float xu_a = mul_rn(cs, static_cast<float>(u0)), yu_a = mul_rn(-sn, static_cast<float>(u0));
  const float xu_b = mul_rn(cs, static_cast<float>(u_last)), yu_b = mul_rn(-sn, static_cast<float>(u_last));
  const float uf_lo = static_cast<float>(u0) - 3.5f, uf_hi = static_cast<float>(u_last) + 3.5f;

  // buffer resources of the slice and of the (up to two) map planes (an opaque type: no arrays)
  const __amdgpu_buffer_rsrc_t rs_v = plane_rsrc(it.v, nn);
  const __amdgpu_buffer_rsrc_t rs_m0 = plane_rsrc(it.mb, nn);
  const __amdgpu_buffer_rsrc_t rs_m1 = plane_rsrc(it.mb + (D - 1) * nn, nn);

  // slots of this thread: flat index s = i * 256 + tid of the box in rows of kSlots float4
  int s_row[kIter], s_col[kIter];
#pragma unroll
  for (int i = 0; i < kIter; ++i) {
    const int sl = i * 256 + tid;
    s_row[i] = sl / kSlots;
    s_col[i] = 4 * (sl - s_row[i] * kSlots);
  }
  const int wave_slot0 = tid & ~63;  // the wave's first slot of each pass: LDS destinations are wave-uniform

  auto issue = [&](const Box& bx) {
    const int th = bx.rhi - bx.rlo + 1;
#pragma unroll
    for (int i = 0; i < kIter; ++i) {
      if (s_row[i] < th) {
        const int r = bx.rlo + s_row[i];
        const int c = bx.clo + s_col[i];
        bool in = c <= bx.chi && static_cast<unsigned>(r) < static_cast<unsigned>(n) &&
                  static_cast<unsigned>(c) < static_cast<unsigned>(n);
        if (TRIM && !bx.border) {
          const float px = static_cast<float>(c) + (1.5f - m02);
          const float py = static_cast<float>(r) - m12;
          const float us = cs * px - sn * py;
          const float ysm = sn * px + cs * py;
          in = in && us >= uf_lo && us <= uf_hi && ysm >= static_cast<float>(bx.ylo) - 3.5f &&
               ysm <= static_cast<float>(bx.yhi) + 3.5f;
        }
        const unsigned off = in ? static_cast<unsigned>(r * n + c) * 4u : kOob;
        float* dst = lds + 4 * (i * 256 + wave_slot0);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs_v, (lds_void_t*)dst, 16, off, 0, 0, 0);
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rs_m0, (lds_void_t*)(dst + kDirectPlane), 16, off, 0, 0, 0);
        if (D == 2)
          __builtin_amdgcn_raw_ptr_buffer_load_lds(rs_m1, (lds_void_t*)(dst + 2 * kDirectPlane), 16, off, 0, 0, 0);
      }
    }
  };

  float acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0.f;

  const int n_chunks = (n + kRows - 1) / kRows;
  const bool lds_boxes = n_chunks <= kMaxChunks;
  int2* boxes = reinterpret_cast<int2*>(lds + (1 + D) * kDirectPlane);
  if (lds_boxes) {
    for (int c = tid; c < n_chunks; c += 256)
      boxes[c] = pack_box(chunk_box<T>(c * kRows, n, fn, xu_a, yu_a, xu_b, yu_b, cs, sn, m02, m12));
    __syncthreads();
  }
  auto box_of = [&](int y) -> Box {
    return lds_boxes ? unpack_box(boxes[y / kRows], y, kRows, n)
                     : chunk_box<T>(y, n, fn, xu_a, yu_a, xu_b, yu_b, cs, sn, m02, m12);
  };

  for (int yc = 0; yc < n; yc += kRows) {
    const Box cur = box_of(yc);
    if (cur.empty) continue;
    __syncthreads();  // the previous chunk's gathers are done with the planes
    issue(cur);
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this thread's loads have reached LDS
    __syncthreads();
    if (u < n) {
      const int yb0 = yc + grp * kRowsPerGroup;
      const int base = cur.rlo * kPitch + cur.clo;
      auto sample = [&](int y, bool clamp) {
        const float yf = static_cast<float>(y);
        float xs = add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
        float ys = add_rn(add_rn(yu, mul_rn(cs, yf)), m12);
        if (clamp) {
          xs = fminf(fmaxf(xs, -1.f), fn);
          ys = fminf(fmaxf(ys, -1.f), fn);
        }
        const float c0f = floorf(xs);
        const float r0f = floorf(ys);
        const float dc = xs - c0f;
        const float dr = ys - r0f;
        const int li = static_cast<int>(fmaf(r0f, static_cast<float>(kPitch), c0f)) - base;
        const float* f = lds + li;
        const float f00 = f[0], f01 = f[1], f10 = f[kPitch], f11 = f[kPitch + 1];
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const float* m = f + (1 + d) * kDirectPlane;
          const float t00 = f00 * m[0], t01 = f01 * m[1], t10 = f10 * m[kPitch], t11 = f11 * m[kPitch + 1];
          const float top = fmaf(dc, t01 - t00, t00);
          const float bot = fmaf(dc, t11 - t10, t10);
          acc[d] += fmaf(dr, bot - top, top);
        }
      };
      const bool full = yb0 + kRowsPerGroup <= n;
      if (!cur.border) {
        if (full) {
#pragma unroll
          for (int k = 0; k < kRowsPerGroup; ++k) sample(yb0 + k, false);
        } else {
#pragma unroll
          for (int k = 0; k < kRowsPerGroup; ++k)
            if (yb0 + k < n) sample(yb0 + k, false);
        }
      } else {
#pragma unroll
        for (int k = 0; k < kRowsPerGroup; ++k)
          if (yb0 + k < n) sample(yb0 + k, true);
      }
    }
  }

  __syncthreads();
#pragma unroll
  for (int d = 0; d < D; ++d) lds[(d * kGroups + grp) * kBins + lane] = acc[d];
  __syncthreads();
  if (grp == 0 && u < n) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      float sum = lds[(d * kGroups) * kBins + lane];
#pragma unroll
      for (int g = 1; g < kGroups; ++g) sum = sum + lds[(d * kGroups + g) * kBins + lane];
      it.out[d * it.det_stride + u] = sum * it.wa[d];
    }
  }
}

template <int D, bool TRIM>
__global__ __launch_bounds__(256) void fp_direct_kernel(const float* __restrict__ vol, const float* __restrict__ maps,
                                                         int64_t maps_slice_stride, float* __restrict__ sino,
                                                         const float4* __restrict__ geom, const float* __restrict__ w,
                                                         int n, int n_angles, int n_det_total, int det0, int n_ublk,
                                                         int angle0, int maps_a0) {
  extern __shared__ float lds[];  // (1 + D) planes [kBoxRows][4 kSlots], then the chunk boxes
  const int lb = logical_block(blockIdx.x, gridDim.x);
  const int la = lb / n_ublk;
  const int a = angle0 + la;
  const int b = blockIdx.y;
  const size_t nn = static_cast<size_t>(n) * n;
  FpItem it;
  it.n = n;
  it.u0 = (lb - la * n_ublk) * BatchTile::kBins;
  it.gm = geom[a];
  it.v = vol + static_cast<size_t>(b) * nn;
  it.mb = maps + static_cast<size_t>(b) * maps_slice_stride + (static_cast<size_t>(a - maps_a0) * n_det_total + det0) * nn;
  it.out = sino + ((static_cast<size_t>(b) * n_det_total + det0) * n_angles + a) * n;
  it.det_stride = static_cast<size_t>(n_angles) * n;
  it.wa = w + static_cast<size_t>(a) * n_det_total + det0;
  fp_direct_block<D, TRIM>(it, lds);
}

template <int D, bool TRIM>
__global__ __launch_bounds__(256) void fp_direct_batch_kernel(const BatchSlice* __restrict__ slices, int n_slices,
                                                               const int64_t* __restrict__ item_start,
                                                               const float* __restrict__ vol,
                                                               const float* __restrict__ maps, float* __restrict__ sino,
                                                               const float4* __restrict__ geom,
                                                               const float* __restrict__ w, int n_angles,
                                                               int n_det_total, int det0) {
  extern __shared__ float lds[];
  const int64_t total = item_start[n_slices];
  const int g = gridDim.x;
  const int64_t lb = logical_block(blockIdx.x, g);
  for (int64_t k = lb; k < total; k += g) {
    const int s = find_slice(item_start, n_slices, k);
    const BatchSlice sl = slices[s];
    const int local = static_cast<int>(k - item_start[s]);
    const int a = local / sl.n_ublk;
    const size_t nn = static_cast<size_t>(sl.n) * sl.n;
    FpItem it;
    it.n = sl.n;
    it.u0 = (local - a * sl.n_ublk) * BatchTile::kBins;
    it.gm = geom[static_cast<size_t>(sl.size_idx) * n_angles + a];
    it.v = vol + sl.vol_off;
    it.mb = maps + sl.maps_off + (static_cast<size_t>(a) * n_det_total + det0) * nn;
    it.out = sino + sl.sino_off + (static_cast<size_t>(det0) * n_angles + a) * sl.n;
    it.det_stride = static_cast<size_t>(n_angles) * sl.n;
    it.wa = w + static_cast<size_t>(a) * n_det_total + det0;
    fp_direct_block<D, TRIM>(it, lds);
    __syncthreads();  // the next item restages the planes
  }
}

template <class T, int D, bool MAPS, bool TRIM, bool PATCH = false>
__global__ __launch_bounds__(256, T::kWaves) void fp_stream_kernel(const float* __restrict__ vol, const float* __restrict__ maps,
                                                        int64_t maps_slice_stride, float* __restrict__ sino,
                                                        const float4* __restrict__ geom, const float* __restrict__ w,
                                                        int n, int n_angles, int n_det_total, int det0, int n_ublk,
                                                        int angle0, int maps_a0) {
  extern __shared__ float lds[];  // [kBoxRows][kPitch][D]
  const int lb = logical_block(blockIdx.x, gridDim.x);
  const int la = lb / n_ublk;
  const int a = angle0 + la;
  const int b = blockIdx.y;
  const size_t nn = static_cast<size_t>(n) * n;
  FpItem it;
  it.n = n;
  it.u0 = (lb - la * n_ublk) * T::kBins;
  it.gm = geom[a];
  it.v = vol + static_cast<size_t>(b) * nn;
  it.mb = MAPS ? maps + static_cast<size_t>(b) * maps_slice_stride +
                     (static_cast<size_t>(a - maps_a0) * n_det_total + det0) * nn
               : nullptr;
  it.out = sino + ((static_cast<size_t>(b) * n_det_total + det0) * n_angles + a) * n;
  it.det_stride = static_cast<size_t>(n_angles) * n;
  it.wa = w + static_cast<size_t>(a) * n_det_total + det0;
  fp_stream_block<T, D, MAPS, TRIM, PATCH>(it, lds);
}

// Persistent form over a batch of slices of different sizes: work item k of
// the batch is (slice s, angle a, bin block) with s found in the item prefix
// sums; workgroup p takes items r * G + logical_block(p, G), r = 0, 1, ...
// (G workgroups), so each XCD walks a contiguous run of bin blocks.
template <int D, bool MAPS, bool TRIM, bool PATCH = false>
__global__ __launch_bounds__(256) void fp_stream_batch_kernel(const BatchSlice* __restrict__ slices, int n_slices,
                                                              const int64_t* __restrict__ item_start,
                                                              const float* __restrict__ vol,
                                                              const float* __restrict__ maps, float* __restrict__ sino,
                                                              const float4* __restrict__ geom,
                                                              const float* __restrict__ w, int n_angles,
                                                              int n_det_total, int det0) {
  extern __shared__ float lds[];  // [kBoxRows][kPitch][D]
  const int64_t total = item_start[n_slices];
  const int g = gridDim.x;
  const int64_t lb = logical_block(blockIdx.x, g);
  for (int64_t k = lb; k < total; k += g) {
    const int s = find_slice(item_start, n_slices, k);
    const BatchSlice sl = slices[s];
    const int local = static_cast<int>(k - item_start[s]);
    const int a = local / sl.n_ublk;
    const size_t nn = static_cast<size_t>(sl.n) * sl.n;
    FpItem it;
    it.n = sl.n;
    it.u0 = (local - a * sl.n_ublk) * BatchTile::kBins;
    it.gm = geom[static_cast<size_t>(sl.size_idx) * n_angles + a];
    it.v = vol + sl.vol_off;
    it.mb = MAPS ? maps + sl.maps_off + (static_cast<size_t>(a) * n_det_total + det0) * nn : nullptr;
    it.out = sino + sl.sino_off + (static_cast<size_t>(det0) * n_angles + a) * sl.n;
    it.det_stride = static_cast<size_t>(n_angles) * sl.n;
    it.wa = w + static_cast<size_t>(a) * n_det_total + det0;
    fp_stream_block<BatchTile, D, MAPS, TRIM, PATCH>(it, lds);
  }
}

template <int D, bool MAPS>
void launch(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch, int det0,
            int angle0, int n_ang, int maps_a0, hipStream_t st) {
  if (MAPS && (p.fp_variant == 7 || p.fp_variant == 8) && p.n * p.n <= (1 << 29)) {
    const int n_ublk = (p.n + BatchTile::kBins - 1) / BatchTile::kBins;
    dim3 grid(n_ublk * n_ang, batch);
    const size_t shmem = sizeof(float) * (1 + D) * kDirectPlane + sizeof(int2) * kMaxChunks;
    if (p.fp_variant == 8)
      hipLaunchKernelGGL((fp_direct_kernel<D, true>), grid, dim3(256), shmem, st, vol, maps, stride, sino,
                         p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, n_ublk, angle0, maps_a0);
    else
      hipLaunchKernelGGL((fp_direct_kernel<D, false>), grid, dim3(256), shmem, st, vol, maps, stride, sino,
                         p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, n_ublk, angle0, maps_a0);
    return;
  }
  if (p.fp_variant == 5 || p.fp_variant == 7 || p.fp_variant == 8 || p.fp_variant == 9) {
    const int n_ublk = (p.n + Tile32::kBins - 1) / Tile32::kBins;
    dim3 grid(n_ublk * n_ang, batch);
    const size_t shmem = sizeof(float) * D * Tile32::kTile + sizeof(int2) * kMaxChunks;
    if (p.fp_variant == 9)
      hipLaunchKernelGGL((fp_stream_kernel<Tile32, D, MAPS, false, true>), grid, dim3(256), shmem, st, vol, maps,
                         stride, sino, p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, n_ublk, angle0, maps_a0);
    else
      hipLaunchKernelGGL((fp_stream_kernel<Tile32, D, MAPS, false>), grid, dim3(256), shmem, st, vol, maps, stride,
                         sino, p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, n_ublk, angle0, maps_a0);
    return;
  }
  const int n_ublk = (p.n + Tile64::kBins - 1) / Tile64::kBins;
  dim3 grid(n_ublk * n_ang, batch);
  const size_t shmem = sizeof(float) * D * Tile64::kTile + sizeof(int2) * kMaxChunks;
  if (p.fp_variant == 4) {
    hipLaunchKernelGGL((fp_stream_kernel<Tile64, D, MAPS, true>), grid, dim3(256), shmem, st, vol, maps, stride, sino,
                       p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, n_ublk, angle0, maps_a0);
  } else {
    hipLaunchKernelGGL((fp_stream_kernel<Tile64, D, MAPS, false>), grid, dim3(256), shmem, st, vol, maps, stride,
                       sino, p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, n_ublk, angle0, maps_a0);
  }
}

template <int D, bool MAPS>
void launch_batch(const BatchPlan& bp, const float* vol, const float* maps, float* sino, int det0, hipStream_t st) {
  const int64_t total = bp.h_item_start_fp.back();
  // fp_wg_per_cu = 0: one workgroup per work item (a flat grid), else a persistent grid
  const int64_t want = bp.fp_wg_per_cu > 0 ? static_cast<int64_t>(bp.fp_wg_per_cu) * bp.n_cu : total;
  const int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(total, want)));
  if (bp.fp_patch) {
    const size_t shmem_p = sizeof(float) * D * BatchTile::kTile + sizeof(int2) * kMaxChunks;
    hipLaunchKernelGGL((fp_stream_batch_kernel<D, MAPS, false, true>), dim3(grid), dim3(256), shmem_p, st,
                       bp.d_slices, bp.n_slices, bp.d_item_start_fp, vol, maps, sino, bp.d_geom, bp.d_w, bp.n_angles,
                       bp.n_det, det0);
    return;
  }
  if (MAPS && bp.fp_direct) {
    const size_t shmem_d = sizeof(float) * (1 + D) * kDirectPlane + sizeof(int2) * kMaxChunks;
    if (bp.fp_trim)
      hipLaunchKernelGGL((fp_direct_batch_kernel<D, true>), dim3(grid), dim3(256), shmem_d, st, bp.d_slices,
                         bp.n_slices, bp.d_item_start_fp, vol, maps, sino, bp.d_geom, bp.d_w, bp.n_angles, bp.n_det,
                         det0);
    else
      hipLaunchKernelGGL((fp_direct_batch_kernel<D, false>), dim3(grid), dim3(256), shmem_d, st, bp.d_slices,
                         bp.n_slices, bp.d_item_start_fp, vol, maps, sino, bp.d_geom, bp.d_w, bp.n_angles, bp.n_det,
                         det0);
    return;
  }
  const size_t shmem = sizeof(float) * D * BatchTile::kTile + sizeof(int2) * kMaxChunks;
  if (bp.fp_trim) {
    hipLaunchKernelGGL((fp_stream_batch_kernel<D, MAPS, true>), dim3(grid), dim3(256), shmem, st, bp.d_slices,
                       bp.n_slices, bp.d_item_start_fp, vol, maps, sino, bp.d_geom, bp.d_w, bp.n_angles, bp.n_det,
                       det0);
  } else {
    hipLaunchKernelGGL((fp_stream_batch_kernel<D, MAPS, false>), dim3(grid), dim3(256), shmem, st, bp.d_slices,
                       bp.n_slices, bp.d_item_start_fp, vol, maps, sino, bp.d_geom, bp.d_w, bp.n_angles, bp.n_det,
                       det0);
  }
}

}  // namespace

void launch_fp_stream_batch(const BatchPlan& bp, const float* vol, const float* maps, float* sino, hipStream_t st) {
  for (int det0 = 0; det0 < bp.n_det; det0 += 2) {
    const bool two = bp.n_det - det0 >= 2;
    if (maps) {
      if (two) launch_batch<2, true>(bp, vol, maps, sino, det0, st);
      else launch_batch<1, true>(bp, vol, maps, sino, det0, st);
    } else {
      if (two) launch_batch<2, false>(bp, vol, nullptr, sino, det0, st);
      else launch_batch<1, false>(bp, vol, nullptr, sino, det0, st);
    }
  }
}

void launch_fp_stream(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch,
                      int det0, int nd, int angle0, int n_ang, int maps_a0, hipStream_t st) {
  if (maps) {
    if (nd == 2) launch<2, true>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, maps_a0, st);
    else launch<1, true>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, maps_a0, st);
  } else {
    // without maps every detector sees the same slice
    if (nd == 2) launch<2, false>(p, vol, nullptr, stride, sino, batch, det0, angle0, n_ang, 0, st);
    else launch<1, false>(p, vol, nullptr, stride, sino, batch, det0, angle0, n_ang, 0, st);
  }
}

}  // namespace chip
