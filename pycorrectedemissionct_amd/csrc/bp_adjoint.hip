// Exact adjoint of the forward projection as a scatter-free gather (gfx950).
//
// The fp (fp_stream.hip) blends four taps of vol * M[a, d] at every sample
// (y, u) of angle a: pixel (i, j) receives the weight
//   hat(Px - j) * hat(Py - i),  hat(t) = max(0, 1 - |t|),
//   (Px, Py) = (c u + s y + m02, -s u + c y + m12),
// from each sample whose footprint covers it. The adjoint therefore is
//   out[i, j] = sum_a sum_d w[a, d] M[a, d][i, j] sum_u K_a(u; i, j) line[d, a, u],
//   K_a(u; i, j) = sum_y hat(Px(u, y) - j) hat(Py(u, y) - i).
// In lattice coordinates the covering samples lie within |u - u*| < |c| + |s|
// <= sqrt(2) (same for y) of the pixel's preimage (u*, y*), i.e. in a 3 x 3
// block of lattice points; each pixel evaluates those nine weights directly
// (no scatter, no atomics, deterministic). Samples outside [0, N) in u or y do
// not exist in the fp and get weight 0. The offsets (Px - j, Py - i) are one
// per-pixel offset plus per-angle lattice constants; they differ from the
// fp's float32 coordinates by rounding only, and the tent weights are
// continuous, so <A x, y> = <x, A^T y> holds to about 1e-7 relative.
//
// A workgroup owns a 32 x 8 pixel tile (rows of 128 contiguous bytes for the
// map reads) and walks all angles; per block of 32 angles the line segments
// the tile needs (<= 40 bins per angle and detector) are staged in LDS.
#include "common.hpp"

#include <type_traits>

namespace chip {
namespace {

constexpr int kTileW = 32;
constexpr int kTileH = 8;
constexpr int kAngBlk = 32;
constexpr int kGrp = 8;  // angles whose map values are loaded together
constexpr int kWin = 40;  // >= 31 |c| + 7 |s| + 2 (|c| + |s|) + 2

// One workgroup's work: the 32 x 8 pixel tile (tx, ty) of S slices of one size
// (their candidate weights are shared, each slice has its own lines, maps and sums).
template <int S = 1>
struct BpItemS {
  const float* sino[S];  // slice sinograms [D_total][A][n]
  const float* maps[S];  // slice maps [A - maps_a0 ...][D_total][n][n], or NULL
  float* vol[S];         // slices [n][n]
  const float4* geom;    // [A] of this slice size
  int n;
  int tx, ty;
};
using BpItem = BpItemS<1>;

// hat(t) = max(0, 1 - |t|): 1 - |t| <= 1, so clamping it to [0, 1] is the same
// value, which the hardware applies as an output modifier of the subtraction
__device__ __forceinline__ float tent(float t) { return __builtin_amdgcn_fmed3f(1.f - fabsf(t), 0.f, 1.f); }

// Per angle of a block (one 16-byte LDS read per angle): the constant terms of the
// first-candidate arguments, the flat window index of bin 0 and whether every
// candidate row of the tile lies in [0, N).
struct __align__(16) BpAngle {
  float au;  // 1 - (|c| + |s|) - c m02 + s m12
  float ay;  // 1 - (|c| + |s|) - s m02 - c m12
  int wbase;  // byte offset into the window block of bin 0 (may be negative)
  int rows_inside;
};

template <int D, bool MAPS, int S = 1>
__device__ __forceinline__ void bp_adjoint_block(const BpItemS<S>& itm, const float* __restrict__ w, int n_angles,
                                                 int n_det_total, int det0, int angle0, int n_ang, int maps_a0,
                                                 int accumulate, float (*win)[S * D][kWin], BpAngle* ang_s) {
  constexpr int SD = S * D;  // channels (slice major)
  // angles per group of map loads (the prefetched group holds kG x S x D values)
  constexpr int kG = S == 1 ? kGrp : (S == 2 ? kGrp / 2 : kGrp / (2 * D));
  const int n = itm.n;
  const float4* __restrict__ geom = itm.geom;
  const int tid = threadIdx.x;
  const int j = itm.tx * kTileW + (tid & (kTileW - 1));
  const int i = itm.ty * kTileH + (tid / kTileW);
  const size_t nn = static_cast<size_t>(n) * n;
  const bool inside = i < n && j < n;
  const size_t pix = static_cast<size_t>(min(i, n - 1)) * n + min(j, n - 1);
  const float fj = static_cast<float>(j), fi = static_cast<float>(i);
  // tile corners for the per-angle bin window and row range
  const float tj0 = static_cast<float>(itm.tx * kTileW), tj1 = tj0 + (kTileW - 1);
  const float ti0 = static_cast<float>(itm.ty * kTileH), ti1 = ti0 + (kTileH - 1);

  float out[S];
#pragma unroll
  for (int sl = 0; sl < S; ++sl) out[sl] = 0.f;
  for (int a0 = angle0; a0 < angle0 + n_ang; a0 += kAngBlk) {
    const int na = min(kAngBlk, angle0 + n_ang - a0);
    __syncthreads();
    // the block's staging outranks the other workgroups' pixel loops on their SIMDs (as in
    // the image-tile fp; within 1 % here)
    __builtin_amdgcn_s_setprio(3);
    if (tid < na) {
      const float4 g = geom[a0 + tid];
      const float cs = g.x, sn = g.y;
      const float r = fabsf(cs) + fabsf(sn);
      // preimage (u*, y*) = (c (x - m02) - s (y - m12), s (x - m02) + c (y - m12)) over the tile corners
      const float u00 = cs * (tj0 - g.z) - sn * (ti0 - g.w), u01 = cs * (tj1 - g.z) - sn * (ti0 - g.w);
      const float u10 = cs * (tj0 - g.z) - sn * (ti1 - g.w), u11 = cs * (tj1 - g.z) - sn * (ti1 - g.w);
      const float y00 = sn * (tj0 - g.z) + cs * (ti0 - g.w), y01 = sn * (tj1 - g.z) + cs * (ti0 - g.w);
      const float y10 = sn * (tj0 - g.z) + cs * (ti1 - g.w), y11 = sn * (tj1 - g.z) + cs * (ti1 - g.w);
      const float umin = fminf(fminf(u00, u01), fminf(u10, u11));
      const float ymin = fminf(fminf(y00, y01), fminf(y10, y11));
      const float ymax = fmaxf(fmaxf(y00, y01), fmaxf(y10, y11));
      BpAngle& ba = ang_s[tid];
      const int wlo = static_cast<int>(floorf(umin - r)) - 1;  // first window bin
      ba.wbase = static_cast<int>(sizeof(float)) * (tid * (SD * kWin) - wlo);
      const float omr = 1.f - r;
      ba.au = omr - cs * g.z + sn * g.w;
      ba.ay = omr - sn * g.z - cs * g.w;
      // candidate rows are floor(y* - r) + 1 .. + 3 (one of slack for rounding)
      ba.rows_inside = (floorf(ymin - r) >= 0.f && floorf(ymax - r) + 4.f <= static_cast<float>(n - 1)) ? 1 : 0;
    }
    __syncthreads();
    // one slice at a time, so that its line pointer is selected at compile time
#pragma unroll
    for (int sl = 0; sl < S; ++sl) {
      const float* __restrict__ line = itm.sino[sl];
      for (int k = tid; k < na * D * kWin; k += 256) {
        const int al = k / (D * kWin);
        const int rem = k - al * (D * kWin);
        const int d = rem / kWin;
        const int q = rem - d * kWin;
        const int c = sl * D + d;
        const int u = al * (SD * kWin) - ang_s[al].wbase / static_cast<int>(sizeof(float)) + q;
        // the line times its weight w[a, d], applied once per bin instead of once per pixel
        win[al][c][q] = (u >= 0 && u < n) ? line[(static_cast<size_t>(det0 + d) * n_angles + a0 + al) * n + u] *
                                                w[static_cast<size_t>(a0 + al) * n_det_total + det0 + d]
                                          : 0.f;
      }
    }
    __builtin_amdgcn_s_setprio(0);
    __syncthreads();
    if (inside) {
      // map values of the next group of kGrp angles are loaded while the current group is
      // computed: with one pixel per thread the per-angle loads would otherwise expose the
      // full memory latency once per angle
      float ma[kG][SD], mb2[kG][SD];
      auto load_maps = [&](float (*m)[SD], int g0) {
#pragma unroll
        for (int k = 0; k < kG; ++k)
#pragma unroll
          for (int c = 0; c < SD; ++c)
            m[k][c] = (MAPS && g0 + k < na)
                          ? itm.maps[c / D][(static_cast<size_t>(a0 + g0 + k - maps_a0) * n_det_total + det0 + c % D) *
                                                nn +
                                            pix]
                          : 1.f;
      };
      const char* wbytes = reinterpret_cast<const char*>(&win[0][0][0]);
      auto angle_group = [&](const float (*m)[SD], int g0) {
#pragma unroll
        for (int k = 0; k < kG; ++k) {
          const int al = g0 + k;
          if (al >= na) break;
          const float4 g = geom[a0 + al];
          const float cs = g.x, sn = g.y;
          const BpAngle ba = ang_s[al];
          // first of <= 3 candidate bins / rows: floor(p* - r) + 1 for the pixel centre's
          // preimage (u*, y*) = (c (j - m02) - s (i - m12), s (j - m02) + c (i - m12)) and
          // r = |c| + |s|; the angle's constant terms are folded into (au, ay)
          const float fu0 = floorf(fmaf(cs, fj, fmaf(-sn, fi, ba.au)));
          const float fy0 = floorf(fmaf(sn, fj, fmaf(cs, fi, ba.ay)));
          // sample offsets from the pixel: (c u + s y + m02 - j, -s u + c y + m12 - i) at (fu0, fy0)
          const float dx0 = fmaf(cs, fu0, fmaf(sn, fy0, g.z - fj));
          const float dy0 = fmaf(-sn, fu0, fmaf(cs, fy0, g.w - fi));
          // candidate (ku, ky) sits at (dx0, dy0) + ku (c, -s) + ky (s, c) from the pixel: the
          // offsets are stepped by adds (rounding of order 1e-7 at |offset| <= 3, below the
          // float32 rounding of the samples' own coordinates)
          float K[3];
          auto kernel_weights = [&](auto masked) {
            const int iy0 = static_cast<int>(fy0);
            float ax = dx0, ay = dy0;
#pragma unroll
            for (int ku = 0; ku < 3; ++ku) {
              float bx = ax, by = ay, ksum = 0.f;
#pragma unroll
              for (int ky = 0; ky < 3; ++ky) {
                const float wx = tent(bx);
                float wy = tent(by);
                // rows outside [0, N) are not samples of the fp and weigh 0
                if (decltype(masked)::value)
                  wy = (static_cast<unsigned>(iy0 + ky) < static_cast<unsigned>(n)) ? wy : 0.f;
                ksum = fmaf(wx, wy, ksum);
                bx += sn;
                by += cs;
              }
              K[ku] = ksum;
              ax += cs;
              ay -= sn;
            }
          };
          if (ba.rows_inside) kernel_weights(std::false_type{});
          else kernel_weights(std::true_type{});
          // window of this angle, bin fu0 (bins outside [0, N) read the zero padding)
          const float* wq =
              reinterpret_cast<const float*>(wbytes + (ba.wbase + static_cast<int>(sizeof(float)) * static_cast<int>(fu0)));
#pragma unroll
          for (int c = 0; c < SD; ++c) {
            const float gd = fmaf(K[0], wq[c * kWin], fmaf(K[1], wq[c * kWin + 1], K[2] * wq[c * kWin + 2]));
            out[c / D] = MAPS ? fmaf(m[k][c], gd, out[c / D]) : out[c / D] + gd;
          }
        }
      };
      load_maps(ma, 0);
      for (int g0 = 0; g0 < na; g0 += kG) {
        if (g0 + kG < na) load_maps(mb2, g0 + kG);
        angle_group(ma, g0);
#pragma unroll
        for (int k = 0; k < kG; ++k)
#pragma unroll
          for (int c = 0; c < SD; ++c) ma[k][c] = mb2[k][c];
      }
    }
  }
  if (inside) {
#pragma unroll
    for (int sl = 0; sl < S; ++sl) {
      float* o = itm.vol[sl] + pix;
      if (accumulate > 1) unsafeAtomicAdd(o, out[sl]);  // angle halves of a zeroed slice (see launch)
      else *o = accumulate ? *o + out[sl] : out[sl];
    }
  }
}

// part != NULL: one slice, blockIdx.z = one of split parts of the angle range, each
// stored to its own plane of part. split = 1 (part NULL): one slice, blockIdx.z = half of
// the angle range; both halves add their sums to the zeroed slice (two addends onto 0
// give the same bits in either order). Otherwise blockIdx.z = slice of a batch.
template <int D, bool MAPS>
__global__ __launch_bounds__(256) void bp_adjoint_tile_kernel(const float* __restrict__ sino,
                                                              const float* __restrict__ maps,
                                                              int64_t maps_slice_stride, float* __restrict__ vol,
                                                              const float4* __restrict__ geom,
                                                              const float* __restrict__ w, int n, int n_angles,
                                                              int n_det_total, int det0, int angle0, int n_ang,
                                                              int maps_a0, int accumulate, int split,
                                                              float* __restrict__ part) {
  __shared__ float win[kAngBlk][D][kWin];
  __shared__ BpAngle ang_s[kAngBlk];
  int b = blockIdx.z;
  if (part) {
    // one slice, blockIdx.z = part s of split parts of the angle range; each part stores its
    // own sum in part[s] (plain stores: the parts are added in order by a later pass)
    const int s = blockIdx.z, lo = n_ang * s / split, hi = n_ang * (s + 1) / split;
    angle0 += lo;
    n_ang = hi - lo;
    b = 0;
    vol = part + static_cast<size_t>(s) * n * n;
    accumulate = 0;
  } else if (split) {
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
  const size_t nn = static_cast<size_t>(n) * n;
  BpItem itm;
  itm.sino[0] = sino + static_cast<size_t>(b) * n_det_total * n_angles * n;
  itm.maps[0] = MAPS ? maps + static_cast<size_t>(b) * maps_slice_stride : nullptr;
  itm.vol[0] = vol + static_cast<size_t>(b) * nn;
  itm.geom = geom;
  itm.n = n;
  itm.tx = blockIdx.x;
  itm.ty = blockIdx.y;
  bp_adjoint_block<D, MAPS>(itm, w, n_angles, n_det_total, det0, angle0, n_ang, maps_a0, accumulate, win, ang_s);
}

// Persistent form over a batch of slices of different sizes: item k is tile
// (tx, ty) of slice s, s found in the tile prefix sums.
template <int D, bool MAPS>
__global__ __launch_bounds__(256) void bp_adjoint_batch_kernel(const BatchSlice* __restrict__ slices, int n_slices,
                                                               const int64_t* __restrict__ item_start,
                                                               const float* __restrict__ sino,
                                                               const float* __restrict__ maps,
                                                               float* __restrict__ vol,
                                                               const float4* __restrict__ geom,
                                                               const float* __restrict__ w, int n_angles,
                                                               int n_det_total, int det0, int accumulate) {
  __shared__ float win[kAngBlk][D][kWin];
  __shared__ BpAngle ang_s[kAngBlk];
  const int64_t total = item_start[n_slices];
  for (int64_t k = blockIdx.x; k < total; k += gridDim.x) {
    const int s = find_slice(item_start, n_slices, k);
    const BatchSlice sl = slices[s];
    const int local = static_cast<int>(k - item_start[s]);
    const int ntx = (sl.n + kTileW - 1) / kTileW;
    BpItem itm;
    itm.sino[0] = sino + sl.sino_off;
    itm.maps[0] = MAPS ? maps + sl.maps_off : nullptr;
    itm.vol[0] = vol + sl.vol_off;
    itm.geom = geom + static_cast<size_t>(sl.size_idx) * n_angles;
    itm.n = sl.n;
    itm.ty = local / ntx;
    itm.tx = local - itm.ty * ntx;
    bp_adjoint_block<D, MAPS>(itm, w, n_angles, n_det_total, det0, 0, n_angles, 0, accumulate, win, ang_s);
    __syncthreads();  // the next item restages the windows
  }
}

// Slices of one size at the given slots: blockIdx.z = unit, unit u serves the S
// slots [slot0 + S u, slot0 + S u + S) (with S = 2 or 4 one workgroup per two or
// four slices: the candidate weights are computed once for all of them; every
// slice gets the bits of the one-slice kernels).
template <int D, bool MAPS, int S>
__global__ __launch_bounds__(256) void bp_adjoint_slots_kernel(const float* __restrict__ sino,
                                                               const float* __restrict__ maps,
                                                               float* __restrict__ vol, TileSlots slots,
                                                               int slot0, const float4* __restrict__ geom,
                                                               const float* __restrict__ w, int n, int n_angles,
                                                               int n_det_total, int det0, int angle0, int n_ang,
                                                               int maps_a0, int accumulate, int parts,
                                                               float* __restrict__ part, int64_t part_plane) {
  __shared__ float win[kAngBlk][S * D][kWin];
  __shared__ BpAngle ang_s[kAngBlk];
  // blockIdx.z = unit x parts + part: with parts > 1 each part sums its share of the angle
  // range into its own plane of `part` (laid out like the volume), added in part order later
  const int unit = static_cast<int>(blockIdx.z) / parts, pi = static_cast<int>(blockIdx.z) - unit * parts;
  if (parts > 1) {
    const int lo = n_ang * pi / parts, hi = n_ang * (pi + 1) / parts;
    angle0 += lo;
    n_ang = hi - lo;
    accumulate = 0;
  }
  BpItemS<S> itm;
#pragma unroll
  for (int sl = 0; sl < S; ++sl) {
    const int b = slot0 + S * unit + sl;
    itm.sino[sl] = sino + slots.sino(b);
    itm.maps[sl] = MAPS ? maps + slots.maps(b) : nullptr;
    itm.vol[sl] = parts > 1 ? part + pi * part_plane + slots.vol(b) : vol + slots.vol(b);
  }
  itm.geom = geom;
  itm.n = n;
  itm.tx = blockIdx.x;
  itm.ty = blockIdx.y;
  bp_adjoint_block<D, MAPS, S>(itm, w, n_angles, n_det_total, det0, angle0, n_ang, maps_a0, accumulate, win, ang_s);
}

// out = the parts' sums added in part order (the same bits on every run)
__global__ void sum_parts_kernel(float* __restrict__ out, const float* __restrict__ part, int n_parts, int64_t nn) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  float acc = part[i];
  for (int s = 1; s < n_parts; ++s) acc += part[s * nn + i];
  out[i] = acc;
}

// out[i] = (accumulate ? out[i] : 0) + the parts' sums in part order, over a packed volume
__global__ void sum_parts_acc_kernel(float* __restrict__ out, const float* __restrict__ part, int n_parts,
                                     int64_t total, int accumulate) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) return;
  float acc = part[i];
  for (int s = 1; s < n_parts; ++s) acc += part[s * total + i];
  out[i] = accumulate ? out[i] + acc : acc;
}

// Parts of one slice's angle range (2 .. max_parts) that make its tiles x parts
// workgroups the closest to whole waves of what the GPU holds at once (ties: fewer
// parts); each part keeps at least 16 angles. 0 when the slice needs no split. The
// count, and with it the order of the sums, follows the kernel's occupancy on this
// device unless chip_plan_tune "bp_parts" pins it.
template <int D, bool MAPS>
int bp_parts(int tiles, int n_ang, int max_parts, int tune) {
  if (tune == 0 || max_parts < 2) return 0;
  if (tune > 0) return (tune <= max_parts && 16 * tune <= n_ang) ? tune : 0;
  static const int resident = [] {
    int per_cu = 0, dev = 0, n_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, bp_adjoint_tile_kernel<D, MAPS>, 256, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return 0;
    return per_cu * n_cu;
  }();
  if (resident <= 0 || tiles >= 4 * resident) return 0;
  int best = 0;
  double best_eff = 0.0;
  for (int s = 2; s <= max_parts && 16 * s <= n_ang; ++s) {
    const int64_t wg = static_cast<int64_t>(tiles) * s;
    const int64_t waves = (wg + resident - 1) / resident;
    const double eff = static_cast<double>(wg) / static_cast<double>(waves * resident);
    if (eff > best_eff + 1e-9) best_eff = eff, best = s;
  }
  return best;
}

__global__ void zero_kernel(float* __restrict__ v, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) v[i] = 0.f;
}

template <int D, bool MAPS>
void launch(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol, int batch, int det0,
            int angle0, int n_ang, int maps_a0, int accumulate, hipStream_t st, const BpScratch& scr) {
  const int tx = (p.n + kTileW - 1) / kTileW, ty = (p.n + kTileH - 1) / kTileH;
  const int64_t nn = static_cast<int64_t>(p.n) * p.n;
  // one slice fills only n^2 / 64 waves (4 per SIMD at 512^2); two halves of its angle range
  // (2048 workgroups at 512^2) leave a second wave of 256 on a GPU that holds 1792 at once.
  // With a fresh output and every detector in this pass, the range is split into the parts
  // that come closest to whole waves, each summed into the plan's scratch, then added in order
  const int max_parts = scr.part ? static_cast<int>(std::min<int64_t>(kBpMaxParts, scr.cap_elems / nn)) : 0;
  const int parts = (batch == 1 && !accumulate && det0 == 0 && p.n_det <= 2 && n_ang >= 2 * kAngBlk)
                        ? bp_parts<D, MAPS>(tx * ty, n_ang, max_parts, scr.tune)
                        : 0;
  if (parts) {
    // each part's sum goes to the plan's scratch; one pass adds them in part order
    dim3 grid(tx, ty, parts);
    hipLaunchKernelGGL((bp_adjoint_tile_kernel<D, MAPS>), grid, dim3(256), 0, st, sino, maps, stride, vol,
                       p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, angle0, n_ang, maps_a0, accumulate, parts,
                       p.d_bp_part);
    hipLaunchKernelGGL(sum_parts_kernel, dim3(static_cast<unsigned>((nn + 255) / 256)), dim3(256), 0, st, vol,
                       static_cast<const float*>(p.d_bp_part), parts, nn);
    return;
  }
  if (batch >= 2) {
    // four or two slices per workgroup (their candidate weights shared), an odd last slice alone
    TileSlots sl;
    sl.vol_stride = nn;
    sl.maps_stride = stride;
    sl.sino_stride = static_cast<int64_t>(p.n_det) * p.n_angles * p.n;
    // (four slices only in one-detector passes: with two detectors the eight channels'
    // windows and map loads cost more occupancy than the shared weights save)
    const int quads = D == 1 ? batch / 4 : 0, pairs = (batch - 4 * quads) / 2, b1 = 4 * quads + 2 * pairs;
    if constexpr (D == 1) {
      if (quads)
        hipLaunchKernelGGL((bp_adjoint_slots_kernel<D, MAPS, 4>), dim3(tx, ty, quads), dim3(256), 0, st, sino, maps,
                           vol, sl, 0, p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, angle0, n_ang, maps_a0,
                           accumulate, 1, static_cast<float*>(nullptr), int64_t(0));
    }
    if (pairs)
      hipLaunchKernelGGL((bp_adjoint_slots_kernel<D, MAPS, 2>), dim3(tx, ty, pairs), dim3(256), 0, st, sino, maps,
                         vol, sl, 4 * quads, p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, angle0, n_ang,
                         maps_a0, accumulate, 1, static_cast<float*>(nullptr), int64_t(0));
    if (b1 < batch)
      hipLaunchKernelGGL((bp_adjoint_slots_kernel<D, MAPS, 1>), dim3(tx, ty, 1), dim3(256), 0, st, sino, maps, vol,
                         sl, b1, p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, angle0, n_ang, maps_a0,
                         accumulate, 1, static_cast<float*>(nullptr), int64_t(0));
    return;
  }
  // slices too large for the scratch: two halves added atomically into the zeroed output
  const int split = (batch == 1 && !accumulate && det0 == 0 && p.n_det <= 2 && n_ang >= 2 * kAngBlk) ? 1 : 0;
  if (split) {
    hipLaunchKernelGGL(zero_kernel, dim3(static_cast<unsigned>((nn + 255) / 256)), dim3(256), 0, st, vol, nn);
  }
  dim3 grid(tx, ty, split ? 2 : batch);
  hipLaunchKernelGGL((bp_adjoint_tile_kernel<D, MAPS>), grid, dim3(256), 0, st, sino, maps, stride, vol, p.d_fp_geom,
                     p.d_w, p.n, p.n_angles, p.n_det, det0, angle0, n_ang, maps_a0, accumulate, split,
                     static_cast<float*>(nullptr));
}

template <int D, bool MAPS>
void launch_batch(const BatchPlan& bp, const float* sino, const float* maps, float* vol, int det0, int accumulate,
                  hipStream_t st) {
  const int64_t total = bp.h_item_start_bpa.back();
  // bp_wg_per_cu = 0: one workgroup per tile (a flat grid), else a persistent grid
  const int64_t want = bp.bp_wg_per_cu > 0 ? static_cast<int64_t>(bp.bp_wg_per_cu) * bp.n_cu : total;
  const int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(total, want)));
  hipLaunchKernelGGL((bp_adjoint_batch_kernel<D, MAPS>), dim3(grid), dim3(256), 0, st, bp.d_slices, bp.n_slices,
                     bp.d_item_start_bpa, sino, maps, vol, bp.d_geom, bp.d_w, bp.n_angles, bp.n_det, det0,
                     accumulate);
}

// Workgroups of one size group of the paired batch bp: ng slices of a size as
// groups of four (with "bp_quads", in one-detector passes), then pairs, then one
// alone. Four slices per workgroup share the candidate weights twice as far: at one
// detector 772 -> 722 us for 32 x 256^2 x 360, 8.36 -> 7.84 ms for 128 slices of
// 128..768 at 360 angles; with two detectors (eight channels) the kernel holds too
// many registers and too much LDS per workgroup and was 15 % slower.
struct SlotUnits {
  int quads, pairs, single;
};
inline SlotUnits slot_units(int ng, int quads_on) {
  SlotUnits u;
  u.quads = quads_on ? ng / 4 : 0;
  const int rest = ng - 4 * u.quads;
  u.pairs = rest / 2;
  u.single = rest % 2;
  return u;
}

template <int D, bool MAPS, int S>
int slots_per_cu() {
  static const int per_cu = [] {
    int k = 0;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&k, bp_adjoint_slots_kernel<D, MAPS, S>, 256, 0) == hipSuccess
               ? k
               : 0;
  }();
  return per_cu;
}

// Angle parts of the paired batch bp: a paired workgroup walks every angle of its tile,
// and with one workgroup per tile and slice pair too few loads are in flight to keep HBM
// busy (a read-only kernel of that shape reaches 2.1 TB/s); splitting the angle range into
// parts multiplies the workgroups. The count aims at about eight waves of what the GPU holds
// at once (at most 8 parts of at least 16 angles each: 2 x 512^2 x 720 x 2 detectors 747 ->
// 654 us with 6 parts), unless "bp_pair_parts" pins it; the count follows the kernels'
// occupancy and the CU count, and with it the grouping of the sums.
template <int D, bool MAPS>
int pair_parts(const BatchPlan& bp) {
  if (bp.bp_pair_parts >= 1) return bp.bp_pair_parts;
  bool any_quads = false;
  int64_t wg = 0;
  for (size_t g = 0; g < bp.size_plans.size(); ++g) {
    const int n = bp.size_plans[g]->p.n;
    const SlotUnits u = slot_units(bp.tile_group_start[g + 1] - bp.tile_group_start[g], D == 1 && bp.bp_quads);
    any_quads = any_quads || u.quads > 0;
    wg += static_cast<int64_t>((n + kTileW - 1) / kTileW) * ((n + kTileH - 1) / kTileH) *
          (u.quads + u.pairs + u.single);
  }
  // the residency of the widest kernel in use (four slices hold the most registers and LDS)
  int per_cu = slots_per_cu<D, MAPS, 2>();
  if constexpr (D == 1) {
    if (any_quads) per_cu = slots_per_cu<D, MAPS, 4>();
  }
  const int64_t resident = static_cast<int64_t>(per_cu) * bp.n_cu;
  if (resident <= 0 || wg <= 0) return 1;
  const int64_t want = (8 * resident + wg - 1) / wg;
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>({want, kBpMaxParts, bp.n_angles / 16})));
}

template <int D, bool MAPS>
void launch_slots(const BatchPlan& bp, const float* sino, const float* maps, float* vol, int det0, int acc,
                  int parts, float* part, hipStream_t st) {
  const int64_t plane = bp.h_vol_start.back();
  for (size_t g = 0; g < bp.size_plans.size(); ++g) {
    const Plan& sp = bp.size_plans[g]->p;
    const int g0 = bp.tile_group_start[g], ng = bp.tile_group_start[g + 1] - g0;
    const SlotUnits u = slot_units(ng, D == 1 && bp.bp_quads);
    const unsigned tx = (sp.n + kTileW - 1) / kTileW, ty = (sp.n + kTileH - 1) / kTileH;
    TileSlots sl;
    sl.offs = bp.d_tile_slots;
    if constexpr (D == 1) {
      if (u.quads)
        hipLaunchKernelGGL((bp_adjoint_slots_kernel<D, MAPS, 4>), dim3(tx, ty, u.quads * parts), dim3(256), 0, st,
                           sino, maps, vol, sl, g0, sp.d_fp_geom, sp.d_w, sp.n, sp.n_angles, sp.n_det, det0, 0,
                           sp.n_angles, 0, acc, parts, part, plane);
    }
    const int p0 = g0 + 4 * u.quads;
    if (u.pairs)
      hipLaunchKernelGGL((bp_adjoint_slots_kernel<D, MAPS, 2>), dim3(tx, ty, u.pairs * parts), dim3(256), 0, st,
                         sino, maps, vol, sl, p0, sp.d_fp_geom, sp.d_w, sp.n, sp.n_angles, sp.n_det, det0, 0,
                         sp.n_angles, 0, acc, parts, part, plane);
    if (u.single)
      hipLaunchKernelGGL((bp_adjoint_slots_kernel<D, MAPS, 1>), dim3(tx, ty, parts), dim3(256), 0, st, sino, maps,
                         vol, sl, p0 + 2 * u.pairs, sp.d_fp_geom, sp.d_w, sp.n, sp.n_angles, sp.n_det, det0, 0,
                         sp.n_angles, 0, acc, parts, part, plane);
  }
  if (parts > 1)
    hipLaunchKernelGGL(sum_parts_acc_kernel, dim3(static_cast<unsigned>((plane + 255) / 256)), dim3(256), 0, st, vol,
                       static_cast<const float*>(part), parts, plane, acc);
}

}  // namespace

int bp_pair_parts_ready(const BatchPlan& bp, bool maps, hipStream_t st, int* parts_out) {
  const bool two = bp.n_det >= 2;
  int parts = maps ? (two ? pair_parts<2, true>(bp) : pair_parts<1, true>(bp))
                   : (two ? pair_parts<2, false>(bp) : pair_parts<1, false>(bp));
  *parts_out = parts;
  if (parts <= 1) return CHIP_OK;
  const size_t need = sizeof(float) * static_cast<size_t>(parts) * bp.h_vol_start.back();
  if (bp.bp_part_bytes >= need) return CHIP_OK;
  if (stream_capturing(st)) {
    set_error("the paired batch bp needs " + std::to_string(need) +
              " bytes of part sums: run it once outside stream capture first");
    return CHIP_EINVAL;
  }
  auto& bm = const_cast<BatchPlan&>(bp);
  return ensure_buffer(bm.d_bp_part, bm.bp_part_bytes, need, st, "paired bp part sums");
}

void launch_bp_adjoint_batch(const BatchPlan& bp, const float* sino, const float* maps, float* vol, hipStream_t st,
                             int parts) {
  if (bp.bp_pairs) {
    // per slice size, four or two slices per workgroup, the angle range in `parts` parts
    for (int det0 = 0; det0 < bp.n_det; det0 += 2) {
      const bool two = bp.n_det - det0 >= 2;
      const int acc = det0 > 0 ? 1 : 0;
      if (maps) {
        if (two) launch_slots<2, true>(bp, sino, maps, vol, det0, acc, parts, bp.d_bp_part, st);
        else launch_slots<1, true>(bp, sino, maps, vol, det0, acc, parts, bp.d_bp_part, st);
      } else {
        if (two) launch_slots<2, false>(bp, sino, nullptr, vol, det0, acc, parts, bp.d_bp_part, st);
        else launch_slots<1, false>(bp, sino, nullptr, vol, det0, acc, parts, bp.d_bp_part, st);
      }
    }
    return;
  }
  for (int det0 = 0; det0 < bp.n_det; det0 += 2) {
    const bool two = bp.n_det - det0 >= 2;
    const int acc = det0 > 0 ? 1 : 0;
    if (maps) {
      if (two) launch_batch<2, true>(bp, sino, maps, vol, det0, acc, st);
      else launch_batch<1, true>(bp, sino, maps, vol, det0, acc, st);
    } else {
      if (two) launch_batch<2, false>(bp, sino, nullptr, vol, det0, acc, st);
      else launch_batch<1, false>(bp, sino, nullptr, vol, det0, acc, st);
    }
  }
}

void launch_bp_adjoint_tiles(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol,
                             int batch, int angle0, int n_ang, int maps_a0, int accumulate, hipStream_t st,
                             const BpScratch& scr) {
  // detectors in passes of two; later passes add to the first
  for (int det0 = 0; det0 < p.n_det; det0 += 2) {
    const bool two = p.n_det - det0 >= 2;
    const int acc = (det0 > 0 || accumulate) ? 1 : 0;
    if (maps) {
      if (two) launch<2, true>(p, sino, maps, stride, vol, batch, det0, angle0, n_ang, maps_a0, acc, st, scr);
      else launch<1, true>(p, sino, maps, stride, vol, batch, det0, angle0, n_ang, maps_a0, acc, st, scr);
    } else {
      if (two) launch<2, false>(p, sino, nullptr, stride, vol, batch, det0, angle0, n_ang, 0, acc, st, scr);
      else launch<1, false>(p, sino, nullptr, stride, vol, batch, det0, angle0, n_ang, 0, acc, st, scr);
    }
  }
}

bool bp_may_split(const Plan& p, int batch, int mode, int n_ang) {
  return mode == CHIP_BP_ADJOINT && batch == 1 && p.n_det <= 2 && n_ang >= 2 * kAngBlk && p.n <= kBpPartsMaxN &&
         p.bp_parts_tune != 0;
}

int bp_adjoint_parts(const Plan& p, bool maps, int max_parts) {
  const int tiles = ((p.n + kTileW - 1) / kTileW) * ((p.n + kTileH - 1) / kTileH);
  const int two = p.n_det >= 2;
  if (maps) return two ? bp_parts<2, true>(tiles, p.n_angles, max_parts, p.bp_parts_tune)
                       : bp_parts<1, true>(tiles, p.n_angles, max_parts, p.bp_parts_tune);
  return two ? bp_parts<2, false>(tiles, p.n_angles, max_parts, p.bp_parts_tune)
             : bp_parts<1, false>(tiles, p.n_angles, max_parts, p.bp_parts_tune);
}


BpPartGuard::BpPartGuard(const Plan& p, int batch, int mode, int n_ang, hipStream_t st) : p_(p), st_(st) {
  scr_.tune = p.bp_parts_tune;
  if (!bp_may_split(p, batch, mode, n_ang)) return;
  lk_ = std::unique_lock<std::mutex>(p.bp_state->mu);
  const bool capturing = stream_capturing(st);
  if (!p.d_bp_part && !capturing) {
    // no room: the call runs as two halves added atomically, which needs no scratch
    if (hipMalloc(&p.d_bp_part, sizeof(float) * kBpMaxParts * static_cast<size_t>(p.n) * p.n) != hipSuccess) {
      p.d_bp_part = nullptr;
      (void)hipGetLastError();
    }
  }
  if (!p.d_bp_part) return;
  BpPartState& s = *p.bp_state;
  if (!capturing) {
    if (s.used && s.last != st && s.done) (void)hipStreamWaitEvent(st, s.done, 0);
    if (!s.done) (void)hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    record_ = s.done != nullptr;
  }
  scr_.part = p.d_bp_part;
  scr_.cap_elems = static_cast<int64_t>(kBpMaxParts) * p.n * p.n;
}

BpPartGuard::~BpPartGuard() {
  if (!record_) return;
  BpPartState& s = *p_.bp_state;
  if (hipEventRecord(s.done, st_) == hipSuccess) {
    s.last = st_;
    s.used = true;
  }
}

}  // namespace chip
