// Whole-sinogram forward projection in image-space bands (gfx950).
//
// Discretisation: scikit-image 0.18.3 radon as called by corrct's skimage
// backend (_projector_backends.py:312-333) on vol * M[a, d]
// (projectors.py:519-525): sample (y, u) of angle a reads
//   x' = (c u + s y) + m02,  y' = (-s u + c y) + m12       (float32, no FMA)
// with an order-1 blend, 0 outside the slice; the line is the sum over y,
// times w[a, d].
//
// Decomposition: a workgroup owns one angle and U consecutive bins. Its
// threads walk the rays in the direction in which the "major" coordinate
// grows: y' where |c| >= |s| (steep rays: the image lines are rows) and x'
// otherwise (shallow rays: the lines are columns). The image is cut into
// bands of T lines along the major axis; band k holds the samples whose major
// floor lies in [T k - 1, T k + T - 1), which tap lines T k - 1 .. T k + T - 1.
// The T new lines of a band are staged in LDS as vol * M[a, d], each only over
// the minor range that the block's samples tap on it (the section of the
// block's sample parallelogram with the slab of major values [j - 1, j + 1]),
// so every map value is read from memory once; the line shared with the
// previous band is copied inside LDS. Loads for the next band are issued into
// registers before the current band's samples are gathered.
//
// Rows are staged in runs along the row; a column band is 32 columns wide so
// that each row of it is one whole 128-byte line of the vector caches (with
// 16 columns every line was fetched twice). Steep and shallow angles run as
// two launches with their own tile shapes: 256 bins x 16 rows, one thread per
// bin, and 128 bins x 32 columns, two threads per bin taking alternate
// samples.
//
// Each thread sums its samples in walk order (two threads of a bin: their
// two sums are added at the end): this differs from the oracle's sequential
// sum over y by rounding only.
#include "common.hpp"

// the sample coordinates are skimage's separate float32 products and sums:
// no contraction of a * b + c into one rounding (explicit fmaf stays fused)
#pragma clang fp contract(off)

namespace chip {
namespace {

__device__ __forceinline__ float mul_rn(float a, float b) { return a * b; }
__device__ __forceinline__ float add_rn(float a, float b) { return a + b; }

constexpr int kThreads = 256;

// Band tile: T lines, U bins. A band's tap box spans at most (U - 1) / |cos| +
// (T + 2)|tan| + 0.2 (range margins) + 1 (floors) + 1 (right tap) + 3 (float4
// alignment) minor positions, with |cos| >= 1 / sqrt(2) and |tan| <= 1 in the
// walk's frame. Row tiles keep the pitch a multiple of 4 (float4 stores);
// column tiles an odd pitch (their per-line stores then spread over the banks).
template <int T, int U, bool ROWS>
struct BandTile {
  static constexpr int kSplit = kThreads / U;  // threads per bin
  static constexpr int kBound = static_cast<int>((U - 1) * 1.41422) + 1 + T + 2 + 7;
  static constexpr int kPitch = ROWS ? ((kBound + 3) & ~3) : (kBound | 1);
  static constexpr int kTile = (T + 1) * kPitch;  // floats per detector
  // rows: kTpl threads per line, each kPer float4 runs of the line
  static constexpr int kTpl = kThreads / T;
  static constexpr int kQ = kPitch / 4;
  // columns: kG float4 groups of 4 lines, kRowsPer rows per pass of the workgroup
  static constexpr int kG = T / 4;
  static constexpr int kRowsPer = kThreads / kG;
  static constexpr int kPer = ROWS ? (kQ + kTpl - 1) / kTpl : (kPitch + kRowsPer - 1) / kRowsPer;
  static_assert(kSplit == 1 || kSplit == 2 || kSplit == 4, "1, 2 or 4 threads per bin");
  static_assert(T % 4 == 0 && kThreads % T == 0, "lines in float4 groups");
  static_assert(kPitch <= 2 * kThreads, "the halo copy covers two minor positions per thread");
};

// The block's samples form a parallelogram: in exact arithmetic
//   major = pa u + pb y + pc,  minor = qa u + qb y + qc,  u in [u0, u1], y in [0, N - 1].
// Along its boundary rays (u = u0, u1) minor = ku M + cu u + c0 and along its
// boundary rows (y = 0, N - 1) minor = ky M + cy y + c1, M the major value. For a
// slab of major values [c - h, c + h] the samples' minor values then lie in
//   [max(ku c - |ku| h + min_u(cu u) + c0, ky c - |ky| h + min_y(cy y) + c1),
//    min(ku c + |ku| h + max_u(cu u) + c0, ky c + |ky| h + max_y(cy y) + c1)]
// (each term bounds one of the two strips whose intersection is the
// parallelogram). The constants are computed once in double and used in float:
// |ku| <= 1 keeps the first pair within 2e-3 of exact for N <= 4096, widened by
// 0.02 with the float32 rounding of the samples; the second pair is used only
// while |ky| <= 64 (a nearly axis-aligned row cuts a line over fewer than 4
// bins), widened by 0.25.
struct BandRange {
  float ku, aku, u_lo, u_hi;  // slope, |slope|, offsets of the ray strip
  float ky, aky, y_lo, y_hi;  // same for the row strip (has_y)
  float mmin, mmax;           // major range of the parallelogram, widened
  int has_y;
};

__device__ BandRange band_range(float cs, float sn, float m02, float m12, bool rows, int u0, int u1, int n) {
  const double c = cs, s = sn;
  double pa, pb, pc, qa, qb, qc;
  if (rows) {
    pa = -s, pb = c, pc = m12, qa = c, qb = s, qc = m02;
  } else {
    pa = c, pb = s, pc = m02, qa = -s, qb = c, qc = m12;
  }
  const double du0 = u0, du1 = u1, y1 = n - 1;
  BandRange r;
  const double k00 = pa * du0 + pc, k10 = pa * du1 + pc, ky1 = pb * y1;
  r.mmin = static_cast<float>(fmin(fmin(k00, k10), fmin(k00 + ky1, k10 + ky1)) - 0.05);
  r.mmax = static_cast<float>(fmax(fmax(k00, k10), fmax(k00 + ky1, k10 + ky1)) + 0.05);
  const double ku = qb / pb, cu = qa - ku * pa, c0 = qc - ku * pc;
  r.ku = static_cast<float>(ku);
  r.aku = static_cast<float>(fabs(ku));
  r.u_lo = static_cast<float>(fmin(cu * du0, cu * du1) + c0 - 0.02);
  r.u_hi = static_cast<float>(fmax(cu * du0, cu * du1) + c0 + 0.02);
  r.has_y = fabs(pa) * 64.0 >= fabs(qa);
  if (r.has_y) {
    const double ky = qa / pa, cy = qb - ky * pb, c1 = qc - ky * pc;
    r.ky = static_cast<float>(ky);
    r.aky = static_cast<float>(fabs(ky));
    r.y_lo = static_cast<float>(fmin(0.0, cy * y1) + c1 - 0.25);
    r.y_hi = static_cast<float>(fmax(0.0, cy * y1) + c1 + 0.25);
  } else {
    r.ky = r.aky = r.y_lo = r.y_hi = 0.f;
  }
  return r;
}

// minor taps [lo, hi] (inclusive) of the samples with major in [c - h, c + h], clamped to
// [-1, N] in minor as the gather clamps them; an empty range (lo > hi, far outside every
// index) if there are none
__device__ __forceinline__ int2 slab_taps(const BandRange& r, float c, float h, int n) {
  const int2 none = make_int2(1 << 30, -(1 << 30));
  if (c + h < r.mmin || c - h > r.mmax) return none;
  float lo = fmaf(r.ku, c, r.u_lo) - r.aku * h;
  float hi = fmaf(r.ku, c, r.u_hi) + r.aku * h;
  if (r.has_y) {
    lo = fmaxf(lo, fmaf(r.ky, c, r.y_lo) - r.aky * h);
    hi = fminf(hi, fmaf(r.ky, c, r.y_hi) + r.aky * h);
  }
  if (lo > hi) return none;
  const float fn = static_cast<float>(n);
  return make_int2(static_cast<int>(floorf(fminf(fmaxf(lo, -1.f), fn))),
                   static_cast<int>(floorf(fminf(fmaxf(hi, -1.f), fn))) + 1);
}

// a line j is tapped by the samples with major floor j - 1 or j (major in [j - 1, j + 1))
constexpr float kLineHalf = 1.02f;

struct BandItem {
  const float* v;     // slice [n][n]
  const float* mb;    // maps of (angle, det0): D planes of n*n, or NULL
  float* out;         // sinogram line of (det0, angle); detector d at out + d * det_stride
  size_t det_stride;  // elements between detectors' lines (A * n)
  const float* wa;    // weights of (angle, det0 .. det0 + D)
  float4 gm;          // (cos, sin, m02, m12) of the angle
  int n;
  int u0;
};

// ROWS: lines are image rows (major = y'); otherwise image columns (major = x').
// LDS: tile [T + 1][kPitch][D] (line slot, minor offset from the band origin, detector).
template <int T, int U, bool ROWS, int D, bool MAPS>
__device__ __forceinline__ void fp_band_block(const BandItem& it, float* lds) {
  static_assert(D == 1 || D == 2, "one or two detectors per pass");
  using BT = BandTile<T, U, ROWS>;
  constexpr int P = BT::kPitch, kPer = BT::kPer, S = BT::kSplit;
  const int n = it.n, u0 = it.u0, tid = threadIdx.x;
  const int u = u0 + tid / S;
  const int u1 = min(u0 + U, n) - 1;
  const float cs = it.gm.x, sn = it.gm.y, m02 = it.gm.z, m12 = it.gm.w;
  const float fn = static_cast<float>(n);
  const size_t nn = static_cast<size_t>(n) * n;
  const float* v = it.v;
  const float* mb = it.mb;

  const BandRange rg = band_range(cs, sn, m02, m12, ROWS, u0, u1, n);
  const int nb = (n + T) / T;  // bands over major floors -1 .. n - 1
  const int kf = max(0, static_cast<int>(floorf((rg.mmin + 1.f) / T)));
  const int kl = min(nb - 1, static_cast<int>(floorf((rg.mmax + 1.f) / T)));
  // origin of band k's box: below every line's taps (0.05 more margin), on a float4
  auto box_origin = [&](int k) {
    const int2 r = slab_taps(rg, static_cast<float>(T * k - 1) + 0.5f * T, 0.5f * T + kLineHalf + 0.05f, n);
    return (r.x <= r.y ? r.x : -1) & ~3;
  };

  // the walk: t = 0, 1, ... along the ray in the direction of growing major
  const float slope = ROWS ? cs : sn;
  const int dir = slope > 0.f ? 1 : -1;
  const float uf = static_cast<float>(u);
  const float xu = mul_rn(cs, uf), yu = mul_rn(-sn, uf);
  auto major_at = [&](int t) {
    const float yf = static_cast<float>(dir > 0 ? t : n - 1 - t);
    return ROWS ? add_rn(add_rn(yu, mul_rn(cs, yf)), m12) : add_rn(add_rn(xu, mul_rn(sn, yf)), m02);
  };
  // first sample with major >= -1 (major is non-decreasing along the walk); the S threads
  // of a bin take every S-th sample from there. A sample is past the ray's end when its
  // major exceeds the last sample's by more than 0.35 (consecutive samples differ by
  // |slope| >= 0.7), so no index is kept: the band's bound is lowered to that cap.
  int t = n;
  float cap = -1e30f;
  if (u <= u1) {
    int lo = 0, hi = n;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (major_at(mid) >= -1.f) hi = mid;
      else lo = mid + 1;
    }
    t = lo + tid % S;
    cap = major_at(n - 1) + 0.35f;
  }
  float yf = static_cast<float>(dir > 0 ? t : n - 1 - t);
  const float ystep = static_cast<float>(dir * S);

  // staging: rows: thread -> line slot 1 + tid / kTpl, float4 runs q0 + kTpl i of it;
  // columns: thread -> line group (4 columns) tid % kG, rows q0 + kRowsPer i. A slot
  // that is needed but lies outside the slice holds zeros.
  float4 fr[kPer];
  float4 mr[D][kPer];
  unsigned need_mask = 0u;
  const int q0 = ROWS ? tid % BT::kTpl : tid / BT::kG;
  const int line0 = ROWS ? 1 + tid / BT::kTpl : 1 + 4 * (tid % BT::kG);
  auto issue = [&](int k, int o) {
    need_mask = 0u;
    // needed slots m in [n_lo, n_hi], loaded from memory in [l_lo, l_hi]
    int n_lo, n_hi, l_lo, l_hi, base;
    if (ROWS) {
      const int j = T * k - 1 + line0;
      const int2 r = slab_taps(rg, static_cast<float>(j), kLineHalf, n);
      // float4 run m covers columns o + 4m .. o + 4m + 3 (o and N are multiples of 4)
      n_lo = (r.x - o) >> 2;
      n_hi = min((r.y - o) >> 2, BT::kQ - 1);
      l_lo = max(n_lo, (-o) >> 2);
      l_hi = (j >= 0 && j < n) ? min(n_hi, ((n - o) >> 2) - 1) : -1;
      base = j * n + o;
    } else {
      const int x = T * k + line0 - 1;  // columns x .. x + 3
      const int2 r = slab_taps(rg, static_cast<float>(x) + 1.5f, kLineHalf + 1.5f, n);
      n_lo = max(r.x - o, 0);
      n_hi = min(r.y - o, P - 1);
      l_lo = max(n_lo, -o);
      l_hi = (x + 3 < n) ? min(n_hi, n - 1 - o) : -1;
      base = o * n + x;
    }
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      const int m = q0 + (ROWS ? BT::kTpl : BT::kRowsPer) * i;
      if (m < n_lo || m > n_hi) continue;
      need_mask |= 1u << i;
      if (m >= l_lo && m <= l_hi) {
        // uniform base + 32-bit byte offset: scalar-base addressing
        const unsigned boff = static_cast<unsigned>(base + (ROWS ? 4 * m : m * n)) * 4u;
        fr[i] = *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(v) + boff);
#pragma unroll
        for (int d = 0; d < D; ++d)
          mr[d][i] = MAPS ? *reinterpret_cast<const float4*>(reinterpret_cast<const char*>(mb + d * nn) + boff)
                          : make_float4(1.f, 1.f, 1.f, 1.f);
      } else {
        fr[i] = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int d = 0; d < D; ++d) mr[d][i] = make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int i = 0; i < kPer; ++i) {
      if (!(need_mask & (1u << i))) continue;
      float p[D][4];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        p[d][0] = MAPS ? fr[i].x * mr[d][i].x : fr[i].x;
        p[d][1] = MAPS ? fr[i].y * mr[d][i].y : fr[i].y;
        p[d][2] = MAPS ? fr[i].z * mr[d][i].z : fr[i].z;
        p[d][3] = MAPS ? fr[i].w * mr[d][i].w : fr[i].w;
      }
      if (ROWS) {
        // four consecutive minor positions of one line: [col][D] interleaved
        const int o = (line0 * P + 4 * (q0 + BT::kTpl * i)) * D;
#pragma unroll
        for (int h = 0; h < D; ++h) {
          float4 w4;
          w4.x = p[(4 * h + 0) % D][(4 * h + 0) / D];
          w4.y = p[(4 * h + 1) % D][(4 * h + 1) / D];
          w4.z = p[(4 * h + 2) % D][(4 * h + 2) / D];
          w4.w = p[(4 * h + 3) % D][(4 * h + 3) / D];
          *reinterpret_cast<float4*>(lds + o + 4 * h) = w4;
        }
      } else {
        // one minor position (row) of four lines (columns)
        const int m = q0 + BT::kRowsPer * i;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int o = ((line0 + e) * P + m) * D;
          if (D == 2) *reinterpret_cast<float2*>(lds + o) = make_float2(p[0][e], p[D - 1][e]);
          else lds[o] = p[0][e];
        }
      }
    }
  };

  // the first band's line T kf - 1 (slot 0) is read from memory once; later bands copy it from slot T
  auto load_line = [&](int j, int o, int2 r, float (&hv)[2][D]) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int mi = o + tid + kThreads * h;
#pragma unroll
      for (int d = 0; d < D; ++d) hv[h][d] = 0.f;
      if (tid + kThreads * h < P && mi >= r.x && mi <= r.y && j >= 0 && j < n && mi >= 0 && mi < n) {
        const size_t px = ROWS ? static_cast<size_t>(j) * n + mi : static_cast<size_t>(mi) * n + j;
        const float vv = v[px];
#pragma unroll
        for (int d = 0; d < D; ++d) hv[h][d] = MAPS ? vv * mb[d * nn + px] : vv;
      }
    }
  };

  float acc[D];
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = 0.f;

  if (kf <= kl) {
    int o_cur = box_origin(kf);
    int2 r_halo = slab_taps(rg, static_cast<float>(T * kf - 1), kLineHalf, n);
    float hv[2][D];
    load_line(T * kf - 1, o_cur, r_halo, hv);
    issue(kf, o_cur);
    int o_prev = o_cur;
    for (int k = kf; k <= kl; ++k) {
      if (k > kf) {
        // slot T of band k - 1 is line T k - 1: keep it for slot 0 (band origins differ)
        r_halo = slab_taps(rg, static_cast<float>(T * k - 1), kLineHalf, n);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int m = tid + kThreads * h;
          const int mi = o_cur + m;
#pragma unroll
          for (int d = 0; d < D; ++d) hv[h][d] = 0.f;
          if (m < P && mi >= r_halo.x && mi <= r_halo.y) {
            const int src = (T * P + (mi - o_prev)) * D;
#pragma unroll
            for (int d = 0; d < D; ++d) hv[h][d] = lds[src + d];
          }
        }
      }
      __syncthreads();  // band k - 1's gathers and the copy reads are done with the tile
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int m = tid + kThreads * h;
        const int mi = o_cur + m;
        if (m < P && mi >= r_halo.x && mi <= r_halo.y) {
#pragma unroll
          for (int d = 0; d < D; ++d) lds[m * D + d] = hv[h][d];
        }
      }
      stage();
      __syncthreads();
      int o_next = o_cur;
      if (k < kl) {
        o_next = box_origin(k + 1);
        issue(k + 1, o_next);
      }

      // gather the samples of band k: major floor in [T k - 1, T k + T - 1)
      const float band_hi = fminf(static_cast<float>(T * k + T - 1), cap);
      const int ibase = (T * k - 1) * P + o_cur;
      // one sample in two steps: fetch = skimage inverse map and the LDS reads of its four
      // taps, blend = the order-1 blend. Two samples are fetched and blended per step and
      // kept only while inside the band, so their LDS reads are in flight together (the
      // sample past the band reads clamped taps inside the tile and is dropped).
      struct Tap {
        float maj, dc, dr;
        float a0[D], a1[D], b0[D], b1[D];  // (M0, m0), (M0, m0 + 1), (M0 + 1, m0), (M0 + 1, m0 + 1)
      };
      auto fetch = [&](float yv, Tap& tp) {
        const float xs = add_rn(add_rn(xu, mul_rn(sn, yv)), m02);
        const float ys = add_rn(add_rn(yu, mul_rn(cs, yv)), m12);
        tp.maj = ROWS ? ys : xs;
        const float mnr = fminf(fmaxf(ROWS ? xs : ys, -1.f), fn);
        const float M0 = floorf(tp.maj), m0 = floorf(mnr);
        tp.dc = ROWS ? mnr - m0 : tp.maj - M0;
        tp.dr = ROWS ? tp.maj - M0 : mnr - m0;
        const int li = min(max(static_cast<int>(M0) * P + static_cast<int>(m0) - ibase, 0), T * P - 2);
        if (D == 2) {
          const float2* tl = reinterpret_cast<const float2*>(lds) + li;
          const float2 p0 = tl[0], p1 = tl[1], q0v = tl[P], q1v = tl[P + 1];
          tp.a0[0] = p0.x, tp.a0[D - 1] = p0.y, tp.a1[0] = p1.x, tp.a1[D - 1] = p1.y;
          tp.b0[0] = q0v.x, tp.b0[D - 1] = q0v.y, tp.b1[0] = q1v.x, tp.b1[D - 1] = q1v.y;
        } else {
          const float* tl = lds + li;
          tp.a0[0] = tl[0], tp.a1[0] = tl[1], tp.b0[0] = tl[P], tp.b1[0] = tl[P + 1];
        }
      };
      auto blend = [&](const Tap& tp, float (&f)[D]) {
        // pixel taps t_rc: rows r0, r0 + 1 and columns c0, c0 + 1
#pragma unroll
        for (int d = 0; d < D; ++d) {
          const float t00 = tp.a0[d];
          const float t01 = ROWS ? tp.a1[d] : tp.b0[d];
          const float t10 = ROWS ? tp.b0[d] : tp.a1[d];
          const float t11 = tp.b1[d];
          const float top = fmaf(tp.dc, t01 - t00, t00);
          const float bot = fmaf(tp.dc, t11 - t10, t10);
          f[d] = fmaf(tp.dr, bot - top, top);
        }
      };
      for (;;) {
        Tap ta, tb;
        fetch(yf, ta);
        fetch(yf + ystep, tb);
        float fa[D], fb[D];
        blend(ta, fa);
        blend(tb, fb);
        const bool oka = ta.maj < band_hi, okb = tb.maj < band_hi;  // okb implies oka
#pragma unroll
        for (int d = 0; d < D; ++d) {
          acc[d] = oka ? acc[d] + fa[d] : acc[d];
          acc[d] = okb ? acc[d] + fb[d] : acc[d];
        }
        yf += (oka ? ystep : 0.f) + (okb ? ystep : 0.f);
        if (!okb) break;
      }
      o_prev = o_cur;
      o_cur = o_next;
    }
  }

  if (S >= 2) {
    // the S threads of a bin are neighbouring lanes
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] += __shfl_xor(acc[d], 1);
  }
  if (S == 4) {
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] += __shfl_xor(acc[d], 2);
  }
  if (tid % S == 0 && u <= u1) {
#pragma unroll
    for (int d = 0; d < D; ++d) it.out[d * it.det_stride + u] = acc[d] * it.wa[d];
  }
}

template <int T, int U, bool ROWS, int D, bool MAPS, int WAVES>
__global__ __launch_bounds__(kThreads, WAVES) void fp_band_kernel(
    const float* __restrict__ vol, const float* __restrict__ maps, int64_t maps_slice_stride,
    float* __restrict__ sino, const float4* __restrict__ geom, const float* __restrict__ w, int n, int n_angles,
    int n_det_total, int det0, int n_ublk, int angle0, int maps_a0) {
  extern __shared__ float lds[];
  // consecutive workgroups go to different XCDs, so every XCD gets both orientations' angles
  const int lb = blockIdx.x;
  const int la = lb / n_ublk;
  const int a = angle0 + la;
  const float4 gm = geom[a];
  if ((fabsf(gm.x) >= fabsf(gm.y)) != ROWS) return;  // the other orientation's launch has this angle
  const int b = blockIdx.y;
  const size_t nn = static_cast<size_t>(n) * n;
  BandItem it;
  it.n = n;
  it.u0 = (lb - la * n_ublk) * U;
  it.gm = gm;
  it.v = vol + static_cast<size_t>(b) * nn;
  it.mb = MAPS ? maps + static_cast<size_t>(b) * maps_slice_stride +
                     (static_cast<size_t>(a - maps_a0) * n_det_total + det0) * nn
               : nullptr;
  it.out = sino + ((static_cast<size_t>(b) * n_det_total + det0) * n_angles + a) * n;
  it.det_stride = static_cast<size_t>(n_angles) * n;
  it.wa = w + static_cast<size_t>(a) * n_det_total + det0;
  fp_band_block<T, U, ROWS, D, MAPS>(it, lds);
}

template <int T, int U, bool ROWS, int D, bool MAPS, int WAVES>
void launch_t(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch, int det0,
              int angle0, int n_ang, int maps_a0, hipStream_t st) {
  using BT = BandTile<T, U, ROWS>;
  const int n_ublk = (p.n + U - 1) / U;
  dim3 grid(n_ublk * n_ang, batch);
  const size_t shmem = sizeof(float) * D * BT::kTile;
  hipLaunchKernelGGL((fp_band_kernel<T, U, ROWS, D, MAPS, WAVES>), grid, dim3(kThreads), shmem, st, vol, maps,
                     stride, sino, p.d_fp_geom, p.d_w, p.n, p.n_angles, p.n_det, det0, n_ublk, angle0, maps_a0);
}

// steep angles: 256 bins x 16 rows (53 KiB for two detectors, 3 workgroups per CU);
// shallow angles: 128 bins x 32 columns (58 KiB, 2 workgroups per CU)
template <int D, bool MAPS>
void launch(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch, int det0,
            int angle0, int n_ang, int maps_a0, hipStream_t st) {
  launch_t<16, 256, true, D, MAPS, 3>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, maps_a0, st);
  launch_t<32, 128, false, D, MAPS, 2>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, maps_a0, st);
}

}  // namespace

void launch_fp_band(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch,
                    int det0, int nd, int angle0, int n_ang, int maps_a0, hipStream_t st) {
  if (maps) {
    if (nd == 2) launch<2, true>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, maps_a0, st);
    else launch<1, true>(p, vol, maps, stride, sino, batch, det0, angle0, n_ang, maps_a0, st);
  } else {
    if (nd == 2) launch<2, false>(p, vol, nullptr, stride, sino, batch, det0, angle0, n_ang, 0, st);
    else launch<1, false>(p, vol, nullptr, stride, sino, batch, det0, angle0, n_ang, 0, st);
  }
}

}  // namespace chip
