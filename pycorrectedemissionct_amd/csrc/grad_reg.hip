// Gradient-family regularizers of libcorrct_hip (gfx950): the dual update,
// dual proximal and primal contribution of Regularizer_Grad and its
// subclasses TV1D / TV2D / TV3D, HubTV2D / HubTV3D, smooth1D / 2D / 3D, VTV and
// TNV (regularizers.py:134-215, 218-433, 1324-1440) for gradients over the last
// 1..3 axes.
//
// Layout: the primal is [C][B][S] with S the product of the gradient axes'
// sizes, C the channels a VTV / TNV prox couples (1 for TV) and B the
// remaining axes, which stay independent; the dual is [nd][C][B][S], the
// reference's TransformGradient.adj_shape (operators.py:781-793). One thread
// owns one (b, s) and all C * nd dual entries there, so each prox is local
// and every dual entry is read and written once (coalesced along s).
//
// Float32 order follows the reference's numpy expressions:
//   dual   q = q + 0.5 * (x[i+1] - x[i])        (0 past the last sample)
//   l12    q = (q / max(|q|_2 over nd, w)) * w   (data_terms.py:549-555, 607-608)
//   l1     q = (q / max(|q|, w)) * w
//   l2     q = q * w,  w = 1 / (1 + 0.5)             (data_terms.py:294-304, no data)
//   Huber  q = q * w,  w = 1 / (1 + 0.5 huber_size); q = q / max(1, |q|_2 over nd)
//          (data_terms.py:482-501 with l2_axis 0); without an l2 axis each
//          component is divided by max(1, |q|) (data_terms.py:497-498)
//   VTV    q = (q / max(dn, w)) * w with dn the point-wise derivative norm
//          (1, 2, inf) followed by the channel norm (1, 2, inf); the dual
//          norms of inf are 1, as in regularizers.py:1395-1411
//   TNV    per (b, s) the C x nd matrix M = U diag(s) V^T becomes
//          U diag(s / max(s, 1)) V^T (data_terms.py:781-820 with the l1 dual
//          prox of weight 1); computed as M V diag(1 / max(s, 1)) V^T from a
//          float64 Jacobi eigen-decomposition of M^T M (nd <= 3)
//   primal upd = upd +- scale * (-div q), div with the first difference 0
//          (operators.py:820-844), summed over the axes in order.
// Padding (np.pad modes of TransformGradient): only the gradient's last
// difference and the divergence's first one depend on it (grad_last,
// div_first); edge is the default.
#include "common.hpp"

namespace chip {
namespace {

constexpr int kGrBlock = 256;
constexpr int kGrMaxChan = 16;

enum GradMode { kGrL12 = 0, kGrL1 = 1, kGrVTV = 2, kGrTNV = 3, kGrL2 = 4, kGrHuber = 5, kGrHuberEl = 6 };

// np.pad modes of TransformGradient (operators.py:796-835)
enum GradPad { kPadEdge = 0, kPadConstant = 1, kPadWrap = 2, kPadReflect = 3, kPadSymmetric = 4 };

struct GradGeom {
  int C;        // coupled channels
  int nd;       // gradient axes (1..3)
  int64_t B;    // independent batch
  int64_t S;    // samples per (c, b)
  int dim[3];   // sizes of the gradient axes, slowest first
  int64_t stride[3];
  int pad;      // GradPad
};

// forward difference at the last sample of an axis: the padded value minus x0
// (edge and symmetric repeat x0, constant pads 0, wrap the first sample,
// reflect the one before the last); p points at x0, the axis has `dim` samples
__device__ __forceinline__ float grad_last(const float* p, float x0, int dim, int64_t stride, int pad) {
  switch (pad) {
    case kPadConstant: return -x0;
    case kPadWrap: return p[-static_cast<int64_t>(dim - 1) * stride] - x0;
    case kPadReflect: return dim > 1 ? p[-stride] - x0 : 0.f;
    default: return 0.f;
  }
}

// backward difference of the divergence at the first sample: q0 minus the padded value
__device__ __forceinline__ float div_first(const float* p, float q0, int dim, int64_t stride, int pad) {
  switch (pad) {
    case kPadConstant: return q0;
    case kPadWrap: return q0 - p[static_cast<int64_t>(dim - 1) * stride];
    case kPadReflect: return dim > 1 ? q0 - p[stride] : 0.f;
    default: return 0.f;
  }
}

inline dim3 gr_grid(int64_t n) {
  int64_t blocks = (n + kGrBlock - 1) / kGrBlock;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  return dim3(static_cast<unsigned>(blocks));
}

// position of sample s along each gradient axis (the axes may be any of the
// block's axes, in any order)
__device__ __forceinline__ void sample_index(const GradGeom& g, int64_t s, int idx[3]) {
  for (int k = 0; k < g.nd; ++k) idx[k] = static_cast<int>((s / g.stride[k]) % g.dim[k]);
}

// Cyclic Jacobi on the symmetric nd x nd matrix a (overwritten by its
// eigenvalues on the diagonal); v receives the eigenvectors as columns.
__device__ void jacobi_sym(double a[3][3], double v[3][3], int nd) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) v[i][j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 12; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < nd; ++p) {
      diag += a[p][p] * a[p][p];
      for (int q = p + 1; q < nd; ++q) off += a[p][q] * a[p][q];
    }
    if (off <= 1e-30 * diag || off == 0.0) break;
    for (int p = 0; p < nd; ++p) {
      for (int q = p + 1; q < nd; ++q) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0);
        const double s = t * c;
        // A <- J^T A J with J = [[c, s], [-s, c]] on (p, q); V <- V J
        for (int k = 0; k < nd; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < nd; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
        for (int k = 0; k < nd; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
    }
  }
}

__device__ __forceinline__ float pnorm_acc(float acc, float v, int ord) {
  // running point-wise norm: ord 2 accumulates squares, 1 and 0 (inf's dual, 1) absolutes
  return ord == 2 ? acc + v * v : acc + fabsf(v);
}

// ND gradient axes and at most CM channels, both compile-time so the C * ND
// dual values stay in registers; g.C <= CM channels are live.
template <int ND, int CM>
__global__ __launch_bounds__(kGrBlock) void grad_dual_kernel(float* __restrict__ q, const float* __restrict__ x,
                                                             GradGeom g, int mode, float w,
                                                             const float* __restrict__ wt, int64_t wt_period,
                                                             int der, int chan, const float* __restrict__ qref) {
  const int64_t total = g.B * g.S;
  const int64_t cs = g.B * g.S;  // channel stride of x and of each q component
  const int64_t qs = static_cast<int64_t>(g.C) * cs;  // component stride of q
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t b = t / g.S;
    const int64_t s = t - b * g.S;
    int idx[3];
    sample_index(g, s, idx);
    const int64_t off = b * g.S + s;
    float v[ND][CM];
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      if (c >= g.C) break;
      const float* xc = x + c * cs + off;
      const float x0 = xc[0];
#pragma unroll
      for (int k = 0; k < ND; ++k) {
        const float gk = (idx[k] + 1 < g.dim[k]) ? xc[g.stride[k]] - x0
                                                 : grad_last(xc, x0, g.dim[k], g.stride[k], g.pad);
        v[k][c] = q[k * qs + c * cs + off] + 0.5f * gk;
      }
    }
    // the weight of channel c at this sample: an array broadcast over the primal (period
    // dividing its size), else the scalar
    auto wch = [&](int c) -> float { return wt ? wt[(c * cs + off) % wt_period] : w; };
    if (mode == kGrL12) {
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (c >= g.C) break;
        const float wc = wch(c);
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < ND; ++k) acc = acc + v[k][c] * v[k][c];
        const float den = fmaxf(sqrtf(acc), wc);
#pragma unroll
        for (int k = 0; k < ND; ++k) v[k][c] = (v[k][c] / den) * wc;
      }
    } else if (mode == kGrL2) {
      // DataFidelity_l2 without data: q = q * (1 / (1 + sigma)), w holding that factor
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (c >= g.C) break;
#pragma unroll
        for (int k = 0; k < ND; ++k) v[k][c] = v[k][c] * w;
      }
    } else if (mode == kGrHuber) {
      // DataFidelity_Huber over l2_axis 0: q = q * w (w = 1 / (1 + sigma huber_size)),
      // then q / max(1, |q|_2 over the nd components)
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (c >= g.C) break;
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < ND; ++k) {
          v[k][c] = v[k][c] * w;
          acc = acc + v[k][c] * v[k][c];
        }
        const float den = fmaxf(1.f, sqrtf(acc));
#pragma unroll
        for (int k = 0; k < ND; ++k) v[k][c] = v[k][c] / den;
      }
    } else if (mode == kGrHuberEl) {
      // DataFidelity_Huber without an l2 axis: q = q * w, then q / max(1, |q|) per component
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (c >= g.C) break;
#pragma unroll
        for (int k = 0; k < ND; ++k) {
          const float t = v[k][c] * w;
          v[k][c] = t / fmaxf(1.f, fabsf(t));
        }
      }
    } else if (mode == kGrL1) {
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (c >= g.C) break;
        const float wc = wch(c);
#pragma unroll
        for (int k = 0; k < ND; ++k) v[k][c] = (v[k][c] / fmaxf(fabsf(v[k][c]), wc)) * wc;
      }
    } else if (mode == kGrVTV) {
      // qref: the reference volume's fixed dual, one more channel after the C channels in
      // every norm, never written (regularizers.py:1410-1411, 1434-1435)
      float r[ND];
#pragma unroll
      for (int k = 0; k < ND; ++k) r[k] = qref ? qref[k * total + off] : 0.f;
      if (der == 1) {
        // gn = |q| per (k, c); the channel norm runs over c for each k
#pragma unroll
        for (int k = 0; k < ND; ++k) {
          float acc = 0.f;
#pragma unroll
          for (int c = 0; c < CM; ++c)
            if (c < g.C) acc = pnorm_acc(acc, fabsf(v[k][c]), chan);
          if (qref) acc = pnorm_acc(acc, fabsf(r[k]), chan);
          const float dn_k = chan == 2 ? sqrtf(acc) : acc;
#pragma unroll
          for (int c = 0; c < CM; ++c) {
            if (c >= g.C) break;
            const float dn = chan == 1 ? fabsf(v[k][c]) : dn_k;
            const float wc = wch(c);
            v[k][c] = (v[k][c] / fmaxf(dn, wc)) * wc;
          }
        }
      } else {
        float gn[CM];
#pragma unroll
        for (int c = 0; c < CM; ++c) {
          float acc = 0.f;
#pragma unroll
          for (int k = 0; k < ND; ++k) acc = pnorm_acc(acc, c < g.C ? v[k][c] : 0.f, der);
          gn[c] = der == 2 ? sqrtf(acc) : acc;
        }
        float acc = 0.f;
#pragma unroll
        for (int c = 0; c < CM; ++c)
          if (c < g.C) acc = pnorm_acc(acc, gn[c], chan);
        if (qref) {
          float ar = 0.f;
#pragma unroll
          for (int k = 0; k < ND; ++k) ar = pnorm_acc(ar, r[k], der);
          acc = pnorm_acc(acc, der == 2 ? sqrtf(ar) : ar, chan);
        }
        const float dn_all = chan == 2 ? sqrtf(acc) : acc;
#pragma unroll
        for (int c = 0; c < CM; ++c) {
          if (c >= g.C) break;
          const float wc = wch(c);
          const float den = fmaxf(chan == 1 ? gn[c] : dn_all, wc);
#pragma unroll
          for (int k = 0; k < ND; ++k) v[k][c] = (v[k][c] / den) * wc;
        }
      }
    } else {  // kGrTNV
      double a[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, ev[3][3];
#pragma unroll
      for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int k = j; k < ND; ++k) {
          double acc = 0.0;
#pragma unroll
          for (int c = 0; c < CM; ++c)
            if (c < g.C) acc += static_cast<double>(v[j][c]) * v[k][c];
          a[j][k] = acc;
          a[k][j] = acc;
        }
      jacobi_sym(a, ev, ND);
      double h[3];
#pragma unroll
      for (int i = 0; i < ND; ++i) h[i] = 1.0 / fmax(sqrt(fmax(a[i][i], 0.0)), 1.0);
      double hm[3][3];
#pragma unroll
      for (int j = 0; j < ND; ++j)
#pragma unroll
        for (int k = 0; k < ND; ++k) {
          double acc = 0.0;
#pragma unroll
          for (int i = 0; i < ND; ++i) acc += ev[j][i] * h[i] * ev[k][i];
          hm[j][k] = acc;
        }
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        if (c >= g.C) break;
        double m[ND];
#pragma unroll
        for (int j = 0; j < ND; ++j) m[j] = v[j][c];
#pragma unroll
        for (int k = 0; k < ND; ++k) {
          double acc = 0.0;
#pragma unroll
          for (int j = 0; j < ND; ++j) acc += m[j] * hm[j][k];
          v[k][c] = static_cast<float>(acc);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      if (c >= g.C) break;
#pragma unroll
      for (int k = 0; k < ND; ++k) q[k * qs + c * cs + off] = v[k][c];
    }
  }
}

template <int ND>
void launch_grad_dual(float* q, const float* x, const GradGeom& g, int mode, float w, const float* wt,
                      int64_t wt_period, int der, int chan, const float* qref, hipStream_t st) {
  const dim3 grid = gr_grid(g.B * g.S);
  if (g.C == 1)
    hipLaunchKernelGGL((grad_dual_kernel<ND, 1>), grid, dim3(kGrBlock), 0, st, q, x, g, mode, w, wt, wt_period, der,
                       chan, qref);
  else if (g.C <= 4)
    hipLaunchKernelGGL((grad_dual_kernel<ND, 4>), grid, dim3(kGrBlock), 0, st, q, x, g, mode, w, wt, wt_period, der,
                       chan, qref);
  else
    hipLaunchKernelGGL((grad_dual_kernel<ND, kGrMaxChan>), grid, dim3(kGrBlock), 0, st, q, x, g, mode, w, wt,
                       wt_period, der, chan, qref);
}

__global__ __launch_bounds__(kGrBlock) void grad_primal_kernel(float* __restrict__ upd, const float* __restrict__ q,
                                                               GradGeom g, float scale,
                                                               const float* __restrict__ scales,
                                                               int64_t scales_period, int apply_scale,
                                                               int subtract, const float* __restrict__ mask,
                                                               int64_t mask_period) {
  const int64_t total = static_cast<int64_t>(g.C) * g.B * g.S;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t s = e % g.S;
    int idx[3];
    sample_index(g, s, idx);
    float d = 0.f;
    for (int k = 0; k < g.nd; ++k) {
      const float* qk = q + k * total + e;
      const float dk = (idx[k] > 0) ? qk[0] - qk[-g.stride[k]] : div_first(qk, qk[0], g.dim[k], g.stride[k], g.pad);
      d = (k == 0) ? dk : d + dk;
    }
    float t = -d;
    if (mask) t = t * mask[e % mask_period];
    if (apply_scale) t = t * (scales ? scales[e % scales_period] : scale);
    upd[e] = subtract ? upd[e] - t : upd[e] + t;
  }
}

int make_geom(GradGeom& g, int n_chan, int64_t n_batch, int nd, const int* dims, const int64_t* strides, int pad) {
  CHIP_CHECK_ARG(nd >= 1 && nd <= 3, "nd must be 1, 2 or 3");
  CHIP_CHECK_ARG(pad >= kPadEdge && pad <= kPadSymmetric, "pad_mode must be 0..4 (edge, constant, wrap, reflect, "
                 "symmetric)");
  g.pad = pad;
  CHIP_CHECK_ARG(dims != nullptr, "NULL dims");
  CHIP_CHECK_ARG(n_chan >= 1 && n_chan <= kGrMaxChan, "n_chan must be in [1, 16]");
  CHIP_CHECK_ARG(n_batch >= 1, "n_batch must be >= 1");
  g.C = n_chan;
  g.nd = nd;
  g.B = n_batch;
  g.S = 1;
  for (int k = 0; k < 3; ++k) {
    g.dim[k] = 1;
    g.stride[k] = 0;
  }
  for (int k = 0; k < nd; ++k) {
    CHIP_CHECK_ARG(dims[k] >= 1, "gradient axis sizes must be >= 1");
    g.dim[k] = dims[k];
    g.S *= dims[k];
  }
  if (strides) {  // the gradient axes at these strides inside one block of n_batch samples
    g.B = 1;
    g.S = n_batch;
    for (int k = 0; k < nd; ++k) {
      CHIP_CHECK_ARG(strides[k] >= 1 && strides[k] * dims[k] <= n_batch && n_batch % (strides[k] * dims[k]) == 0,
                     "each gradient axis must tile the block: stride * size dividing the block size");
      for (int j = 0; j < k; ++j) CHIP_CHECK_ARG(strides[j] != strides[k], "gradient axes must be distinct");
      g.stride[k] = strides[k];
    }
    return CHIP_OK;
  }
  int64_t st = 1;
  for (int k = nd - 1; k >= 0; --k) {
    g.stride[k] = st;
    st *= g.dim[k];
  }
  return CHIP_OK;
}

}  // namespace
}  // namespace chip

using namespace chip;

extern "C" {

int chip_grad_dual(float* q, const float* x, int n_chan, int64_t n_batch, int nd, const int* dims,
                   const int64_t* strides, int pad_mode, int mode, float weight, const float* weights, int64_t weights_period, int der_norm,
                   int chan_norm, const float* qref, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(q != nullptr && x != nullptr, "NULL argument");
  CHIP_CHECK_ARG(mode >= kGrL12 && mode <= kGrHuberEl, "unknown mode");
  CHIP_CHECK_ARG(mode != kGrVTV || ((der_norm == 0 || der_norm == 1 || der_norm == 2) &&
                                    (chan_norm == 0 || chan_norm == 1 || chan_norm == 2)),
                 "VTV point-wise norms must be 1, 2 or 0 (inf)");
  CHIP_CHECK_ARG(mode == kGrVTV || mode == kGrTNV || n_chan == 1, "coupled channels need the VTV or TNV mode");
  CHIP_CHECK_ARG(qref == nullptr || mode == kGrVTV, "a reference dual (qref) is taken by the VTV mode only");
  CHIP_CHECK_ARG(weights == nullptr || mode == kGrL12 || mode == kGrL1 || mode == kGrVTV,
                 "a weight array is taken by the l12, l1 and VTV modes");
  GradGeom g;
  const int rc = make_geom(g, n_chan, n_batch, nd, dims, strides, pad_mode);
  if (rc != CHIP_OK) return rc;
  const int64_t total = static_cast<int64_t>(g.C) * g.B * g.S;
  CHIP_CHECK_ARG(weights == nullptr || (weights_period >= 1 && total % weights_period == 0),
                 "weights_period must divide the primal's element count");
  const int64_t wp = weights ? weights_period : 1;
  hipStream_t st = as_stream(stream);
  if (nd == 1) launch_grad_dual<1>(q, x, g, mode, weight, weights, wp, der_norm, chan_norm, qref, st);
  else if (nd == 2) launch_grad_dual<2>(q, x, g, mode, weight, weights, wp, der_norm, chan_norm, qref, st);
  else launch_grad_dual<3>(q, x, g, mode, weight, weights, wp, der_norm, chan_norm, qref, st);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_grad_update_primal(float* upd, const float* q, int n_chan, int64_t n_batch, int nd, const int* dims,
                            const int64_t* strides, int pad_mode, float scale, const float* scales, int64_t scales_period, int apply_scale,
                            int subtract, const float* mask, int64_t mask_period, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(upd != nullptr && q != nullptr, "NULL argument");
  GradGeom g;
  const int rc = make_geom(g, n_chan, n_batch, nd, dims, strides, pad_mode);
  if (rc != CHIP_OK) return rc;
  const int64_t total = static_cast<int64_t>(g.C) * g.B * g.S;
  CHIP_CHECK_ARG(mask == nullptr || (mask_period >= 1 && total % mask_period == 0),
                 "mask_period must divide the primal's element count");
  CHIP_CHECK_ARG(scales == nullptr || (scales_period >= 1 && total % scales_period == 0),
                 "scales_period must divide the primal's element count");
  hipLaunchKernelGGL(grad_primal_kernel, gr_grid(total), dim3(kGrBlock), 0, as_stream(stream), upd, q, g, scale,
                     scales, scales ? scales_period : int64_t{1}, apply_scale, subtract, mask, mask_period);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

}  // extern "C"
