// Fused solver kernels of libcorrct_hip (gfx950): the elementwise parts of
// corrct's SIRT / MLEM / PDHG iterations (solvers.py:60-117, 688-1206),
// each a single HBM pass, plus deterministic reductions.
//
// Float32 operation order follows the reference's numpy expressions so
// that a few hundred iterations stay within 1e-4 of a float32 CPU run:
//   l2 dual   p = (p + (ax + bg) * sigma - sigma * b) * (1 / (1 + sigma))
//   KL dual   p = (1 + p - sqrt((p - 1)^2 + 4 sigma max(b, 0))) / 2
//   TV dual   q = q + 0.5 grad(x);  q = (q / max(|q|, lam)) * lam
//   primal    x_new = clip(x - (atp - div q) * tau) * mask; xr = x_new + (x_new - x)
#include <cfloat>

#include "common.hpp"
#include "solver_elem.hpp"

namespace chip {

constexpr int kEwBlock = 256;

static inline dim3 ew_grid(int64_t n) {
  int64_t blocks = (n + kEwBlock - 1) / kEwBlock;
  if (blocks > 65536) blocks = 65536;  // grid-stride beyond this
  if (blocks < 1) blocks = 1;
  return dim3(static_cast<unsigned>(blocks));
}

#define GRID_STRIDE(i, n) \
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < (n); i += static_cast<int64_t>(gridDim.x) * blockDim.x)

__device__ __forceinline__ float clipf(float v, int hl, float lo, int hu, float hi) {
  if (hl) v = fmaxf(v, lo);
  if (hu) v = fminf(v, hi);
  return v;
}

// ---- reductions -------------------------------------------------------------

__global__ void zero_u32_kernel(unsigned* p) { *p = 0u; }

// max |v| via atomicMax on the float bit pattern (order independent for v >= 0)
__global__ __launch_bounds__(kEwBlock) void absmax_kernel(const float* __restrict__ v, int64_t n, unsigned* out) {
  float m = 0.f;
  GRID_STRIDE(i, n) { m = fmaxf(m, fabsf(v[i])); }
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  __shared__ float s[kEwBlock / 64];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < kEwBlock / 64; ++k) m = fmaxf(m, s[k]);
    atomicMax(out, __float_as_uint(m));
  }
}

// single-workgroup double-precision sum of f(v): fixed order -> deterministic
template <int ABS>
__global__ __launch_bounds__(1024) void sum_kernel(const float* __restrict__ v, int64_t n, double* out) {
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) {
    const double x = v[i];
    acc += ABS ? fabs(x) : x * x;
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
  __shared__ double s[16];
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int k = 0; k < 16; ++k) t += s[k];
    *out = t;
  }
}

// ---- preconditioner -----------------------------------------------------------

// v = rho / fix(|v| + add); fix(t) = 1 if t / max < 1e-5 (solvers.py:98-116)
__global__ __launch_bounds__(kEwBlock) void precond_fix_kernel(float* __restrict__ v, int64_t n, float add,
                                                               const float* __restrict__ maxval, float rho) {
  // max(|v| + add) == max|v| + add: float rounding is monotonic
  const float mx = (add != 0.f) ? *maxval + add : *maxval;
  GRID_STRIDE(i, n) {
    float t = fabsf(v[i]);
    if (add != 0.f) t = t + add;
    if (t / mx < 1e-5f) t = 1.f;
    v[i] = rho / t;
  }
}

// y[i] += a * x[i % period] (a regularizer's masked step-size term added to tau)
__global__ __launch_bounds__(kEwBlock) void axpy_kernel(float* __restrict__ y, float a, const float* __restrict__ x,
                                                        int64_t period, int64_t n) {
  GRID_STRIDE(i, n) { y[i] = y[i] + a * x[i % period]; }
}

// v[i] = rho / v[i]
__global__ __launch_bounds__(kEwBlock) void rdiv_kernel(float* __restrict__ v, float rho, int64_t n) {
  GRID_STRIDE(i, n) { v[i] = rho / v[i]; }
}

// ---- PDHG data duals ----------------------------------------------------------

__global__ __launch_bounds__(kEwBlock) void dual_l2_kernel(float* __restrict__ p, const float* __restrict__ ax,
                                                           const float* __restrict__ sigma,
                                                           const float* __restrict__ b, const float* __restrict__ bg,
                                                           const float* __restrict__ mask, int64_t n) {
  GRID_STRIDE(i, n) {
    float a = ax[i];
    if (bg) a = a + bg[i];
    float v = elem_dual_l2(p[i], a, sigma[i], b[i]);
    if (mask) v = v * mask[i];
    p[i] = v;
  }
}

__global__ __launch_bounds__(kEwBlock) void dual_kl_kernel(float* __restrict__ p, const float* __restrict__ ax,
                                                           const float* __restrict__ sigma,
                                                           const float* __restrict__ b, const float* __restrict__ bg,
                                                           const float* __restrict__ mask, int64_t n) {
  GRID_STRIDE(i, n) {
    float a = ax[i];
    if (bg) a = a + bg[i];
    float r = elem_dual_kl(p[i], a, sigma[i], b[i]);
    if (mask) r = r * mask[i];
    p[i] = r;
  }
}

// ---- norm-type data duals (enum chip_data_term) -------------------------------

// sign(v) * max(|v| - t, 0)
__device__ __forceinline__ float soft_threshold(float v, float t) {
  const float m = fmaxf(fabsf(v) - t, 0.f);
  return v > 0.f ? m : (v < 0.f ? -m : 0.f);
}

// v = p + sigma (ax + bg) - sigma b, then the elementwise part of the term's
// dual proximal; for the group norms only the part before the group scaling.
__device__ __forceinline__ float dual_norm_pre(int term, float v, float s, float bv, float e) {
  v = v - s * bv;
  switch (term) {
    case CHIP_DATA_L2: return v * (1.f / (1.f + s));
    case CHIP_DATA_WL2: {
      const float z = e == 0.f ? 1.f : 0.f;  // e holds the weight
      return v * ((1.f / (1.f + s / (e + z))) * (1.f - z));
    }
    case CHIP_DATA_L2B: return soft_threshold(v, s * sqrtf(e)) * (1.f / (1.f + s * e));
    case CHIP_DATA_HUBER: return v * (1.f / (1.f + s * e));
    case CHIP_DATA_L1B: return soft_threshold(v, e);
    default: return v;  // L1, L12
  }
}

__device__ __forceinline__ bool dual_norm_projects(int term) {
  return term == CHIP_DATA_HUBER || term == CHIP_DATA_L1 || term == CHIP_DATA_L12 || term == CHIP_DATA_L1B;
}

__global__ __launch_bounds__(kEwBlock) void dual_norm_kernel(float* __restrict__ p, const float* __restrict__ ax,
                                                             const float* __restrict__ sigma,
                                                             const float* __restrict__ b,
                                                             const float* __restrict__ bg,
                                                             const float* __restrict__ mask, int64_t n, int term,
                                                             const float* __restrict__ prm, int64_t period) {
  const bool proj = dual_norm_projects(term);
  GRID_STRIDE(i, n) {
    const float s = sigma[i];
    float a = ax[i];
    if (bg) a = a + bg[i];
    const float e = prm ? prm[i % period] : 0.f;
    float v = dual_norm_pre(term, p[i] + a * s, s, b[i], e);
    if (proj) v = v / fmaxf(fabsf(v), 1.f);
    if (mask) v = v * mask[i];
    p[i] = v;
  }
}

// one thread per group (o, i): members (o * G + g) * S + i, g < G; the
// pre-scaled values go to p in a first pass while the squared norm builds up
// in member order, then a second pass divides by max(|group|_2, 1)
__global__ __launch_bounds__(kEwBlock) void dual_group_kernel(float* __restrict__ p, const float* __restrict__ ax,
                                                              const float* __restrict__ sigma,
                                                              const float* __restrict__ b,
                                                              const float* __restrict__ bg,
                                                              const float* __restrict__ mask, int64_t n_groups,
                                                              int term, const float* __restrict__ prm,
                                                              int64_t period, int64_t G, int64_t S) {
  GRID_STRIDE(t, n_groups) {
    const int64_t o = t / S;
    const int64_t base = o * G * S + (t - o * S);
    float acc = 0.f;
    for (int64_t g = 0; g < G; ++g) {
      const int64_t k = base + g * S;
      const float s = sigma[k];
      float a = ax[k];
      if (bg) a = a + bg[k];
      const float e = prm ? prm[k % period] : 0.f;
      const float v = dual_norm_pre(term, p[k] + a * s, s, b[k], e);
      acc = acc + v * v;
      p[k] = v;
    }
    const float den = fmaxf(sqrtf(acc), 1.f);
    for (int64_t g = 0; g < G; ++g) {
      const int64_t k = base + g * S;
      float v = p[k] / den;
      if (mask) v = v * mask[k];
      p[k] = v;
    }
  }
}

// compute_residual of the l2-based terms and KL, times the solver's sigma
__global__ __launch_bounds__(kEwBlock) void data_residual_kernel(float* __restrict__ res, const float* __restrict__ b,
                                                                 const float* __restrict__ ax,
                                                                 const float* __restrict__ bg,
                                                                 const float* __restrict__ mask,
                                                                 const float* __restrict__ asig,
                                                                 const float* __restrict__ scale, int64_t n,
                                                                 int term, const float* __restrict__ prm,
                                                                 int64_t period) {
  GRID_STRIDE(i, n) {
    float a = ax[i];
    if (bg) a = a + bg[i];
    const float sa = asig ? asig[i] : 1.f;
    float r;
    if (term == CHIP_DATA_KL) {
      const float v = fmaxf(a, FLT_EPSILON) * sa;
      const float sd = (4.f * sa) * fmaxf(b[i], 0.f);
      const float dm = v - 1.f;
      r = ((1.f + v) - sqrtf(dm * dm + sd)) / 2.f;
      if (mask) r = r * mask[i];
      r = -r;
    } else {
      r = b[i] - a;
      if (term == CHIP_DATA_WL2) r = r * prm[i % period];
      if (mask) r = r * mask[i];
      if (term == CHIP_DATA_L2B) r = soft_threshold(r, sa * sqrtf(prm[i % period]));
    }
    if (scale) r = r * scale[i];
    res[i] = r;
  }
}

// dual variable consistent with a primal estimate: the data term's gradient at
// A x (the fixed point of the dual updates above): l2 p = (A x + bg) - b,
// KL p = 1 - b / (A x + bg) (0 where A x + bg <= 0)
__global__ __launch_bounds__(kEwBlock) void dual_init_kernel(float* __restrict__ p, const float* __restrict__ ax,
                                                             const float* __restrict__ b,
                                                             const float* __restrict__ bg,
                                                             const float* __restrict__ mask, int64_t n, int kl) {
  GRID_STRIDE(i, n) {
    float a = ax[i];
    if (bg) a = a + bg[i];
    float r;
    if (kl) r = a > 0.f ? 1.f - fmaxf(b[i], 0.f) / a : 0.f;
    else r = a - b[i];
    if (mask) r = r * mask[i];
    p[i] = r;
  }
}

// ---- TV2D -----------------------------------------------------------------------

// q[b][0] += 0.5 * (x[i+1][j] - x[i][j]); q[b][1] += 0.5 * (x[i][j+1] - x[i][j]);
// then q <- (q / max(|q|_2, lam)) * lam
__global__ __launch_bounds__(kEwBlock) void tv_dual_kernel(float* __restrict__ q, const float* __restrict__ x, int n,
                                                           int batch, float lam) {
  const int64_t nn = static_cast<int64_t>(n) * n;
  const int64_t total = nn * batch;
  GRID_STRIDE(k, total) {
    const int64_t b = k / nn;
    const int64_t pix = k - b * nn;
    const int i = static_cast<int>(pix / n);
    const int j = static_cast<int>(pix - static_cast<int64_t>(i) * n);
    const float* xb = x + b * nn;
    const float xc = xb[pix];
    const float g0 = (i + 1 < n) ? xb[pix + n] - xc : 0.f;
    const float g1 = (j + 1 < n) ? xb[pix + 1] - xc : 0.f;
    float* q0 = q + (2 * b) * nn + pix;
    float* q1 = q0 + nn;
    float a0 = *q0 + 0.5f * g0;
    float a1 = *q1 + 0.5f * g1;
    const float nrm = sqrtf(a0 * a0 + a1 * a1);
    const float den = fmaxf(nrm, lam);
    *q0 = (a0 / den) * lam;
    *q1 = (a1 / den) * lam;
  }
}

// -div q with edge padding: d[k] = q[k] - q[k-1] (k >= 1), d[0] = 0 per axis
__device__ __forceinline__ float neg_div(const float* __restrict__ q0, const float* __restrict__ q1, int64_t pix, int i,
                                         int j, int n) {
  const float d0 = (i > 0) ? q0[pix] - q0[pix - n] : 0.f;
  const float d1 = (j > 0) ? q1[pix] - q1[pix - 1] : 0.f;
  return -(d0 + d1);
}

__global__ __launch_bounds__(kEwBlock) void pdhg_primal_kernel(float* __restrict__ x, float* __restrict__ xbar,
                                                               const float* __restrict__ atp,
                                                               const float* __restrict__ q,
                                                               const float* __restrict__ tau, int hl, float lo,
                                                               int hu, float hi, const float* __restrict__ mask,
                                                               int n, int batch) {
  const int64_t nn = static_cast<int64_t>(n) * n;
  const int64_t total = nn * batch;
  GRID_STRIDE(k, total) {
    const int64_t b = k / nn;
    const int64_t pix = k - b * nn;
    float upd = atp[k];
    if (q) {
      const int i = static_cast<int>(pix / n);
      const int j = static_cast<int>(pix - static_cast<int64_t>(i) * n);
      const float* q0 = q + (2 * b) * nn;
      upd = upd + neg_div(q0, q0 + nn, pix, i, j, n);
    }
    const float xo = x[k];
    float xn = xo - upd * tau[k];
    xn = clipf(xn, hl, lo, hu, hi);
    if (mask) xn = xn * mask[k];
    xbar[k] = xn + (xn - xo);
    x[k] = xn;
  }
}

// ---- SIRT -----------------------------------------------------------------------

__global__ __launch_bounds__(kEwBlock) void sirt_residual_kernel(float* __restrict__ res, const float* __restrict__ b,
                                                                 const float* __restrict__ ax,
                                                                 const float* __restrict__ bg,
                                                                 const float* __restrict__ mask,
                                                                 const float* __restrict__ sigma, int64_t n) {
  GRID_STRIDE(i, n) {
    float a = ax[i];
    if (bg) a = a + bg[i];
    float r = elem_residual(a, b[i]);
    if (mask) r = r * mask[i];
    if (sigma) r = r * sigma[i];
    res[i] = r;
  }
}

__global__ __launch_bounds__(kEwBlock) void sirt_update_kernel(float* __restrict__ x, const float* __restrict__ upd,
                                                               const float* __restrict__ q,
                                                               const float* __restrict__ tau, int hl, float lo,
                                                               int hu, float hi, const float* __restrict__ mask,
                                                               int n, int batch) {
  const int64_t nn = static_cast<int64_t>(n) * n;
  const int64_t total = nn * batch;
  GRID_STRIDE(k, total) {
    const int64_t b = k / nn;
    const int64_t pix = k - b * nn;
    float u = upd[k];
    if (q) {
      const int i = static_cast<int>(pix / n);
      const int j = static_cast<int>(pix - static_cast<int64_t>(i) * n);
      const float* q0 = q + (2 * b) * nn;
      u = u - neg_div(q0, q0 + nn, pix, i, j, n);
    }
    float xn = x[k] + u * tau[k];
    xn = clipf(xn, hl, lo, hu, hi);
    if (mask) xn = xn * mask[k];
    x[k] = xn;
  }
}

// ---- FBP filtering ----------------------------------------------------------------

// out[l][i] = sum_{j < n} in[l][j] h[(i - j) mod p]: the zero-padded circular
// convolution of length p (a power of two >= 2n) that corrct's FFT filtering
// computes; one workgroup per line, the line and h in LDS, float64 sums.
constexpr int kFiltMaxP = 4096;

__global__ __launch_bounds__(256) void filter_lines_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                           const double* __restrict__ h, int n, int p) {
  __shared__ double hs[kFiltMaxP];
  __shared__ float xs[kFiltMaxP / 2];
  const int64_t l = blockIdx.x;
  for (int k = threadIdx.x; k < p; k += blockDim.x) hs[k] = h[k];
  for (int k = threadIdx.x; k < n; k += blockDim.x) xs[k] = in[l * n + k];
  __syncthreads();
  const int mask = p - 1;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double acc = 0.0;
    for (int j = 0; j < n; ++j) acc = fma(static_cast<double>(xs[j]), hs[(i - j) & mask], acc);
    out[l * n + i] = static_cast<float>(acc);
  }
}

// ---- detector PSF -----------------------------------------------------------------

// out[l][i] = sum_{k<K} xe[i + (K-1)/2 - k] kern[k]: "same" convolution of each
// line extended by (K-1)/2 edge copies on each side and zero beyond
__global__ __launch_bounds__(256) void psf_lines_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                        const float* __restrict__ kern, int k_len, int64_t n_lines,
                                                        int n) {
  const int pw = (k_len - 1) / 2;
  const int64_t total = n_lines * n;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t l = t / n;
    const int i = static_cast<int>(t - l * n);
    const float* row = in + l * n;
    float acc = 0.f;
    for (int k = 0; k < k_len; ++k) {
      const int m = i + pw - k;
      if (m < -pw || m > n - 1 + pw) continue;
      acc = fmaf(row[min(max(m, 0), n - 1)], kern[k], acc);
    }
    out[t] = acc;
  }
}

// ---- FISTA ----------------------------------------------------------------------

// x_new = y + tau * grad; optional l1 soft threshold at tau * l1w; clip; mask;
// y = x_new + momentum * (x_new - x); x = x_new   (tau per element or scalar)
__global__ __launch_bounds__(kEwBlock) void fista_update_kernel(float* __restrict__ x, float* __restrict__ y,
                                                                const float* __restrict__ grad,
                                                                const float* __restrict__ tau, float tau_s, float l1w,
                                                                const float* __restrict__ l1wa, int64_t l1w_period,
                                                                const float* __restrict__ l1m, int64_t l1m_period,
                                                                int hl, float lo, int hu, float hi,
                                                                const float* __restrict__ mask, float momentum,
                                                                int64_t n) {
  const bool prox = l1wa != nullptr || l1w > 0.f;
  GRID_STRIDE(i, n) {
    const float t = tau ? tau[i] : tau_s;
    float xn = y[i] + t * grad[i];
    if (prox) {
      const float thr = t * (l1wa ? l1wa[i % l1w_period] : l1w);
      if (l1m) {  // thresholded inside the update mask, blended outside
        const float m = l1m[i % l1m_period];
        const float xm = soft_threshold(xn * m, thr);
        xn = xn * (1.f - m) + xm;
      } else {
        xn = soft_threshold(xn, thr);
      }
    }
    xn = clipf(xn, hl, lo, hu, hi);
    if (mask) xn = xn * mask[i];
    y[i] = xn + momentum * (xn - x[i]);
    x[i] = xn;
  }
}

// ---- Regularizer_l1 (identity transform) ------------------------------------------

__global__ __launch_bounds__(kEwBlock) void l1_dual_kernel(float* __restrict__ q, const float* __restrict__ x,
                                                           int64_t n, float w, const float* __restrict__ wa,
                                                           int64_t w_period) {
  GRID_STRIDE(i, n) {
    const float wi = wa ? wa[i % w_period] : w;
    const float v = q[i] + x[i];
    q[i] = (v / fmaxf(fabsf(v), wi)) * wi;
  }
}

__global__ __launch_bounds__(kEwBlock) void l1_primal_kernel(float* __restrict__ upd, const float* __restrict__ q,
                                                             int64_t n, const float* __restrict__ mask,
                                                             int64_t m_period, int subtract) {
  GRID_STRIDE(i, n) {
    float t = q[i];
    if (mask) t = t * mask[i % m_period];
    upd[i] = subtract ? upd[i] - t : upd[i] + t;
  }
}

// ---- MLEM -----------------------------------------------------------------------

__global__ __launch_bounds__(kEwBlock) void mlem_ratio_kernel(float* __restrict__ r, const float* __restrict__ b,
                                                              const float* __restrict__ ax,
                                                              const float* __restrict__ bg, int64_t n) {
  GRID_STRIDE(i, n) {
    float a = ax[i];
    if (bg) a = a + bg[i];
    r[i] = elem_mlem_ratio(a, b[i]);
  }
}

__global__ __launch_bounds__(kEwBlock) void mlem_update_kernel(float* __restrict__ x, const float* __restrict__ upd,
                                                               const float* __restrict__ tau, int hl, float lo,
                                                               int hu, float hi, const float* __restrict__ mask,
                                                               int64_t n) {
  GRID_STRIDE(i, n) {
    float xn = x[i] * (upd[i] / tau[i]);
    xn = clipf(xn, hl, lo, hu, hi);
    if (mask) xn = xn * mask[i];
    x[i] = xn;
  }
}

// ---- array-valued limits ----------------------------------------------------------

// x = clip(x, lower, upper) * mask, each limit an array (lo_a / hi_a), a scalar
// (hl / hu without an array) or absent; then, with xbar, xbar = x + momentum (x - x_old)
__global__ __launch_bounds__(kEwBlock) void clip_relax_kernel(float* __restrict__ x, const float* __restrict__ lo_a,
                                                              const float* __restrict__ hi_a, int hl, float lo,
                                                              int hu, float hi, const float* __restrict__ mask,
                                                              float* __restrict__ xbar,
                                                              const float* __restrict__ x_old, float momentum,
                                                              int64_t n) {
  GRID_STRIDE(i, n) {
    float v = x[i];
    if (lo_a) v = fmaxf(v, lo_a[i]);
    else if (hl) v = fmaxf(v, lo);
    if (hi_a) v = fminf(v, hi_a[i]);
    else if (hu) v = fminf(v, hi);
    if (mask) v = v * mask[i];
    x[i] = v;
    if (xbar) xbar[i] = v + momentum * (v - x_old[i]);
  }
}

}  // namespace chip

using namespace chip;

extern "C" {

int chip_clip_relax(float* x, const float* lower, const float* upper, int has_lower, float lower_s, int has_upper,
                    float upper_s, const float* x_mask, float* xbar, const float* x_old, float momentum, int64_t n,
                    void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x != nullptr, "x is NULL");
  CHIP_CHECK_ARG(n >= 0, "n must be >= 0");
  CHIP_CHECK_ARG(xbar == nullptr || x_old != nullptr, "xbar needs x_old");
  if (n == 0) return CHIP_OK;
  hipLaunchKernelGGL(clip_relax_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), x, lower, upper, has_lower,
                     lower_s, has_upper, upper_s, x_mask, xbar, x_old, momentum, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_absmax(const float* v, int64_t n, float* out, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(v != nullptr && out != nullptr, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipStream_t st = as_stream(stream);
  unsigned* o = reinterpret_cast<unsigned*>(out);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, st, o);
  dim3 g = ew_grid(n);
  if (g.x > 2048) g.x = 2048;
  hipLaunchKernelGGL(absmax_kernel, g, dim3(kEwBlock), 0, st, v, n, o);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_precond_fix(float* v, int64_t n, float add, const float* maxval, float rho, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(v != nullptr && maxval != nullptr, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(precond_fix_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), v, n, add, maxval, rho);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_axpy(float* y, float a, const float* x, int64_t x_period, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(y != nullptr && x != nullptr, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1 && x_period >= 1 && n % x_period == 0, "x_period must divide n");
  hipLaunchKernelGGL(axpy_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), y, a, x, x_period, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_rdiv(float* v, float rho, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(v != nullptr && n >= 1, "bad argument");
  hipLaunchKernelGGL(rdiv_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), v, rho, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_pdhg_dual_l2(float* p, const float* ax, const float* sigma, const float* b, const float* bg,
                      const float* b_mask, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(p && ax && sigma && b, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(dual_l2_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), p, ax, sigma, b, bg, b_mask, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_pdhg_dual_kl(float* p, const float* ax, const float* sigma, const float* b, const float* bg,
                      const float* b_mask, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(p && ax && sigma && b, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(dual_kl_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), p, ax, sigma, b, bg, b_mask, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_pdhg_dual_init(float* p, const float* ax, const float* b, const float* bg, const float* b_mask, int64_t n,
                        int kl, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(p && ax && b, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(dual_init_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), p, ax, b, bg, b_mask, n, kl);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_tv_dual(float* q, const float* x, int n, int batch, float lam, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(q && x, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1 && batch >= 1, "bad shape");
  CHIP_CHECK_ARG(lam > 0.f, "lambda must be > 0");
  const int64_t total = static_cast<int64_t>(n) * n * batch;
  hipLaunchKernelGGL(tv_dual_kernel, ew_grid(total), dim3(kEwBlock), 0, as_stream(stream), q, x, n, batch, lam);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_pdhg_primal(float* x, float* xbar, const float* atp, const float* q, const float* tau, int has_lower,
                     float lower, int has_upper, float upper, const float* x_mask, int n, int batch, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && xbar && atp && tau, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1 && batch >= 1, "bad shape");
  const int64_t total = static_cast<int64_t>(n) * n * batch;
  hipLaunchKernelGGL(pdhg_primal_kernel, ew_grid(total), dim3(kEwBlock), 0, as_stream(stream), x, xbar, atp, q, tau,
                     has_lower, lower, has_upper, upper, x_mask, n, batch);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_pdhg_dual_norm(float* p, const float* ax, const float* sigma, const float* b, const float* bg,
                        const float* b_mask, int64_t n, int term, const float* param, int64_t param_period,
                        int64_t group_len, int64_t group_stride, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(p && ax && sigma && b, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  CHIP_CHECK_ARG(term == CHIP_DATA_L2 || (term >= CHIP_DATA_WL2 && term <= CHIP_DATA_L1B),
                 "term must be L2, WL2, L2B, HUBER, L1, L12 or L1B");
  const bool needs_param = term == CHIP_DATA_WL2 || term == CHIP_DATA_L2B || term == CHIP_DATA_HUBER ||
                           term == CHIP_DATA_L1B;
  CHIP_CHECK_ARG(!needs_param || param != nullptr, "this data term needs its parameter array");
  CHIP_CHECK_ARG(!param || (param_period >= 1 && n % param_period == 0), "param_period must divide n");
  CHIP_CHECK_ARG(group_len >= 1 && group_stride >= 1 && n % (group_len * group_stride) == 0,
                 "n must be a multiple of group_len * group_stride");
  CHIP_CHECK_ARG(group_len == 1 || term == CHIP_DATA_L12 || term == CHIP_DATA_HUBER,
                 "groups apply to L12 and HUBER only");
  const int64_t period = param ? param_period : 1;
  if (group_len == 1) {
    hipLaunchKernelGGL(dual_norm_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), p, ax, sigma, b, bg,
                       b_mask, n, term, param, period);
  } else {
    const int64_t groups = n / group_len;
    hipLaunchKernelGGL(dual_group_kernel, ew_grid(groups), dim3(kEwBlock), 0, as_stream(stream), p, ax, sigma, b,
                       bg, b_mask, groups, term, param, period, group_len, group_stride);
  }
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_data_residual(float* res, const float* b, const float* ax, const float* bg, const float* b_mask,
                       const float* asig, const float* scale, int64_t n, int term, const float* param,
                       int64_t param_period, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(res && b && ax, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  CHIP_CHECK_ARG(term == CHIP_DATA_L2 || term == CHIP_DATA_KL || term == CHIP_DATA_WL2 || term == CHIP_DATA_L2B,
                 "term must be L2, KL, WL2 or L2B");
  const bool needs_param = term == CHIP_DATA_WL2 || term == CHIP_DATA_L2B;
  CHIP_CHECK_ARG(!needs_param || param != nullptr, "this data term needs its parameter array");
  CHIP_CHECK_ARG(!param || (param_period >= 1 && n % param_period == 0), "param_period must divide n");
  hipLaunchKernelGGL(data_residual_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), res, b, ax, bg, b_mask,
                     asig, scale, n, term, param, param ? param_period : int64_t{1});
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_sirt_residual(float* res, const float* b, const float* ax, const float* bg, const float* b_mask,
                       const float* sigma, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(res && b && ax, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(sirt_residual_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), res, b, ax, bg, b_mask,
                     sigma, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_sirt_update(float* x, const float* upd, const float* q, const float* tau, int has_lower, float lower,
                     int has_upper, float upper, const float* x_mask, int n, int batch, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && upd && tau, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1 && batch >= 1, "bad shape");
  const int64_t total = static_cast<int64_t>(n) * n * batch;
  hipLaunchKernelGGL(sirt_update_kernel, ew_grid(total), dim3(kEwBlock), 0, as_stream(stream), x, upd, q, tau,
                     has_lower, lower, has_upper, upper, x_mask, n, batch);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_filter_lines(const float* in, float* out, const double* h, int64_t n_lines, int n, int p, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(in && out && h, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1 && n_lines >= 1 && n_lines <= 2147483647, "bad shape");
  CHIP_CHECK_ARG(p >= 2 * n && (p & (p - 1)) == 0 && p <= kFiltMaxP,
                 "p must be a power of two with 2 n <= p <= " + std::to_string(kFiltMaxP));
  CHIP_CHECK_ARG(in != out, "in and out must not alias");
  hipLaunchKernelGGL(filter_lines_kernel, dim3(static_cast<unsigned>(n_lines)), dim3(256), 0, as_stream(stream), in,
                     out, h, n, p);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_psf_lines(const float* in, float* out, const float* kern, int k_len, int64_t n_lines, int n,
                   void* stream) {
  clear_error();
  CHIP_CHECK_ARG(in && out && kern, "NULL pointer");
  CHIP_CHECK_ARG(in != out, "in and out must not alias");
  CHIP_CHECK_ARG(k_len >= 1 && n >= 1 && n_lines >= 1, "bad shape");
  hipLaunchKernelGGL(psf_lines_kernel, ew_grid(n_lines * n), dim3(kEwBlock), 0, as_stream(stream), in, out, kern,
                     k_len, n_lines, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_fista_update(float* x, float* y, const float* grad, const float* tau, float tau_scalar, float l1_weight,
                      const float* l1_weights, int64_t l1_weights_period, const float* l1_mask,
                      int64_t l1_mask_period, int has_lower, float lower, int has_upper, float upper,
                      const float* x_mask, float momentum, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && y && grad, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  CHIP_CHECK_ARG(l1_weight >= 0.f, "l1_weight must be >= 0");
  CHIP_CHECK_ARG(!l1_weights || (l1_weights_period >= 1 && n % l1_weights_period == 0),
                 "l1_weights_period must divide n");
  CHIP_CHECK_ARG(!l1_mask || (l1_mask_period >= 1 && n % l1_mask_period == 0), "l1_mask_period must divide n");
  hipLaunchKernelGGL(fista_update_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), x, y, grad, tau,
                     tau_scalar, l1_weight, l1_weights, l1_weights ? l1_weights_period : int64_t{1}, l1_mask,
                     l1_mask ? l1_mask_period : int64_t{1}, has_lower, lower, has_upper, upper, x_mask, momentum, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_l1_dual(float* q, const float* x, int64_t n, float weight, const float* weights, int64_t weights_period,
                 void* stream) {
  clear_error();
  CHIP_CHECK_ARG(q && x, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  CHIP_CHECK_ARG(!weights || (weights_period >= 1 && n % weights_period == 0), "weights_period must divide n");
  hipLaunchKernelGGL(l1_dual_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), q, x, n, weight, weights,
                     weights ? weights_period : int64_t{1});
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_l1_update_primal(float* upd, const float* q, int64_t n, const float* mask, int64_t mask_period, int subtract,
                          void* stream) {
  clear_error();
  CHIP_CHECK_ARG(upd && q, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  CHIP_CHECK_ARG(!mask || (mask_period >= 1 && n % mask_period == 0), "mask_period must divide n");
  hipLaunchKernelGGL(l1_primal_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), upd, q, n, mask,
                     mask ? mask_period : int64_t{1}, subtract);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_mlem_ratio(float* r, const float* b, const float* ax, const float* bg, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(r && b && ax, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(mlem_ratio_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), r, b, ax, bg, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_mlem_update(float* x, const float* upd, const float* tau, int has_lower, float lower, int has_upper,
                     float upper, const float* x_mask, int64_t n, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(x && upd && tau, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(mlem_update_kernel, ew_grid(n), dim3(kEwBlock), 0, as_stream(stream), x, upd, tau, has_lower,
                     lower, has_upper, upper, x_mask, n);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_sumsq(const float* v, int64_t n, double* out, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(v && out, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(sum_kernel<0>, dim3(1), dim3(1024), 0, as_stream(stream), v, n, out);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_sumabs(const float* v, int64_t n, double* out, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(v && out, "NULL pointer");
  CHIP_CHECK_ARG(n >= 1, "n must be >= 1");
  hipLaunchKernelGGL(sum_kernel<1>, dim3(1), dim3(1024), 0, as_stream(stream), v, n, out);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

}  // extern "C"
