/*
 * C restatement of corrct's skimage-backend projector for the CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY: the oracle's timed CPU path (bench.py
 * cpu_baseline) and a second CPU implementation the tests compare with the
 * NumPy oracle. Never linked into the product library.
 *
 * Same float32 semantics as oracle/skimage_ops.py:
 *   radon (skimage radon_transform.py:104-114 as called by
 *   _projector_backends.py:333): inverse map (c u + s y) + m02, (-s u + c y) + m12
 *   in float32, order-1 blend (float32 fractions, float32 dc * right taps,
 *   the rest double, rounded to float32), float32 sum over rows in order;
 *   XRF fp (projectors.py:519-525): w[a,d] * radon(vol * M[a,d]).
 *   iradon bp (radon_transform.py:295-314, _projector_backends.py:350-359):
 *   t = (j - k) c - (i - k) s in double from float32 c, s; linear
 *   interpolation of the line, 0 outside [-k, n-1-k]; float32 accumulation.
 *   exact adjoint: the transpose of the fp as a scatter of each sample's bilinear
 *   weights (double), no reference counterpart (the north star's exact-adjoint bp).
 * Parallelised with OpenMP over (angle, detector) for fp and the adjoint, over rows
 * for the skimage bp.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline float tap(const float* img, int n, long r, long c) {
  return (r >= 0 && r < n && c >= 0 && c < n) ? img[r * n + c] : 0.0f;
}

static float bilinear(const float* img, int n, float ys, float xs) {
  const float r0f = floorf(ys), c0f = floorf(xs);
  const long r0 = (long)r0f, c0 = (long)c0f, r1 = (long)ceilf(ys), c1 = (long)ceilf(xs);
  const float drf = ys - r0f, dcf = xs - c0f;
  const double dr = drf, dc = dcf;
  const float tl = tap(img, n, r0, c0), tr = tap(img, n, r0, c1);
  const float bl = tap(img, n, r1, c0), br = tap(img, n, r1, c1);
  const float trw = dcf * tr, brw = dcf * br;
  const double top = (1.0 - dc) * (double)tl + (double)trw;
  const double bot = (1.0 - dc) * (double)bl + (double)brw;
  return (float)((1.0 - dr) * top + dr * bot);
}

/* geom[a] = (cos, sin, m02, m12); w[a*D + d]; maps may be NULL */
void oracle_fp(const float* vol, const float* maps, const float* geom, const float* w, int n, int A, int D,
               float* sino) {
  const size_t nn = (size_t)n * n;
#pragma omp parallel
  {
    float* tmp = (float*)malloc(nn * sizeof(float));
#pragma omp for schedule(dynamic)
    for (long ad = 0; ad < (long)A * D; ++ad) {
      const int a = (int)(ad / D), d = (int)(ad % D);
      const float* m = maps ? maps + ((size_t)a * D + d) * nn : NULL;
      for (size_t k = 0; k < nn; ++k) tmp[k] = m ? vol[k] * m[k] : vol[k];
      const float cs = geom[4 * a], sn = geom[4 * a + 1], m02 = geom[4 * a + 2], m12 = geom[4 * a + 3];
      float* out = sino + ((size_t)d * A + a) * n;
      for (int u = 0; u < n; ++u) out[u] = 0.0f;
      for (int y = 0; y < n; ++y) {
        const float yf = (float)y;
        for (int u = 0; u < n; ++u) {
          const float uf = (float)u;
          /* built with -ffp-contract=off: each product and sum rounds to float32 */
          const float xs = (cs * uf + sn * yf) + m02;
          const float ys = (-sn * uf + cs * yf) + m12;
          out[u] += bilinear(tmp, n, ys, xs);
        }
      }
      const float wt = w ? w[(size_t)a * D + d] : 1.0f;
      for (int u = 0; u < n; ++u) out[u] *= wt;
    }
    free(tmp);
  }
}

/* trig[a] = (cos, sin) as float32; sino [D][A][n]; vol [n][n] = sum_d sum_a interp */
void oracle_bp_skimage(const float* sino, const float* trig, int n, int A, int D, float* vol) {
  const int k = n / 2;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) vol[(size_t)i * n + j] = 0.0f;
    for (int d = 0; d < D; ++d) {
      for (int a = 0; a < A; ++a) {
        const double c = trig[2 * a], s = trig[2 * a + 1];
        const float* line = sino + ((size_t)d * A + a) * n;
        for (int j = 0; j < n; ++j) {
          const double t = (double)(j - k) * c - (double)(i - k) * s;
          float v = 0.0f;
          if (t >= -k && t <= n - 1 - k) {
            const double tf = floor(t);
            const int kk = (int)tf + k;
            const double fr = t - tf;
            const double s0 = line[kk], s1 = (kk + 1 < n) ? line[kk + 1] : 0.0;
            v = (float)(s0 + fr * (s1 - s0));
          }
          vol[(size_t)i * n + j] += v;
        }
      }
    }
  }
}

/* Exact transpose of oracle_fp, written as a scatter: every sample (a, d, y, u)
 * of the fp adds w[a,d] line[u] times its four bilinear tap weights (in double)
 * times M[a,d] at the tap into the pixel the fp read. geom / w / maps as in
 * oracle_fp; sino [D][A][n]; vol [n][n] (the double sums rounded once). */
void oracle_bp_adjoint(const float* sino, const float* maps, const float* geom, const float* w, int n, int A, int D,
                       float* vol) {
  const size_t nn = (size_t)n * n;
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  double* acc = (double*)calloc((size_t)nt * nn, sizeof(double));
#pragma omp parallel
  {
    int t = 0;
#ifdef _OPENMP
    t = omp_get_thread_num();
#endif
    double* my = acc + (size_t)t * nn;
#pragma omp for schedule(dynamic)
    for (long ad = 0; ad < (long)A * D; ++ad) {
      const int a = (int)(ad / D), d = (int)(ad % D);
      const float* m = maps ? maps + ((size_t)a * D + d) * nn : NULL;
      const float cs = geom[4 * a], sn = geom[4 * a + 1], m02 = geom[4 * a + 2], m12 = geom[4 * a + 3];
      const float* line = sino + ((size_t)d * A + a) * n;
      const double wt = w ? (double)w[(size_t)a * D + d] : 1.0;
      for (int y = 0; y < n; ++y) {
        const float yf = (float)y;
        for (int u = 0; u < n; ++u) {
          const double lv = wt * (double)line[u];
          if (lv == 0.0) continue;
          const float uf = (float)u;
          const float xs = (cs * uf + sn * yf) + m02;
          const float ys = (-sn * uf + cs * yf) + m12;
          const double r0f = floor((double)ys), c0f = floor((double)xs);
          const double dr = (double)ys - r0f, dc = (double)xs - c0f;
          const long r0 = (long)r0f, c0 = (long)c0f;
          const double tw[4] = {(1.0 - dr) * (1.0 - dc), (1.0 - dr) * dc, dr * (1.0 - dc), dr * dc};
          for (int k = 0; k < 4; ++k) {
            const long r = r0 + (k >> 1), c = c0 + (k & 1);
            if (r < 0 || r >= n || c < 0 || c >= n || tw[k] == 0.0) continue;
            const size_t pix = (size_t)r * n + (size_t)c;
            my[pix] += tw[k] * lv * (m ? (double)m[pix] : 1.0);
          }
        }
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)nn; ++k) {
    double s = 0.0;
    for (int t = 0; t < nt; ++t) s += acc[(size_t)t * nn + k];
    vol[k] = (float)s;
  }
  free(acc);
}

int oracle_num_threads(void) {
  int t = 1;
#pragma omp parallel
  {
#pragma omp single
    {
#ifdef _OPENMP
      t = omp_get_num_threads();
#endif
    }
  }
  return t;
}
