/*
 * corrct_hip.h -- C ABI of libcorrct_hip.so, the MI355X (gfx950) projector
 * backend for attenuation-corrected XRF tomography.
 *
 * Conventions
 *   - every function returns an int status (CHIP_OK = 0); on error
 *     chip_last_error() returns a per-thread message;
 *   - array pointers are DEVICE pointers owned by the caller; the library
 *     never returns views of internal memory;
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream);
 *     all work is enqueued on it, nothing is synchronised unless stated;
 *   - layouts are C-contiguous float32: volume [B][N][N] (rows = Y, cols = X),
 *     maps [A][D][N][N] per slice, sinogram [B][D][A][N].
 *
 * Each entry point names the reference (corrct 2.0.0) interface it replaces,
 * as file:line relative to the corrct source tree.
 */
#ifndef CORRCT_HIP_H
#define CORRCT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chip_plan chip_plan;
typedef struct chip_batch_plan chip_batch_plan;

enum chip_status {
  CHIP_OK = 0,
  CHIP_EINVAL = 1, /* bad argument (shape, count, null pointer, mode) */
  CHIP_EHIP = 2,   /* HIP runtime error */
  CHIP_ENOMEM = 3  /* device allocation failed */
};

enum chip_bp_mode {
  /* sum_d bp_skimage(sino[d]): iradon(filter None) discretisation, no maps,
   * no weights -- ProjectorAttenuationXRF.bp with is_symmetric=False
   * (src/corrct/projectors.py:619-622) and ProjectorUncorrected.bp
   * (projectors.py:323-341 -> _projector_backends.py:335-359) */
  CHIP_BP_SKIMAGE = 0,
  /* sum_a sum_d w[a,d] M[a,d] * bp_skimage_a(sino[d,a]) -- is_symmetric=True
   * (projectors.py:533-568, 609-618) */
  CHIP_BP_SKIMAGE_SYM = 1,
  /* exact transpose of chip_fp (bilinear weights, maps, weights); no
   * reference counterpart: used for the adjoint (dot) test */
  CHIP_BP_ADJOINT = 2
};

/* Data steps of the fused on-the-fly iteration (chip_otf_fp_row_bp), each the
 * elementwise kernel of the same name applied to one angle block's rows. */
enum chip_row_op {
  CHIP_ROW_PDHG_L2 = 0, /* p = prox_l2(p + sigma (ax + bg)) * mask (chip_pdhg_dual_l2) */
  CHIP_ROW_PDHG_KL = 1, /* p = prox_KL(p + sigma (ax + bg)) * mask (chip_pdhg_dual_kl) */
  CHIP_ROW_SIRT = 2,    /* p = (b - (ax + bg)) * mask * sigma (chip_sirt_residual; sigma may be NULL) */
  CHIP_ROW_MLEM = 3     /* p = b / max(ax + bg, eps) (chip_mlem_ratio; sigma and mask unused) */
};

/* Data-fidelity terms of chip_pdhg_dual_norm / chip_data_residual
 * (src/corrct/data_terms.py). `e` is the term's local_error, `w` its weights,
 * both read as param[i % param_period]. */
enum chip_data_term {
  CHIP_DATA_L2 = 0,    /* DataFidelity_l2     data_terms.py:283-331 */
  CHIP_DATA_KL = 1,    /* DataFidelity_KL     data_terms.py:677-757 */
  CHIP_DATA_WL2 = 2,   /* DataFidelity_wl2    data_terms.py:334-395 (param = w) */
  CHIP_DATA_L2B = 3,   /* DataFidelity_l2b    data_terms.py:398-464 (param = e) */
  CHIP_DATA_HUBER = 4, /* DataFidelity_Huber  data_terms.py:467-532 (param = e) */
  CHIP_DATA_L1 = 5,    /* DataFidelity_l1     data_terms.py:535-597 */
  CHIP_DATA_L12 = 6,   /* DataFidelity_l12    data_terms.py:598-625 */
  CHIP_DATA_L1B = 7    /* DataFidelity_l1b    data_terms.py:628-674 (param = e) */
};

/* Library version (major*10000 + minor*100 + patch). */
int chip_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* chip_last_error(void);

/* Number of exported entry points and their names (for load checks). */
int chip_num_symbols(void);
const char* chip_symbol_name(int i);

/*
 * Geometry plan for an n x n slice, n_angles rotation angles (float32
 * radians, as ProjectorBackend.initialize_geometry stores them:
 * _projector_backends.py:145-160), n_det detectors at det_angles_rad
 * (float64, physics/attenuation.py:97-109), detector weights w_det[n_det]
 * (NULL = ones) and per-angle weights w_ang[n_angles][n_det] (NULL = ones)
 * -- ProjectorAttenuationXRF.__init__ (projectors.py:439-459).
 * Replaces ProjectorBackendSKimage.initialize_geometry
 * (_projector_backends.py:259-299) and the weight handling above.
 * Allocates small per-angle tables on the current device.
 */
int chip_plan_create(chip_plan** out, int n, int n_angles, const float* angles_rad, int n_det,
                     const double* det_angles_rad, const double* w_det, const double* w_ang);
int chip_plan_destroy(chip_plan* plan);

/*
 * Radon tables from the caller's float32 cos / sin: cos_sin[a] = (cos, sin) of
 * deg2rad(rad2deg(angles_rad[a])) in float32. scikit-image's radon / iradon
 * take them from numpy's float32 cos / sin (skimage radon_transform.py:104-112,
 * 295-308, as ProjectorBackendSKimage calls them, _projector_backends.py:299-359),
 * which can differ by one ulp from the correctly rounded values chip_plan_create
 * uses; at N = 1024 a one-ulp cosine moves a whole projection line by ~3e-5
 * pixels. The Python layer passes numpy's values. Synchronises the device
 * (queued work may read the old tables) and rewrites the plan's fp / bp
 * geometry; the attenuation maps are unaffected (they use float64 angles).
 */
int chip_plan_set_radon_trig(chip_plan* plan, const float* cos_sin);

/*
 * Detector-surface sampling of the emitted beam: detector d's outgoing maps
 * become the float32 mean, in order, of exp(-cumulative mu) over its
 * n_sub_det[d] (1 <= n_sub_det[d] <= n_sub_max <= 64) angles
 * ((theta_a + phi_d) + sub_angles[d][k]) + pi; sub_angles is
 * [n_det][n_sub_max] float64 (entries past n_sub_det[d] unused), angles_rad
 * and det_angles_rad as given to chip_plan_create. Replaces the sub-angle
 * loop of AttenuationVolume._compute_attenuation_angle_out
 * (physics/attenuation.py:110-119) with DetectorXRF.angle_range_rad
 * (physics/xrf.py:238-246) evaluated by the caller. Synchronises the device
 * (the previous tables may be in use) and replaces the plan's tables.
 */
int chip_plan_set_subangles(chip_plan* plan, int n_sub_max, const int* n_sub_det, const double* sub_angles,
                            const float* angles_rad, const double* det_angles_rad);

/*
 * Attenuation maps maps[a][d][n][n] = E_in(theta_a) * E_out(theta_a + phi_d + pi)
 * with E(phi) = exp(-trapezoid cumulative sum of mu along the beam), for
 * `batch` slices (mu_* and maps advance by n*n and n_angles*n_det*n*n per
 * slice). mu_in or mu_out may be NULL (factor 1). Replaces
 * AttenuationVolume.compute_maps (physics/attenuation.py:122-180) over
 * compute_attenuation (_projector_backends.py:43-99). Uses a plan-owned
 * scratch buffer (allocated on first use, reused, freed with the plan).
 */
int chip_att_maps(chip_plan* plan, const float* mu_in, const float* mu_out, float* maps, int batch,
                  void* stream);

/*
 * One solver iteration's projector work with on-the-fly maps, angle block by
 * angle block: build the block's maps once, fp x over the block into ax, apply
 * row_op to the block's rows of p (enum chip_row_op; b, sigma, bg, b_mask as the
 * op needs, NULL where absent), and back-project the block's rows of p into vol
 * (bp_mode CHIP_BP_ADJOINT or CHIP_BP_SKIMAGE_SYM, accumulated over the blocks
 * in order). The results equal chip_fp_otf, the elementwise kernel and
 * chip_bp_otf called in turn, bit for bit, with half the map building: the data
 * steps of SIRT, MLEM and PDHG (solvers.py:775-803, 951-982, 1165-1183) are
 * elementwise on the sinogram, so each block's rows are final once its fp is.
 * No reference counterpart (the reference always precomputes the maps).
 * Layouts as chip_fp_otf / chip_bp_otf; p, ax and the data are [batch][D][A][N].
 */
int chip_otf_fp_row_bp(chip_plan* plan, const float* x, const float* mu_in, const float* mu_out, int batch,
                       int row_op, float* p, const float* sigma, const float* b, const float* bg,
                       const float* b_mask, float* ax, float* vol, int bp_mode, void* stream);

/*
 * Binned maps: each of n_img images of nd (2 or 3) axes dims[nd] (slowest
 * first) is cropped to a multiple of `binning` per axis, dropping excess - excess/2
 * samples at the start and excess/2 at the end, and averaged over blocks of
 * binning^nd samples into out[n_img][dims[k] / binning ...]. Replaces
 * AttenuationVolume.get_maps(binning=...) (physics/attenuation.py:294-295) over
 * bin_imgs(auto_crop=True) (processing/pre.py:180-220: skimage block_reduce
 * with np.mean). Sums in row-major block order in float32.
 */
int chip_bin_images(const float* in, float* out, int64_t n_img, int nd, const int* dims, int binning, void* stream);

/*
 * Forward projection sino[b][d][a][u] = w[a,d] * radon_a(vol[b] * M[a,d])[u]
 * for `batch` slices. maps may be NULL (M = 1: ProjectorUncorrected.fp,
 * projectors.py:302-321); maps_slice_stride is the element stride between
 * slices' map sets (0 = all slices share one set). Replaces
 * ProjectorAttenuationXRF.fp (projectors.py:570-596) over fp_angle
 * (projectors.py:502-531) and ProjectorBackendSKimage.fp
 * (_projector_backends.py:312-333).
 */
int chip_fp(const chip_plan* plan, const float* vol, const float* maps, int64_t maps_slice_stride,
            float* sino, int batch, void* stream);

/*
 * Back-projection vol[b] = BP_mode(sino[b]); see enum chip_bp_mode. Replaces
 * ProjectorAttenuationXRF.bp (projectors.py:598-622) and
 * ProjectorBackendSKimage.bp (_projector_backends.py:335-359).
 * One slice with 64 or more angles is split over its angle range to fill the
 * GPU: adjoint mode with N <= 1024 and at most two detectors sums 2-8 parts
 * in plan scratch (8 N^2 floats, allocated by the first such call unless the
 * stream is being captured) and adds them in order; otherwise two halves are
 * added atomically into the zeroed vol. The part count follows the kernel's
 * occupancy on the device, so the bits can differ between GPUs and builds
 * unless chip_plan_tune "bp_parts" pins it. Calls on one plan from different
 * streams are ordered on the scratch by an event; a captured graph that uses
 * the scratch must not replay while another call on the plan runs. vol must
 * not alias sino.
 */
int chip_bp(const chip_plan* plan, const float* sino, const float* maps, int64_t maps_slice_stride,
            float* vol, int batch, int mode, void* stream);

/*
 * chip_bp without device allocation. No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc). Same arguments, kernels
 * and results as chip_bp; the part sums of a split single-slice adjoint bp go
 * to a static 16 MiB device array of the library (one per GPU, held until the
 * process ends) instead of plan memory that chip_bp allocates on first use.
 * A call whose parts x n^2 floats exceed 4 Mi (e.g. 5 or more parts at
 * n = 1024) returns CHIP_EINVAL; chip_plan_tune "bp_parts" can pin fewer.
 * All calls on one GPU share the array: calls that overlap on one GPU (other
 * streams, other threads, other plans) race and must be serialised by the
 * caller. Nothing in the array needs clearing: each part is fully written.
 */
int chip_bp_noalloc(const chip_plan* plan, const float* sino, const float* maps, int64_t maps_slice_stride,
                    float* vol, int batch, int mode, void* stream);

/*
 * Single-angle variants: line[d][u] for angle a (fp), and the back-projection
 * of line[d][u] at angle a (bp, accumulate = 0 overwrites vol, 1 adds).
 * maps points at the full [A][D][n][n] set. Replace fp_angle / bp_angle
 * (projectors.py:259-300, 502-568).
 */
int chip_fp_angle(const chip_plan* plan, const float* vol, const float* maps, float* line, int angle,
                  void* stream);
int chip_bp_angle(const chip_plan* plan, const float* line, const float* maps, float* vol, int angle,
                  int mode, void* stream);

/*
 * On-the-fly attenuation: fp / bp with the maps evaluated from mu_in and
 * mu_out (batch slices of n*n each, either may be NULL) instead of read from
 * a precomputed [A][D][n][n] set. Angles run in blocks; each block's maps are
 * built by the same kernels as chip_att_maps into plan-owned scratch bounded
 * by the "otf_budget_mib" tune key (default 4096), then projected, so the
 * results equal chip_fp / chip_bp over chip_att_maps' output. For
 * configurations whose maps do not fit in memory (the fused on-the-fly mode
 * of AttenuationVolume.compute_maps, physics/attenuation.py:122-180, with
 * ProjectorAttenuationXRF.fp / bp, projectors.py:570-622). chip_bp_otf takes
 * mode CHIP_BP_SKIMAGE_SYM or CHIP_BP_ADJOINT (CHIP_BP_SKIMAGE reads no maps:
 * use chip_bp). Calls on one plan must not overlap: they share its scratch.
 */
int chip_fp_otf(chip_plan* plan, const float* vol, const float* mu_in, const float* mu_out, float* sino,
                int batch, void* stream);
int chip_bp_otf(chip_plan* plan, const float* sino, const float* mu_in, const float* mu_out, float* vol,
                int batch, int mode, void* stream);

/*
 * Kernel selection / tile parameters of a plan, set by the autotuner or by
 * hand; no reference counterpart (the reference has one CPU code path).
 * Keys: "fp_kernel" = 0 (per-sample gather through the caches),
 * 1 (LDS-staged whole-sinogram kernel), 2 (LDS-staged with register
 * prefetch), 3 (streamed LDS tiles of 64 bins x 32 rows, the detectors'
 * products interleaved), 4 (3 staging only the footprint of interior chunks:
 * less HBM traffic, no faster), 5 (3 with 32-bin tiles: a third of the LDS,
 * twice the workgroups per CU) or 6 (workgroups walking
 * image-space bands of 16 rows or 32 columns that stage each map value once:
 * 27 % less HBM traffic than 5, slower at 512^2), 7 (5 with the slice and map
 * boxes staged by gfx950's direct-to-LDS buffer loads, the products formed at
 * the gathers; the same bits as 5), 8 (7 staging only the footprint of
 * interior chunks) or 9 (5 with each wave gathering 8 x 8 sample patches; sums
 * in another order) or 10 (the default: workgroups owning a 64 x 32 image
 * tile of one or two slices and a group of angles, each map value read once,
 * partial line sums per tile added in a fixed order by a second kernel; a bin
 * table per plan built by the first whole-sinogram fp, partial sums in plan
 * scratch, so under stream capture it needs one earlier eager call; results
 * within 1e-6 of 5); 2..10 need N % 4 == 0 and 16-byte aligned vol / maps,
 * else 1 runs (10 with more than 8192 tiles per slice runs 5). The shape of
 * 10: "fp_tile_kind" -1..3 (0: the next angle's map tiles fetched into
 * registers, two samples per lane and trip; 1: the same, one sample; 2 / 3: map
 * tiles loaded at staging, two / one; -1, default: 1 in two-detector passes, 3
 * in one-detector passes), "fp_tile_group" (angles per
 * workgroup, default 8), "fp_tile_pairs" (1: two slices of a batch per
 * workgroup, default); every shape gives the same bits. "bp_parts" = -1 (default: from
 * the adjoint bp kernel's occupancy), 0 (never split into parts) or 2..8
 * (pinned: the same bits on every device). "otf_budget_mib" bounds the
 * on-the-fly scratch.
 */
int chip_plan_tune(chip_plan* plan, const char* key, int value);

/* ---- solver building blocks (solvers.py:60-117, 688-1206) ------------------
 * n = number of elements. Optional arrays may be NULL. All elementwise
 * updates are fused single-pass kernels; reductions are deterministic.
 */

/* out[0] = max_i |v[i]| (one float, device). Replaces np.max(np.abs(.)) in
 * compute_diagonal_scaling (solvers.py:98-116). */
int chip_absmax(const float* v, int64_t n, float* out, void* stream);

/* v[i] = rho / fix(|v[i]| + add), fix(t) = 1 where t / max < 1e-5, with
 * max = maxval[0] + add and maxval the device scalar chip_absmax(v) wrote.
 * compute_diagonal_scaling (solvers.py:98-116). */
int chip_precond_fix(float* v, int64_t n, float add, const float* maxval, float rho, void* stream);
/* y[i] += a * x[i % x_period]: an array-valued regularizer step-size term (tau * upd_mask,
 * regularizers.py:260-273) added to tau in compute_diagonal_scaling / compute_Lipschitz_scaling
 * (solvers.py:99-101, 150-153). */
int chip_axpy(float* y, float a, const float* x, int64_t x_period, int64_t n, void* stream);
/* v[i] = rho / v[i]: compute_Lipschitz_scaling's tau = relaxation_tau / tau for an array tau
 * (solvers.py:155). */
int chip_rdiv(float* v, float rho, int64_t n, void* stream);

/* p = prox_l2(p + sigma * (ax + bg)) * mask, prox_l2(p) = (p - sigma*b) / (1 + sigma).
 * DataFidelity_l2.update_dual / apply_proximal_dual (data_terms.py:217-230,
 * 294-304) and the b_mask product (solvers.py:1172-1173). */
int chip_pdhg_dual_l2(float* p, const float* ax, const float* sigma, const float* b, const float* bg,
                      const float* b_mask, int64_t n, void* stream);

/* p = prox_KL(p + sigma * (ax + bg)) * mask,
 * prox_KL(p) = (1 + p - sqrt((p - 1)^2 + 4 sigma max(b, 0))) / 2.
 * DataFidelity_KL (data_terms.py:677-692). */
int chip_pdhg_dual_kl(float* p, const float* ax, const float* sigma, const float* b, const float* bg,
                      const float* b_mask, int64_t n, void* stream);

/* TV2D dual on `batch` n x n slices: q += 0.5 grad(x); q <- lam q / max(|q|_2, lam).
 * q is [batch][2][n][n]. Regularizer_Grad.update_dual + DataFidelity_l12
 * prox (regularizers.py:134-162; data_terms.py:549-555, 607-608;
 * operators.py:796-816 edge padding). */
int chip_tv_dual(float* q, const float* x, int n, int batch, float lam, void* stream);

/* Dual variable consistent with a primal estimate x (for a warm start; no
 * reference counterpart: corrct starts the duals at 0): the data term's
 * gradient at ax = A x, p = (ax + bg) - b (l2) or p = 1 - b / (ax + bg) (kl = 1;
 * 0 where ax + bg <= 0), times b_mask. */
int chip_pdhg_dual_init(float* p, const float* ax, const float* b, const float* bg, const float* b_mask, int64_t n,
                        int kl, void* stream);

/* PDHG primal: upd = atp - div(q) (q may be NULL); x_new = clip(x - tau upd) * mask;
 * xbar = x_new + (x_new - x); x = x_new. tau is elementwise. lower/upper are
 * used when has_lower / has_upper. solvers.py:1175-1190, operators.py:818-844. */
int chip_pdhg_primal(float* x, float* xbar, const float* atp, const float* q, const float* tau,
                     int has_lower, float lower, int has_upper, float upper, const float* x_mask, int n,
                     int batch, void* stream);

/* SIRT residual r = (b - (ax + bg)) * mask * sigma  (written to `res`).
 * DataFidelityBase.compute_residual (data_terms.py:136-161), solvers.py:954, 975. */
int chip_sirt_residual(float* res, const float* b, const float* ax, const float* bg, const float* b_mask,
                       const float* sigma, int64_t n, void* stream);

/* PDHG data dual of the norm-type terms: v = p + sigma (ax + bg) - sigma b, then
 * apply_proximal_dual of `term` (update_dual data_terms.py:217-230), times b_mask
 * (solvers.py:1172-1173):
 *   L2     v / (1 + sigma)
 *   WL2    v (1 - z) / (1 + sigma / (w + z)),  z = (w == 0)       (data_terms.py:353-354, 301-304)
 *   L2B    soft(v, sigma sqrt(e)) / (1 + sigma e)                (data_terms.py:410-415, 425-429)
 *   HUBER  u = v / (1 + sigma e);  u / max(1, |u|)               (data_terms.py:480-501)
 *   L1     v / max(|v|, 1)                                       (data_terms.py:549-555)
 *   L12    v / max(|v|_2 over the group, 1)                      (data_terms.py:607-608)
 *   L1B    s = soft(v, e);  s / max(|s|, 1)                      (data_terms.py:644-645, 549-555)
 * soft(v, t) = sign(v) max(|v| - t, 0). Groups (L12, and HUBER with an l2_axis)
 * hold group_len elements group_stride apart: element (o, g, i) sits at
 * (o * group_len + g) * group_stride + i, n a multiple of group_len * group_stride.
 * param: w or e as the term needs it (NULL otherwise), param_period >= 1 dividing n.
 * KL has its own entry (chip_pdhg_dual_kl). */
int chip_pdhg_dual_norm(float* p, const float* ax, const float* sigma, const float* b, const float* bg,
                        const float* b_mask, int64_t n, int term, const float* param, int64_t param_period,
                        int64_t group_len, int64_t group_stride, void* stream);

/* res = compute_residual(ax) * scale for the terms SIRT and FISTA accept
 * (l2-based and KL; Solver._initialize_data_fidelity_function, solvers.py:299-310),
 * with a = ax + bg and the term's assigned sigma `asig` (NULL = 1):
 *   L2   (b - a) mask                                  (data_terms.py:136-161)
 *   WL2  (b - a) w mask                                (data_terms.py:356-365)
 *   L2B  soft((b - a) mask, asig sqrt(e))              (data_terms.py:419-423)
 *   KL   -prox_KL(max(a, eps) asig) mask               (data_terms.py:721-740)
 * `scale` (NULL = 1) is the solver's sigma of upd = A^T(res * sigma)
 * (solvers.py:975, 1471). */
int chip_data_residual(float* res, const float* b, const float* ax, const float* bg, const float* b_mask,
                       const float* asig, const float* scale, int64_t n, int term, const float* param,
                       int64_t param_period, void* stream);

/* SIRT update x = clip(x + (upd - reg_upd) * tau) * mask, reg_upd = -div(q)
 * of a TV dual q (NULL = none). solvers.py:975-982. */
int chip_sirt_update(float* x, const float* upd, const float* q, const float* tau, int has_lower,
                     float lower, int has_upper, float upper, const float* x_mask, int n, int batch,
                     void* stream);

/* FBP line filter (Filter.apply_filter with zero padding, filters.py:346-390):
 * out[l][i] = sum_{j<n} in[l][j] h[(i - j) mod p] for n_lines lines of n samples,
 * h = the real-space kernel (irfft of the half-spectrum filter, length p, a power
 * of two with 2 n <= p <= 4096); float64 sums. */
int chip_filter_lines(const float* in, float* out, const double* h, int64_t n_lines, int n, int p, void* stream);

/* Detector PSF on each line of n samples (TransformConvolution with edge
 * padding, operators.py:424-496, as ProjectorUncorrected applies it after
 * fp, projectors.py:229-244, 275-321): out[l][i] = sum_k xe[i + (K-1)/2 - k] kern[k], xe the
 * line extended by (K-1)/2 edge copies on each side and zero beyond. */
int chip_psf_lines(const float* in, float* out, const float* kern, int k_len, int64_t n_lines, int n,
                   void* stream);

/*
 * Gradient-family regularizers over the last nd (1..3) axes, sizes dims[nd]
 * (slowest first), of a primal laid out [n_chan][n_batch][dims...]; the dual
 * q is [nd][n_chan][n_batch][dims...] (TransformGradient.adj_shape,
 * operators.py:767-844).
 * strides: NULL when the gradient axes are the trailing nd axes in order (as
 * above); else the gradient axes are any nd distinct axes of a channel's block
 * of n_batch samples (C-contiguous), given by their sizes dims[k] and element
 * strides strides[k] in the order of Regularizer_Grad.axes (the dual's
 * component k differentiates along axis k; regularizers.py:247-258,
 * operators.py:781-816), and the block is not split into batches.
 * pad_mode: TransformGradient's np.pad mode (operators.py:796-835): 0 = edge
 * (the gradient's last difference is 0 and so is the divergence's first),
 * 1 = constant (last difference -x, first divergence term q), 2 = wrap,
 * 3 = reflect, 4 = symmetric (as edge).
 * chip_grad_dual: q += 0.5 * grad(x), then the dual proximal of `mode`: 0 = l12 (isotropic TV,
 * Regularizer_TV1D/2D/3D, regularizers.py:218-337 with data_terms.py:598-608),
 * 1 = l1 (anisotropic), 2 = VTV with point-wise derivative and channel norms
 * der_norm / chan_norm in {1, 2, 0 = inf} (Regularizer_VTV.apply_proximal_dual,
 * regularizers.py:1362-1430), 3 = TNV (singular values of each C x nd matrix
 * clipped at 1, Regularizer_TNV with DataFidelity_ln, regularizers.py:1324-1359,
 * data_terms.py:781-820), 4 = l2 (q *= weight, the caller passing
 * 1 / (1 + 0.5): Regularizer_smooth1D/2D/3D, regularizers.py:386-433 with
 * data_terms.py:294-304), 5 = Huber over the gradient components (q *= weight,
 * the caller passing 1 / (1 + 0.5 huber_size); q /= max(1, |q|_2):
 * Regularizer_HubTV2D/3D, regularizers.py:340-384 with data_terms.py:482-501),
 * 6 = Huber without an l2 axis (q *= weight as in 5, then each component
 * q /= max(1, |q|), data_terms.py:497-498).
 * Modes 0, 1, 4, 5 and 6 need n_chan == 1; n_chan <= 16.
 * weights (modes 0, 1, 2; else NULL): the regularizer's weight as an array
 * broadcast over the primal, weights[e % weights_period] for primal element e
 * (weights_period divides the primal's element count); it replaces `weight` in
 * the prox of every dual component at that sample (the reference's array
 * weight broadcast against the dual, data_terms.py:549-555,
 * regularizers.py:1413-1430).
 * qref (VTV only, else NULL): the dual of Regularizer_VTV's reference volume,
 * [nd][n_batch][dims...], a fixed extra channel after the n_chan channels in
 * every point-wise norm, never updated (regularizers.py:1386-1392, 1410-1411).
 * chip_grad_update_primal: upd += t (or -= t when subtract) with
 * t = -div(q), times mask[e % mask_period] when mask is not NULL (the
 * regularizer's upd_mask, broadcast over leading axes: mask_period divides
 * the primal's element count), times scale (or scales[e % scales_period]
 * when scales is not NULL: an array weight) when apply_scale
 * (compute_update_primal, regularizers.py:194-210; PDHG adds, SIRT
 * subtracts, solvers.py:953-958, 1178-1180).
 */
int chip_grad_dual(float* q, const float* x, int n_chan, int64_t n_batch, int nd, const int* dims,
                   const int64_t* strides, int pad_mode, int mode, float weight, const float* weights,
                   int64_t weights_period, int der_norm, int chan_norm, const float* qref, void* stream);
int chip_grad_update_primal(float* upd, const float* q, int n_chan, int64_t n_batch, int nd, const int* dims,
                            const int64_t* strides, int pad_mode, float scale, const float* scales,
                            int64_t scales_period, int apply_scale, int subtract, const float* mask,
                            int64_t mask_period, void* stream);

/* Regularizer_l1 in PDHG / SIRT (identity transform, regularizers.py:531-557 and
 * the base class's dual / primal steps, regularizers.py:134-210):
 * chip_l1_dual: q = q + x, then q = (q / max(|q|, w)) * w with w = weight, or
 * weights[i % weights_period] when weights is not NULL (data_terms.py:549-555).
 * chip_l1_update_primal: upd += q (or -= q when subtract), times
 * mask[i % mask_period] when mask is not NULL (upd_mask). */
int chip_l1_dual(float* q, const float* x, int64_t n, float weight, const float* weights, int64_t weights_period,
                 void* stream);
int chip_l1_update_primal(float* upd, const float* q, int64_t n, const float* mask, int64_t mask_period, int subtract,
                          void* stream);

/* FISTA step (solvers.py:1458-1490): x_new = y + tau * grad (tau per element,
 * or tau_scalar when tau is NULL); then the l1 prox with weight w = l1_weight, or
 * l1_weights[i % l1_weights_period] when l1_weights is not NULL, when l1_weight > 0
 * or l1_weights is given: x_new = soft(x_new, tau * w), or with an update mask m =
 * l1_mask[i % l1_mask_period], x_new = x_new (1 - m) + soft(x_new m, tau * w)
 * (Regularizer_l1.apply_proximal_primal, regularizers.py:559-585); then
 * clipped, masked; then y = x_new + momentum (x_new - x) and x = x_new. */
int chip_fista_update(float* x, float* y, const float* grad, const float* tau, float tau_scalar, float l1_weight,
                      const float* l1_weights, int64_t l1_weights_period, const float* l1_mask,
                      int64_t l1_mask_period, int has_lower, float lower, int has_upper, float upper,
                      const float* x_mask, float momentum, int64_t n, void* stream);

/* Array-valued clipping (x_new.clip(lower_limit, upper_limit) with array limits,
 * solvers.py:797-803, 975-982, 1185-1190): x = clip(x, lower, upper) * x_mask,
 * each limit an array of n elements (lower / upper non-NULL), a scalar (has_* with
 * a NULL array) or absent; then, when xbar is given, xbar = x + momentum (x - x_old)
 * (PDHG's relaxation with momentum 1, FISTA's extrapolation). The fused update
 * kernels run without limits and mask first, so the order of operations is the
 * reference's. */
int chip_clip_relax(float* x, const float* lower, const float* upper, int has_lower, float lower_s, int has_upper,
                    float upper_s, const float* x_mask, float* xbar, const float* x_old, float momentum, int64_t n,
                    void* stream);

/* MLEM ratio r = b / max(ax + bg, eps_f32). solvers.py:793-797. */
int chip_mlem_ratio(float* r, const float* b, const float* ax, const float* bg, int64_t n, void* stream);

/* MLEM update x = clip(x * (upd / tau)) * mask. solvers.py:797-803. */
int chip_mlem_update(float* x, const float* upd, const float* tau, int has_lower, float lower,
                     int has_upper, float upper, const float* x_mask, int64_t n, void* stream);

/* Residual norms (double accumulation, deterministic): out[0] = sum r^2 (l2,
 * data_terms.py:298-299) or sum |r| (KL/l1, data_terms.py:741-742). */
int chip_sumsq(const float* v, int64_t n, double* out, void* stream);
int chip_sumabs(const float* v, int64_t n, double* out, void* stream);

/* ---- batches of slices of different sizes ----------------------------------
 * corrct reconstructs a stack slice by slice (ProjectorAttenuationXRF over a
 * 3-D volume, _projector_backends.py:86-92 for the maps); a batch plan runs
 * many independent slices, each of its own size N_s (N_s % 4 == 0), as one
 * persistent launch per pass. Buffers hold the slices back to back in plan
 * order: volumes [N_s][N_s], sinograms [D][A][N_s], maps [A][D][N_s][N_s],
 * TV duals [2][N_s][N_s] (offsets = running sums; chip_batch_info gives the
 * totals). Angles, detectors and weights are shared by all slices, with the
 * meaning of chip_plan_create.
 */
int chip_batch_plan_create(chip_batch_plan** out, int n_slices, const int* sizes, int n_angles,
                           const float* angles_rad, int n_det, const double* det_angles_rad, const double* w_det,
                           const double* w_ang);
int chip_batch_plan_destroy(chip_batch_plan* plan);
/* chip_plan_set_radon_trig for every slice size of a batch plan. */
int chip_batch_set_radon_trig(chip_batch_plan* plan, const float* cos_sin);
/* Total elements of the packed volume, sinogram and maps buffers (NULL outputs are skipped). */
int chip_batch_info(const chip_batch_plan* plan, int64_t* vol_elems, int64_t* sino_elems, int64_t* maps_elems);
/* "fp_tile" (1, default: the image-tile fp of chip_plan_tune's fp_kernel 10,
 * one launch per slice size over the batch-plan's slots with its partial sums
 * in batch-plan scratch grown on first use; 0: the angle x bin-block kernels
 * below) with its shape keys "fp_tile_kind", "fp_tile_group", "fp_tile_pairs";
 * "bp_pairs" (1, default: the exact-adjoint bp runs two slices of a size per
 * workgroup, one launch per size; 0: one slice per workgroup in one launch over
 * every slice; the same bits), "bp_quads" (1, default: in one-detector passes
 * four slices of a size per workgroup first; the same bits) and
 * "bp_pair_parts" (0, default: the angle
 * range of a pair split into parts by the kernel's occupancy, each part summed
 * into batch-plan scratch grown on first use, then added in part order, which
 * changes the grouping of the sum; 1..8 pins the count, 1 gives the bits of
 * one workgroup per slice). Grid sizes of the one-launch kernels:
 * "fp_wg_per_cu" (0..8, default 5) and "bp_wg_per_cu" (0..16, default 4)
 * workgroups per CU of a persistent grid, 0 = one workgroup per work item; fp
 * variants of fp_tile 0: "fp_trim" (stage only the footprint of interior
 * chunks), "fp_direct" (direct-to-LDS staging) and "fp_patch" (8 x 8 sample
 * patches per wave), each 0 / 1, default 0; set by the autotuner.
 * "otf_budget_mib" (>= 16, default 4096) bounds the on-the-fly projectors'
 * scratch. */
int chip_batch_tune(chip_batch_plan* plan, const char* key, int value);
/* Maps of every slice (chip_att_maps per slice; mu_* packed like volumes), through a prefix-sum
 * scratch of at most 1 GiB shared by all sizes. */
int chip_batch_att_maps(chip_batch_plan* plan, const float* mu_in, const float* mu_out, float* maps, void* stream);
/* fp / bp of every slice (chip_fp / chip_bp semantics per slice); maps may be NULL. */
int chip_batch_fp(const chip_batch_plan* plan, const float* vol, const float* maps, float* sino, void* stream);
int chip_batch_bp(const chip_batch_plan* plan, const float* sino, const float* maps, float* vol, int mode,
                  void* stream);
/* fp / bp of every slice with its maps evaluated on the fly from its own mu_in / mu_out (packed
 * like volumes, either may be NULL), a block of angles at a time (chip_fp_otf / chip_bp_otf per
 * slice). All slice sizes share one scratch (prefix sums + maps block) of at most about the batch's
 * "otf_budget_mib" (default 4096), grown on the first call. For batches whose maps do not fit in
 * memory (a few hundred slices of up to 768^2 at 360 angles need about 320 GB of maps). The fp
 * equals chip_batch_fp over chip_batch_att_maps' maps bit for bit with the default fp tiles; the
 * bp agrees to rounding (a lone slice's adjoint bp may split its angle range into parts). Mode
 * CHIP_BP_SKIMAGE_SYM or CHIP_BP_ADJOINT. Calls on one batch plan must not overlap: they share
 * its scratch. */
int chip_batch_fp_otf(chip_batch_plan* plan, const float* vol, const float* mu_in, const float* mu_out,
                      float* sino, void* stream);
int chip_batch_bp_otf(chip_batch_plan* plan, const float* sino, const float* mu_in, const float* mu_out,
                      float* vol, int mode, void* stream);

/* The fused on-the-fly iteration step of chip_otf_fp_row_bp on every slice of
 * the batch (packed like chip_batch_fp / chip_batch_bp; no background or mask):
 * per slice and angle block, the maps once, the fp of x into ax, row_op on p's
 * rows, the bp of p's rows into vol. Equals chip_batch_fp_otf, the elementwise
 * kernel and chip_batch_bp_otf in turn, bit for bit. No reference counterpart. */
int chip_batch_otf_fp_row_bp(chip_batch_plan* plan, const float* x, const float* mu_in, const float* mu_out,
                             int row_op, float* p, const float* sigma, const float* b, float* ax, float* vol,
                             int bp_mode, void* stream);
/* Spatial solver updates per slice (chip_tv_dual, chip_pdhg_primal,
 * chip_sirt_update semantics); x, tau, masks packed like volumes. */
int chip_batch_tv_dual(const chip_batch_plan* plan, float* q, const float* x, float lam, void* stream);
int chip_batch_pdhg_primal(const chip_batch_plan* plan, float* x, float* xbar, const float* atp, const float* q,
                           const float* tau, int has_lower, float lower, int has_upper, float upper,
                           const float* x_mask, void* stream);
int chip_batch_sirt_update(const chip_batch_plan* plan, float* x, const float* upd, const float* q, const float* tau,
                           int has_lower, float lower, int has_upper, float upper, const float* x_mask,
                           void* stream);

/* ---- warm-start network (no reference counterpart) --------------------------
 * Building blocks of the U-Net that seeds PDHG with an estimate of its
 * solution from a short SIRT reconstruction. Activations are NHWC bfloat16
 * (uint16_t bit patterns), conv weights [9][Cout][Cin] bfloat16 (tap =
 * 3 (dh + 1) + (dw + 1)), biases, master weights and gradients float32; the
 * 3x3 convolutions run on the matrix cores with float32 accumulation.
 * Reductions use caller-provided float32 scratch `part` of `part_elems`
 * elements and add partial sums in a fixed order.
 */
/* y = act(bias + conv3x3(concat(x1[c1], x2[c2]))), "same" zero padding; the
 * first o1 output channels go to y1, the rest to y2 (NULL when o1 = cout). */
int chip_nn_conv3x3(const uint16_t* x1, int c1, const uint16_t* x2, int c2, const uint16_t* wt, const float* bias,
                    uint16_t* y1, int o1, uint16_t* y2, int n, int h, int w, int cout, int relu, void* stream);
/* dw[tap][co][ci] = sum_p dy[p][co] * concat(x1, x2)[p + offset(tap)][ci] */
int chip_nn_conv3x3_wgrad(const uint16_t* dy, const uint16_t* x1, int c1, const uint16_t* x2, int c2, int n, int h,
                          int w, int cout, float* part, int64_t part_elems, float* dw, void* stream);
/* db[ch] = sum_p dy[p][ch] */
int chip_nn_bias_grad(const uint16_t* dy, int64_t n_pix, int c, float* part, int64_t part_elems, float* db,
                      void* stream);
/* wt_t[tap][ci][co] = wt[8 - tap][co][ci]: weights of the data-gradient convolution */
int chip_nn_flip_weights(const uint16_t* wt, int cout, int cin, uint16_t* wt_t, void* stream);
/* d[k] = 0 where y[k] <= 0 (gradient through a ReLU whose output is y) */
int chip_nn_relu_mask(uint16_t* d, const uint16_t* y, int64_t count, void* stream);
/* 2 x 2 average pooling of an n x h x w x c input (backward = 1: x is dy of
 * the pooled shape, y is dx of the input shape) */
int chip_nn_avgpool2(const uint16_t* x, uint16_t* y, int n, int h, int w, int c, int backward, void* stream);
/* 2x nearest upsampling of an n x h x w x c input (backward = 1: x is dy of the
 * upsampled shape, y is dx of the input shape) */
int chip_nn_upsample2(const uint16_t* x, uint16_t* y, int n, int h, int w, int c, int backward, void* stream);
/* out[p] = xin[p] + wb[c] + sum_ci wb[ci] feat[p][ci] (1 x 1 head on a residual, float32) */
int chip_nn_head_fwd(const uint16_t* feat, int c, const float* wb, const float* xin, float* out, int64_t n_pix,
                     void* stream);
/* dfeat = dout x wb[:c]; dwb[ci] = sum_p dout[p] feat[p][ci], dwb[c] = sum_p dout[p] */
int chip_nn_head_bwd(const float* dout, const uint16_t* feat, int c, const float* wb, uint16_t* dfeat, float* part,
                     int64_t part_elems, float* dwb, int64_t n_pix, void* stream);
/* loss[0] = mean (out - tgt)^2 (double), dout = 2 (out - tgt) / n_pix; part: >= 1 doubles */
int chip_nn_mse(const float* out, const float* tgt, int64_t n_pix, float* dout, double* part, int part_elems,
                double* loss, void* stream);
/* Adam step `step` (>= 1) on float32 weights w with gradient g and moments m, v;
 * w_bf16 (nullable) receives the bfloat16 copy */
int chip_nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
                 float eps, int step, uint16_t* w_bf16, void* stream);
/* y = bf16(y + x) (gradients of a tensor with two consumers) */
int chip_nn_add_bf16(uint16_t* y, const uint16_t* x, int64_t n, void* stream);
int chip_nn_to_bf16(const float* x, uint16_t* y, int64_t n, void* stream);
int chip_nn_to_f32(const uint16_t* x, float* y, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CORRCT_HIP_H */
