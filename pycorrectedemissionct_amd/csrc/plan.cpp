// Geometry plan, error strings and symbol table of libcorrct_hip.
//
// The per-angle tables reproduce how corrct + scikit-image 0.18.3 derive
// their float32 inverse maps from the float32 rotation angles:
//   radon:  deg = rad2deg(float32 angle)  (_projector_backends.py:148, 299)
//           theta = deg2rad(deg) in float32, cos/sin of theta, and
//           R = [[c, s, -k(c + s - 1)], [-s, c, -k(c - s - 1)]], k = N // 2,
//           evaluated in float32 (skimage radon_transform.py:104-112).
//   rotate: tform = T(-ctr) + R(r) + T(ctr) composed in float64, ctr =
//           (N/2 - 0.5, N/2 - 0.5), then cast to float32
//           (skimage _warps.py:366-404, _geometric.py:705-716, 1119-1148).
//   maps:   in-beam angles stay float32 (attenuation.py:105-106); out-beam
//           angles are theta_a + phi_d (float64) + pi (attenuation.py:108-120,
//           _projector_backends.py:79-80).
#include <cmath>
#include <cstring>
#include <mutex>

#include "common.hpp"

namespace chip {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }

// numpy float32 rad2deg / deg2rad (multiply by a float32 constant)
static inline float rad2deg_f32(float x) {
  const float k = 180.0f / static_cast<float>(M_PI);
  return x * k;
}
static inline float deg2rad_f32(float x) {
  const float k = static_cast<float>(M_PI / 180.0);
  return x * k;
}

// tform = T(-ctr) + R(r) + T(ctr) -> T1 @ (R @ T3), float64, then float32
static Affine rotate_affine(int side, double r) {
  const double cx = side / 2.0 - 0.5;
  const double cy = side / 2.0 - 0.5;
  const double c = std::cos(r), s = std::sin(r);
  // R @ T3 with T3 = [[1,0,-cx],[0,1,-cy]]
  const double a02 = c * (-cx) + (-s) * (-cy);
  const double a12 = s * (-cx) + c * (-cy);
  Affine m;
  m.m00 = static_cast<float>(c);
  m.m01 = static_cast<float>(-s);
  m.m02 = static_cast<float>(a02 + cx);
  m.m10 = static_cast<float>(s);
  m.m11 = static_cast<float>(c);
  m.m12 = static_cast<float>(a12 + cy);
  return m;
}

// radon inverse-map table h_fp_geom[a] = (c, s, -k (c + s - 1), -k (c - s - 1)) in float32
// from the angle's float32 (cos, sin)
static void fill_radon_geom(Plan& p, int a, float c, float s) {
  const float kf = static_cast<float>(-(p.n / 2));
  p.h_fp_geom[4 * a + 0] = c;
  p.h_fp_geom[4 * a + 1] = s;
  p.h_fp_geom[4 * a + 2] = kf * ((c + s) - 1.0f);
  p.h_fp_geom[4 * a + 3] = kf * ((c - s) - 1.0f);
}

// device copies of the radon table and of its row terms (fl(s y), fl(c y)), synchronous
static hipError_t upload_radon_geom(Plan& p) {
  hipError_t e;
  const int n = p.n, n_angles = p.n_angles;
  if (!p.d_fp_geom && (e = hipMalloc(&p.d_fp_geom, sizeof(float4) * n_angles)) != hipSuccess) return e;
  if ((e = hipMemcpy(p.d_fp_geom, p.h_fp_geom.data(), sizeof(float4) * n_angles, hipMemcpyHostToDevice)) != hipSuccess)
    return e;
  // row terms of the fp inverse map, one float32 product each (as the kernels form them)
  std::vector<float2> ytab(static_cast<size_t>(n_angles) * n);
  for (int a = 0; a < n_angles; ++a) {
    const float c = p.h_fp_geom[4 * a + 0], s = p.h_fp_geom[4 * a + 1];
    for (int y = 0; y < n; ++y) {
      const float yf = static_cast<float>(y);
      ytab[static_cast<size_t>(a) * n + y] = make_float2(s * yf, c * yf);
    }
  }
  if (!p.d_ytab && (e = hipMalloc(&p.d_ytab, sizeof(float2) * ytab.size())) != hipSuccess) return e;
  return hipMemcpy(p.d_ytab, ytab.data(), sizeof(float2) * ytab.size(), hipMemcpyHostToDevice);
}

}  // namespace chip

using namespace chip;

static const char* kSymbols[] = {
    "chip_version",       "chip_last_error",   "chip_num_symbols",  "chip_symbol_name",
    "chip_plan_create",   "chip_plan_destroy", "chip_att_maps",     "chip_bin_images",   "chip_fp",
    "chip_bp",            "chip_fp_angle",     "chip_bp_angle",     "chip_absmax",
    "chip_precond_fix",   "chip_pdhg_dual_l2", "chip_pdhg_dual_kl", "chip_pdhg_dual_norm",
    "chip_data_residual", "chip_l1_dual",      "chip_l1_update_primal", "chip_tv_dual",
    "chip_pdhg_primal",   "chip_sirt_residual", "chip_sirt_update", "chip_mlem_ratio",
    "chip_mlem_update",   "chip_sumsq",        "chip_sumabs",       "chip_plan_tune",
    "chip_fp_otf",        "chip_bp_otf",       "chip_otf_fp_row_bp",       "chip_batch_plan_create", "chip_batch_plan_destroy",
    "chip_batch_info",    "chip_batch_tune",   "chip_batch_att_maps", "chip_batch_otf_fp_row_bp", "chip_batch_fp",
    "chip_batch_bp",      "chip_batch_tv_dual", "chip_batch_pdhg_primal", "chip_batch_sirt_update",
    "chip_nn_conv3x3",    "chip_nn_conv3x3_wgrad", "chip_nn_bias_grad", "chip_nn_flip_weights",
    "chip_nn_relu_mask",  "chip_nn_avgpool2",  "chip_nn_upsample2", "chip_nn_head_fwd",
    "chip_nn_head_bwd",   "chip_nn_mse",       "chip_nn_adam",      "chip_nn_to_bf16",
    "chip_nn_to_f32",     "chip_nn_add_bf16",  "chip_pdhg_dual_init", "chip_fista_update",  "chip_filter_lines",  "chip_psf_lines",     "chip_plan_set_subangles", "chip_grad_dual", "chip_grad_update_primal", "chip_bp_noalloc", "chip_plan_set_radon_trig", "chip_batch_set_radon_trig", "chip_clip_relax", "chip_batch_fp_otf", "chip_batch_bp_otf", "chip_axpy", "chip_rdiv",
};

extern "C" {

int chip_version(void) { return 100; }

const char* chip_last_error(void) { return g_last_error.c_str(); }

int chip_num_symbols(void) { return static_cast<int>(sizeof(kSymbols) / sizeof(kSymbols[0])); }

const char* chip_symbol_name(int i) {
  if (i < 0 || i >= chip_num_symbols()) return nullptr;
  return kSymbols[i];
}

int chip_plan_create(chip_plan** out, int n, int n_angles, const float* angles_rad, int n_det,
                     const double* det_angles_rad, const double* w_det, const double* w_ang) {
  clear_error();
  CHIP_CHECK_ARG(out != nullptr, "out is NULL");
  *out = nullptr;
  CHIP_CHECK_ARG(n >= 2 && n <= 16384, "slice side must be in [2, 16384]");
  CHIP_CHECK_ARG(n_angles >= 1, "need at least one angle");
  CHIP_CHECK_ARG(angles_rad != nullptr, "angles_rad is NULL");
  CHIP_CHECK_ARG(n_det >= 1 && n_det <= 64, "detector count must be in [1, 64]");
  CHIP_CHECK_ARG(det_angles_rad != nullptr, "det_angles_rad is NULL");

  auto* plan = new (std::nothrow) chip_plan();
  if (!plan) {
    set_error("chip_plan_create: host allocation failed");
    return CHIP_ENOMEM;
  }
  Plan& p = plan->p;
  p.n = n;
  p.n_angles = n_angles;
  p.n_det = n_det;
  const double min_size = std::ceil(std::sqrt(2.0 * n * static_cast<double>(n)));
  p.pad = static_cast<int>(std::ceil((min_size - n) / 2.0));
  p.np = n + 2 * p.pad;
  (void)hipGetDevice(&p.device);

  p.h_fp_geom.resize(static_cast<size_t>(n_angles) * 4);
  p.h_w.resize(static_cast<size_t>(n_angles) * n_det);
  p.h_rot_in_fwd.resize(n_angles);
  p.h_rot_in_back.resize(n_angles);
  p.h_rot_out_fwd.resize(static_cast<size_t>(n_angles) * n_det);
  p.h_rot_out_back.resize(static_cast<size_t>(n_angles) * n_det);
  for (int a = 0; a < n_angles; ++a) {
    const float deg = rad2deg_f32(angles_rad[a]);
    const float th = deg2rad_f32(deg);
    // correctly rounded float32 cos / sin; chip_plan_set_radon_trig can substitute the
    // values of the caller's float32 math library
    fill_radon_geom(p, a, static_cast<float>(std::cos(static_cast<double>(th))),
                    static_cast<float>(std::sin(static_cast<double>(th))));
    for (int d = 0; d < n_det; ++d) {
      const double wd = w_det ? w_det[d] : 1.0;
      const double wa = w_ang ? w_ang[static_cast<size_t>(a) * n_det + d] : 1.0;
      p.h_w[static_cast<size_t>(a) * n_det + d] = static_cast<float>(wd * wa);
    }
    // in-beam: float32 angle through float32 rad2deg / deg2rad
    p.h_rot_in_fwd[a] = rotate_affine(p.np, static_cast<double>(deg2rad_f32(-deg)));
    p.h_rot_in_back[a] = rotate_affine(p.np, static_cast<double>(deg2rad_f32(deg)));
    // out-beam: float64 angle
    for (int d = 0; d < n_det; ++d) {
      const double phi = static_cast<double>(angles_rad[a]) + det_angles_rad[d] + M_PI;
      const double dg = phi * (180.0 / M_PI);
      p.h_rot_out_fwd[static_cast<size_t>(a) * n_det + d] = rotate_affine(p.np, (-dg) * (M_PI / 180.0));
      p.h_rot_out_back[static_cast<size_t>(a) * n_det + d] = rotate_affine(p.np, dg * (M_PI / 180.0));
    }
  }

  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string("chip_plan_create: ") + what + " -> " + hipGetErrorString(e));
    chip_plan_destroy(plan);
    return e == hipErrorOutOfMemory ? CHIP_ENOMEM : CHIP_EHIP;
  };
  hipError_t e;
  if ((e = upload_radon_geom(p)) != hipSuccess) return fail(e, "radon tables");
  if ((e = hipMalloc(&p.d_w, sizeof(float) * p.h_w.size())) != hipSuccess) return fail(e, "hipMalloc w");
  const size_t n_rot = 2 * static_cast<size_t>(n_angles) + 2 * static_cast<size_t>(n_angles) * n_det;
  if ((e = hipMalloc(&p.d_rot, sizeof(Affine) * n_rot)) != hipSuccess) return fail(e, "hipMalloc rot");
  std::vector<Affine> rot;
  rot.reserve(n_rot);
  rot.insert(rot.end(), p.h_rot_in_fwd.begin(), p.h_rot_in_fwd.end());
  rot.insert(rot.end(), p.h_rot_in_back.begin(), p.h_rot_in_back.end());
  rot.insert(rot.end(), p.h_rot_out_fwd.begin(), p.h_rot_out_fwd.end());
  rot.insert(rot.end(), p.h_rot_out_back.begin(), p.h_rot_out_back.end());
  if ((e = hipMemcpy(p.d_w, p.h_w.data(), sizeof(float) * p.h_w.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy w");
  if ((e = hipMemcpy(p.d_rot, rot.data(), sizeof(Affine) * n_rot, hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e, "hipMemcpy rot");
  *out = plan;
  return CHIP_OK;
}

int chip_plan_destroy(chip_plan* plan) {
  if (!plan) return CHIP_OK;
  Plan& p = plan->p;
  if (p.d_fp_geom) (void)hipFree(p.d_fp_geom);
  if (p.d_w) (void)hipFree(p.d_w);
  if (p.d_ytab) (void)hipFree(p.d_ytab);
  if (p.d_rot) (void)hipFree(p.d_rot);
  if (p.d_nsub) (void)hipFree(p.d_nsub);
  if (p.d_scratch) (void)hipFree(p.d_scratch);
  if (p.d_mblock) (void)hipFree(p.d_mblock);
  if (p.d_bp_part) (void)hipFree(p.d_bp_part);
  if (p.d_fp_tile) (void)hipFree(p.d_fp_tile);
  if (p.d_fp_tab) (void)hipFree(p.d_fp_tab);
  if (p.bp_state && p.bp_state->done) (void)hipEventDestroy(p.bp_state->done);
  if (p.fp_state && p.fp_state->done) (void)hipEventDestroy(p.fp_state->done);
  delete plan;
  return CHIP_OK;
}

int chip_plan_set_radon_trig(chip_plan* plan, const float* cos_sin) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && cos_sin != nullptr, "NULL argument");
  Plan& p = plan->p;
  for (int a = 0; a < p.n_angles; ++a) {
    const float c = cos_sin[2 * a], s = cos_sin[2 * a + 1];
    CHIP_CHECK_ARG(std::isfinite(c) && std::isfinite(s) && std::fabs(c) <= 1.f && std::fabs(s) <= 1.f,
                   "cos / sin entries must be finite and in [-1, 1]");
  }
  CHIP_HIP(hipDeviceSynchronize());  // queued work may still read the old tables
  for (int a = 0; a < p.n_angles; ++a) fill_radon_geom(p, a, cos_sin[2 * a], cos_sin[2 * a + 1]);
  CHIP_HIP(upload_radon_geom(p));
  p.fp_tab_ok = false;  // the image-tile fp's bin table follows the radon table
  return CHIP_OK;
}

int chip_plan_set_subangles(chip_plan* plan, int n_sub_max, const int* n_sub_det, const double* sub_angles,
                            const float* angles_rad, const double* det_angles_rad) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && n_sub_det != nullptr && sub_angles != nullptr, "NULL argument");
  CHIP_CHECK_ARG(angles_rad != nullptr && det_angles_rad != nullptr, "NULL angles");
  CHIP_CHECK_ARG(n_sub_max >= 1 && n_sub_max <= 64, "n_sub_max must be in [1, 64]");
  Plan& p = plan->p;
  const int A = p.n_angles, D = p.n_det, S = n_sub_max;
  for (int d = 0; d < D; ++d)
    CHIP_CHECK_ARG(n_sub_det[d] >= 1 && n_sub_det[d] <= S, "n_sub_det entries must be in [1, n_sub_max]");
  std::vector<Affine> fwd(static_cast<size_t>(A) * D * S), back(static_cast<size_t>(A) * D * S);
  for (int a = 0; a < A; ++a) {
    for (int d = 0; d < D; ++d) {
      // the reference's order: ((theta + phi_d) + sub) + pi, float64 (physics/attenuation.py:109-118,
      // _projector_backends.py:79-80)
      const double base = static_cast<double>(angles_rad[a]) + det_angles_rad[d];
      for (int k = 0; k < S; ++k) {
        const double sub = k < n_sub_det[d] ? sub_angles[static_cast<size_t>(d) * S + k] : 0.0;
        const double dg = ((base + sub) + M_PI) * (180.0 / M_PI);
        const size_t idx = (static_cast<size_t>(a) * D + d) * S + k;
        fwd[idx] = rotate_affine(p.np, (-dg) * (M_PI / 180.0));
        back[idx] = rotate_affine(p.np, dg * (M_PI / 180.0));
      }
    }
  }
  std::vector<Affine> rot;
  rot.reserve(2 * static_cast<size_t>(A) + 2 * fwd.size());
  rot.insert(rot.end(), p.h_rot_in_fwd.begin(), p.h_rot_in_fwd.end());
  rot.insert(rot.end(), p.h_rot_in_back.begin(), p.h_rot_in_back.end());
  rot.insert(rot.end(), fwd.begin(), fwd.end());
  rot.insert(rot.end(), back.begin(), back.end());
  Affine* d_rot = nullptr;
  int* d_nsub = nullptr;
  CHIP_HIP(hipDeviceSynchronize());  // the old tables may still be read by queued work
  CHIP_HIP(hipMalloc(&d_rot, sizeof(Affine) * rot.size()));
  if (hipMalloc(&d_nsub, sizeof(int) * D) != hipSuccess) {
    (void)hipFree(d_rot);
    set_error("chip_plan_set_subangles: hipMalloc failed");
    return CHIP_ENOMEM;
  }
  CHIP_HIP(hipMemcpy(d_rot, rot.data(), sizeof(Affine) * rot.size(), hipMemcpyHostToDevice));
  CHIP_HIP(hipMemcpy(d_nsub, n_sub_det, sizeof(int) * D, hipMemcpyHostToDevice));
  if (p.d_rot) (void)hipFree(p.d_rot);
  if (p.d_nsub) (void)hipFree(p.d_nsub);
  p.d_rot = d_rot;
  p.d_nsub = d_nsub;
  p.h_rot_out_fwd.swap(fwd);
  p.h_rot_out_back.swap(back);
  p.n_sub = S;
  return CHIP_OK;
}

}  // extern "C"
