// Batches of slices of different sizes (gfx950).
//
// A batch plan packs n_slices slices back to back in every buffer kind:
// volumes [n][n], sinograms [D][A][n], maps [A][D][n][n], TV duals [2][n][n],
// each slice at the running sum of its predecessors' sizes. All slices share
// the angles, detectors and weights; slices of equal size share one
// single-slice plan (geometry tables, map-builder scratch).
//
// fp and bp run as one persistent launch per detector pass over the work
// items of every slice (fp: angle x 64-bin block, bp: pixel tile); the
// workgroup count is (workgroups per CU) x (CUs), set by chip_batch_tune.
// The spatial elementwise updates (TV dual, PDHG primal, SIRT update) find
// each pixel's slice in the pixel prefix sums.
#include <algorithm>
#include <cfloat>
#include <map>

#include "common.hpp"

struct chip_batch_plan {
  chip::BatchPlan b;
};

namespace chip {
namespace {

constexpr int kEw = 256;

__device__ __forceinline__ float clip_value(float v, int hl, float lo, int hu, float hi) {
  if (hl) v = fmaxf(v, lo);
  if (hu) v = fminf(v, hi);
  return v;
}

struct PixelRef {
  int n, i, j;
  int64_t off;  // slice's pixel offset
};

__device__ __forceinline__ PixelRef locate(const BatchSlice* __restrict__ slices, const int64_t* __restrict__ vstart,
                                           int n_slices, int64_t k) {
  const int s = find_slice(vstart, n_slices, k);
  PixelRef r;
  r.n = slices[s].n;
  r.off = vstart[s];
  const int pix = static_cast<int>(k - r.off);
  r.i = pix / r.n;
  r.j = pix - r.i * r.n;
  return r;
}

// q <- proj_lambda(q + 1/2 grad x) per slice (forward differences, edge-padded)
__global__ __launch_bounds__(kEw) void tv_dual_batch_kernel(float* __restrict__ q, const float* __restrict__ x,
                                                            const BatchSlice* __restrict__ slices,
                                                            const int64_t* __restrict__ vstart, int n_slices,
                                                            float lam) {
  const int64_t total = vstart[n_slices];
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kEw + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * kEw) {
    const PixelRef r = locate(slices, vstart, n_slices, k);
    const int64_t nn = static_cast<int64_t>(r.n) * r.n;
    const float xc = x[k];
    const float g0 = (r.i + 1 < r.n) ? x[k + r.n] - xc : 0.f;
    const float g1 = (r.j + 1 < r.n) ? x[k + 1] - xc : 0.f;
    float* q0 = q + 2 * r.off + (k - r.off);
    float* q1 = q0 + nn;
    const float a0 = *q0 + 0.5f * g0;
    const float a1 = *q1 + 0.5f * g1;
    const float den = fmaxf(sqrtf(a0 * a0 + a1 * a1), lam);
    *q0 = (a0 / den) * lam;
    *q1 = (a1 / den) * lam;
  }
}

__device__ __forceinline__ float neg_div_at(const float* __restrict__ q, const PixelRef& r, int64_t k) {
  const int64_t nn = static_cast<int64_t>(r.n) * r.n;
  const float* q0 = q + 2 * r.off + (k - r.off);
  const float* q1 = q0 + nn;
  const float d0 = (r.i > 0) ? q0[0] - q0[-r.n] : 0.f;
  const float d1 = (r.j > 0) ? q1[0] - q1[-1] : 0.f;
  return -(d0 + d1);
}

__global__ __launch_bounds__(kEw) void pdhg_primal_batch_kernel(float* __restrict__ x, float* __restrict__ xbar,
                                                                const float* __restrict__ atp,
                                                                const float* __restrict__ q,
                                                                const float* __restrict__ tau, int hl, float lo,
                                                                int hu, float hi, const float* __restrict__ mask,
                                                                const BatchSlice* __restrict__ slices,
                                                                const int64_t* __restrict__ vstart, int n_slices) {
  const int64_t total = vstart[n_slices];
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kEw + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * kEw) {
    float upd = atp[k];
    if (q) upd = upd + neg_div_at(q, locate(slices, vstart, n_slices, k), k);
    const float xo = x[k];
    float xn = clip_value(xo - upd * tau[k], hl, lo, hu, hi);
    if (mask) xn = xn * mask[k];
    xbar[k] = xn + (xn - xo);
    x[k] = xn;
  }
}

__global__ __launch_bounds__(kEw) void sirt_update_batch_kernel(float* __restrict__ x, const float* __restrict__ upd,
                                                                const float* __restrict__ q,
                                                                const float* __restrict__ tau, int hl, float lo,
                                                                int hu, float hi, const float* __restrict__ mask,
                                                                const BatchSlice* __restrict__ slices,
                                                                const int64_t* __restrict__ vstart, int n_slices) {
  const int64_t total = vstart[n_slices];
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * kEw + threadIdx.x; k < total;
       k += static_cast<int64_t>(gridDim.x) * kEw) {
    float u = upd[k];
    if (q) u = u - neg_div_at(q, locate(slices, vstart, n_slices, k), k);
    float xn = clip_value(x[k] + u * tau[k], hl, lo, hu, hi);
    if (mask) xn = xn * mask[k];
    x[k] = xn;
  }
}

dim3 ew_grid(int64_t total) {
  const int64_t blocks = (total + kEw - 1) / kEw;
  return dim3(static_cast<unsigned>(std::min<int64_t>(std::max<int64_t>(blocks, 1), 65536)));
}

template <typename T>
int upload(T*& dst, const std::vector<T>& src) {
  CHIP_HIP(hipMalloc(&dst, sizeof(T) * src.size()));
  CHIP_HIP(hipMemcpy(dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return CHIP_OK;
}

void release(BatchPlan& b) {
  if (b.d_slices) (void)hipFree(b.d_slices);
  if (b.d_item_start_fp) (void)hipFree(b.d_item_start_fp);
  if (b.d_item_start_bpa) (void)hipFree(b.d_item_start_bpa);
  if (b.d_item_start_bps) (void)hipFree(b.d_item_start_bps);
  if (b.d_vol_start) (void)hipFree(b.d_vol_start);
  if (b.d_geom) (void)hipFree(b.d_geom);
  if (b.d_scan) (void)hipFree(b.d_scan);
  if (b.d_mblock) (void)hipFree(b.d_mblock);
  if (b.d_tile_slots) (void)hipFree(b.d_tile_slots);
  if (b.d_fp_part) (void)hipFree(b.d_fp_part);
  if (b.d_bp_part) (void)hipFree(b.d_bp_part);
  if (b.bp_state && b.bp_state->done) (void)hipEventDestroy(b.bp_state->done);
  if (b.fp_state && b.fp_state->done) (void)hipEventDestroy(b.fp_state->done);
  for (chip_plan* p : b.size_plans) chip_plan_destroy(p);
  b = BatchPlan();
}

// d_geom[size][A] = the size plans' radon tables (allocated on first call)
int upload_geom(BatchPlan& b) {
  std::vector<float4> geom;
  for (chip_plan* p : b.size_plans) {
    for (int a = 0; a < b.n_angles; ++a) {
      const float* g = &p->p.h_fp_geom[4 * static_cast<size_t>(a)];
      geom.push_back(make_float4(g[0], g[1], g[2], g[3]));
    }
  }
  if (!b.d_geom) return upload(b.d_geom, geom);
  CHIP_HIP(hipMemcpy(b.d_geom, geom.data(), sizeof(float4) * geom.size(), hipMemcpyHostToDevice));
  return CHIP_OK;
}

int create(BatchPlan& b, int n_slices, const int* sizes, int n_angles, const float* angles_rad, int n_det,
           const double* det_angles_rad, const double* w_det, const double* w_ang) {
  b.n_slices = n_slices;
  b.n_angles = n_angles;
  b.n_det = n_det;
  CHIP_HIP(hipGetDevice(&b.device));
  int n_cu = 0;
  CHIP_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, b.device));
  b.n_cu = std::max(n_cu, 1);
  std::map<int, int> size_index;
  for (int s = 0; s < n_slices; ++s) {
    if (size_index.count(sizes[s])) continue;
    chip_plan* p = nullptr;
    const int rc = chip_plan_create(&p, sizes[s], n_angles, angles_rad, n_det, det_angles_rad, w_det, w_ang);
    if (rc != CHIP_OK) return rc;
    size_index[sizes[s]] = static_cast<int>(b.size_plans.size());
    b.size_plans.push_back(p);
  }
  b.h_item_start_fp.assign(1, 0);
  b.h_item_start_bpa.assign(1, 0);
  b.h_item_start_bps.assign(1, 0);
  b.h_vol_start.assign(1, 0);
  b.h_sino_start.assign(1, 0);
  b.h_maps_start.assign(1, 0);
  for (int s = 0; s < n_slices; ++s) {
    const int n = sizes[s];
    BatchSlice sl;
    sl.n = n;
    sl.n_ublk = (n + kFpBatchBins - 1) / kFpBatchBins;
    sl.size_idx = size_index[n];
    sl.pad = b.size_plans[sl.size_idx]->p.pad;
    sl.vol_off = b.h_vol_start.back();
    sl.sino_off = b.h_sino_start.back();
    sl.maps_off = b.h_maps_start.back();
    b.h_slices.push_back(sl);
    const int64_t nn = static_cast<int64_t>(n) * n;
    b.h_item_start_fp.push_back(b.h_item_start_fp.back() + static_cast<int64_t>(n_angles) * sl.n_ublk);
    b.h_item_start_bpa.push_back(b.h_item_start_bpa.back() + static_cast<int64_t>((n + 31) / 32) * ((n + 7) / 8));
    b.h_item_start_bps.push_back(b.h_item_start_bps.back() + static_cast<int64_t>((n + 15) / 16) * ((n + 15) / 16));
    b.h_vol_start.push_back(sl.vol_off + nn);
    b.h_sino_start.push_back(sl.sino_off + static_cast<int64_t>(n_det) * n_angles * n);
    b.h_maps_start.push_back(sl.maps_off + static_cast<int64_t>(n_angles) * n_det * nn);
  }
  // image-tile fp slots: the slices of each size in plan order
  std::vector<int64_t> offs;
  b.tile_group_start.assign(1, 0);
  for (size_t g = 0; g < b.size_plans.size(); ++g) {
    for (int s = 0; s < n_slices; ++s) {
      const BatchSlice& sl = b.h_slices[s];
      if (sl.size_idx != static_cast<int>(g)) continue;
      b.h_tile_slot_slice.push_back(s);
      offs.push_back(sl.vol_off);
      offs.push_back(sl.maps_off);
      offs.push_back(sl.sino_off);
    }
    b.tile_group_start.push_back(static_cast<int>(b.h_tile_slot_slice.size()));
  }
  int rc;
  if ((rc = upload(b.d_tile_slots, offs)) != CHIP_OK) return rc;
  if ((rc = upload(b.d_slices, b.h_slices)) != CHIP_OK) return rc;
  if ((rc = upload(b.d_item_start_fp, b.h_item_start_fp)) != CHIP_OK) return rc;
  if ((rc = upload(b.d_item_start_bpa, b.h_item_start_bpa)) != CHIP_OK) return rc;
  if ((rc = upload(b.d_item_start_bps, b.h_item_start_bps)) != CHIP_OK) return rc;
  if ((rc = upload(b.d_vol_start, b.h_vol_start)) != CHIP_OK) return rc;
  if ((rc = upload_geom(b)) != CHIP_OK) return rc;
  b.d_w = b.size_plans[0]->p.d_w;
  return CHIP_OK;
}

bool aligned16(const void* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; }

}  // namespace
}  // namespace chip

using namespace chip;

extern "C" {

int chip_batch_plan_create(chip_batch_plan** out, int n_slices, const int* sizes, int n_angles,
                           const float* angles_rad, int n_det, const double* det_angles_rad, const double* w_det,
                           const double* w_ang) {
  clear_error();
  CHIP_CHECK_ARG(out != nullptr && sizes != nullptr && angles_rad != nullptr && det_angles_rad != nullptr,
                 "NULL argument");
  CHIP_CHECK_ARG(n_slices >= 1, "n_slices must be >= 1");
  CHIP_CHECK_ARG(n_angles >= 1 && n_det >= 1, "n_angles and n_det must be >= 1");
  for (int s = 0; s < n_slices; ++s) {
    CHIP_CHECK_ARG(sizes[s] >= 4 && sizes[s] % 4 == 0,
                   "batched slices need a side N >= 4 with N % 4 == 0 (slice " + std::to_string(s) + " has " +
                       std::to_string(sizes[s]) + ")");
  }
  *out = nullptr;
  auto* bp = new (std::nothrow) chip_batch_plan();
  if (!bp) {
    set_error("chip_batch_plan_create: host allocation failed");
    return CHIP_ENOMEM;
  }
  const int rc = create(bp->b, n_slices, sizes, n_angles, angles_rad, n_det, det_angles_rad, w_det, w_ang);
  if (rc != CHIP_OK) {
    const std::string msg = chip_last_error();
    release(bp->b);
    delete bp;
    set_error(msg);
    return rc;
  }
  *out = bp;
  return CHIP_OK;
}

int chip_batch_set_radon_trig(chip_batch_plan* plan, const float* cos_sin) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && cos_sin != nullptr, "NULL argument");
  for (chip_plan* p : plan->b.size_plans) {
    const int rc = chip_plan_set_radon_trig(p, cos_sin);
    if (rc != CHIP_OK) return rc;
  }
  return upload_geom(plan->b);
}

int chip_batch_plan_destroy(chip_batch_plan* plan) {
  clear_error();
  if (plan) {
    release(plan->b);
    delete plan;
  }
  return CHIP_OK;
}

int chip_batch_info(const chip_batch_plan* plan, int64_t* vol_elems, int64_t* sino_elems, int64_t* maps_elems) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr, "plan is NULL");
  const BatchPlan& b = plan->b;
  if (vol_elems) *vol_elems = b.h_vol_start.back();
  if (sino_elems) *sino_elems = b.h_sino_start.back();
  if (maps_elems) *maps_elems = b.h_maps_start.back();
  return CHIP_OK;
}

int chip_batch_tune(chip_batch_plan* plan, const char* key, int value) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && key != nullptr, "NULL argument");
  const std::string k(key);
  if (k == "bp_pair_parts") {
    CHIP_CHECK_ARG(value >= 0 && value <= kBpMaxParts, "bp_pair_parts must be in 0..8 (0: from occupancy)");
    plan->b.bp_pair_parts = value;
    return CHIP_OK;
  }
  if (k == "bp_quads") {
    CHIP_CHECK_ARG(value == 0 || value == 1, "bp_quads must be 0 or 1");
    plan->b.bp_quads = value;
    return CHIP_OK;
  }
  if (k == "bp_pairs") {
    CHIP_CHECK_ARG(value == 0 || value == 1, "bp_pairs must be 0 or 1");
    plan->b.bp_pairs = value;
    return CHIP_OK;
  }
  if (k == "fp_tile") {
    CHIP_CHECK_ARG(value == 0 || value == 1, "fp_tile must be 0 or 1");
    plan->b.fp_tile = value;
    return CHIP_OK;
  }
  if (k == "fp_tile_kind" || k == "fp_tile_group" || k == "fp_tile_pairs")
    return tune_tile_shape(plan->b.fp_tile_shape, k, value);
  if (k == "fp_wg_per_cu") {
    CHIP_CHECK_ARG(value >= 0 && value <= 8, "fp_wg_per_cu must be in 0..8 (0: one workgroup per work item)");
    plan->b.fp_wg_per_cu = value;
    return CHIP_OK;
  }
  if (k == "fp_patch") {
    CHIP_CHECK_ARG(value == 0 || value == 1, "fp_patch must be 0 or 1");
    plan->b.fp_patch = value;
    return CHIP_OK;
  }
  if (k == "fp_direct") {
    CHIP_CHECK_ARG(value == 0 || value == 1, "fp_direct must be 0 or 1");
    plan->b.fp_direct = value;
    return CHIP_OK;
  }
  if (k == "fp_trim") {
    CHIP_CHECK_ARG(value == 0 || value == 1, "fp_trim must be 0 or 1");
    plan->b.fp_trim = value;
    return CHIP_OK;
  }
  if (k == "otf_budget_mib") {
    CHIP_CHECK_ARG(value >= 16, "otf_budget_mib must be >= 16");
    plan->b.otf_budget = static_cast<size_t>(value) << 20;
    return CHIP_OK;
  }
  if (k == "bp_wg_per_cu") {
    CHIP_CHECK_ARG(value >= 0 && value <= 16, "bp_wg_per_cu must be in 0..16 (0: one workgroup per tile)");
    plan->b.bp_wg_per_cu = value;
    return CHIP_OK;
  }
  set_error("chip_batch_tune: unknown key '" + k + "'");
  return CHIP_EINVAL;
}

int chip_batch_att_maps(chip_batch_plan* plan, const float* mu_in, const float* mu_out, float* maps, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && maps != nullptr, "NULL argument");
  CHIP_CHECK_ARG(mu_in != nullptr || mu_out != nullptr, "no attenuation volumes were given");
  BatchPlan& b = plan->b;
  hipStream_t st = as_stream(stream);
  std::vector<int> chunk(b.size_plans.size());
  size_t need = 0;
  for (size_t z = 0; z < b.size_plans.size(); ++z) {
    size_t scan = 0;
    chunk[z] = maps_block_angles(b.size_plans[z]->p, mu_in != nullptr, mu_out != nullptr, false, kMapsScanBudget,
                                 &scan, nullptr);
    need = std::max(need, scan);
  }
  int rc = ensure_buffer(b.d_scan, b.scan_bytes, need, st, "chip_batch_att_maps");
  if (rc != CHIP_OK) return rc;
  for (const BatchSlice& sl : b.h_slices) {
    rc = att_maps_slices(b.size_plans[sl.size_idx]->p, mu_in ? mu_in + sl.vol_off : nullptr,
                         mu_out ? mu_out + sl.vol_off : nullptr, maps + sl.maps_off, 1, chunk[sl.size_idx], b.d_scan,
                         st);
    if (rc != CHIP_OK) return rc;
  }
  return CHIP_OK;
}

int chip_batch_fp(const chip_batch_plan* plan, const float* vol, const float* maps, float* sino, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && vol != nullptr && sino != nullptr, "NULL argument");
  CHIP_CHECK_ARG(aligned16(vol) && (maps == nullptr || aligned16(maps)), "vol / maps must be 16-byte aligned");
  const BatchPlan& b = plan->b;
  const hipStream_t st = as_stream(stream);
  bool tile = b.fp_tile != 0;
  for (chip_plan* sp : b.size_plans) tile = tile && fp_tile_count(sp->p.n) <= kFpTileMaxTiles;
  if (tile) {
    // one image-tile launch per slice size, through one partial-sum scratch
    ScratchUse use(*b.fp_state, st);
    size_t need = 0;
    for (size_t g = 0; g < b.size_plans.size(); ++g) {
      const Plan& sp = b.size_plans[g]->p;
      const int rc = fp_tile_table(sp, st);
      if (rc != CHIP_OK) return rc;
      need = std::max(need, fp_tile_part_bytes(sp, b.tile_group_start[g + 1] - b.tile_group_start[g], b.n_angles));
    }
    if (b.fp_part_bytes < need) {
      CHIP_CHECK_ARG(!stream_capturing(st), "the image-tile fp needs " + std::to_string(need) +
                                                " bytes of partial sums: run it once outside stream capture first");
      auto& bm = const_cast<BatchPlan&>(b);
      const int rc = ensure_buffer(bm.d_fp_part, bm.fp_part_bytes, need, st, "image-tile fp partial sums");
      if (rc != CHIP_OK) return rc;
    }
    for (size_t g = 0; g < b.size_plans.size(); ++g) {
      const int g0 = b.tile_group_start[g], ng = b.tile_group_start[g + 1] - g0;
      if (ng == 0) continue;
      TileSlots sl;
      sl.offs = b.d_tile_slots + 3 * static_cast<size_t>(g0);
      launch_fp_tile(b.size_plans[g]->p, vol, maps, sl, sino, ng, 0, b.n_angles, 0, b.d_fp_part, b.fp_tile_shape, st);
    }
    CHIP_LAUNCH_CHECK();
    return CHIP_OK;
  }
  launch_fp_stream_batch(b, vol, maps, sino, st);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_batch_bp(const chip_batch_plan* plan, const float* sino, const float* maps, float* vol, int mode,
                  void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && vol != nullptr && sino != nullptr, "NULL argument");
  CHIP_CHECK_ARG(mode >= CHIP_BP_SKIMAGE && mode <= CHIP_BP_ADJOINT, "unknown bp mode");
  hipStream_t st = as_stream(stream);
  if (mode == CHIP_BP_ADJOINT) {
    const BatchPlan& b = plan->b;
    if (b.bp_pairs) {
      ScratchUse use(*b.bp_state, st);
      int parts = 1;
      const int rc = bp_pair_parts_ready(b, maps != nullptr, st, &parts);
      if (rc != CHIP_OK) return rc;
      launch_bp_adjoint_batch(b, sino, maps, vol, st, parts);
    } else {
      launch_bp_adjoint_batch(b, sino, maps, vol, st, 1);
    }
  }
  else launch_bp_skimage_batch(plan->b, sino, maps, vol, mode == CHIP_BP_SKIMAGE_SYM ? 1 : 0, st);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

}  // extern "C"

namespace chip {
namespace {

// Block sizes of every slice size within the batch's budget and both shared
// scratch buffers grown to the largest need.
int batch_otf_prepare(BatchPlan& b, bool in, bool out, hipStream_t st, std::vector<int>& chunk) {
  chunk.resize(b.size_plans.size());
  size_t need_c = 0, need_m = 0;
  for (size_t z = 0; z < b.size_plans.size(); ++z) {
    size_t c = 0, m = 0;
    chunk[z] = maps_block_angles(b.size_plans[z]->p, in, out, true, b.otf_budget, &c, &m);
    need_c = std::max(need_c, c);
    need_m = std::max(need_m, m);
  }
  const int rc = ensure_buffer(b.d_scan, b.scan_bytes, need_c, st, "otf prefix sums");
  if (rc != CHIP_OK) return rc;
  return ensure_buffer(b.d_mblock, b.mblock_bytes, need_m, st, "otf maps block");
}

}  // namespace
}  // namespace chip

extern "C" {

int chip_batch_fp_otf(chip_batch_plan* plan, const float* vol, const float* mu_in, const float* mu_out, float* sino,
                      void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && vol != nullptr && sino != nullptr, "NULL argument");
  CHIP_CHECK_ARG(mu_in != nullptr || mu_out != nullptr, "no attenuation volumes were given");
  BatchPlan& b = plan->b;
  hipStream_t st = as_stream(stream);
  std::vector<int> chunk;
  int rc = batch_otf_prepare(b, mu_in != nullptr, mu_out != nullptr, st, chunk);
  if (rc != CHIP_OK) return rc;
  for (const BatchSlice& sl : b.h_slices) {
    rc = otf_fp(b.size_plans[sl.size_idx]->p, vol + sl.vol_off, mu_in ? mu_in + sl.vol_off : nullptr,
                mu_out ? mu_out + sl.vol_off : nullptr, sino + sl.sino_off, 1, chunk[sl.size_idx], b.d_scan,
                b.d_mblock, st);
    if (rc != CHIP_OK) return rc;
  }
  return CHIP_OK;
}

int chip_batch_bp_otf(chip_batch_plan* plan, const float* sino, const float* mu_in, const float* mu_out, float* vol,
                      int mode, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && vol != nullptr && sino != nullptr, "NULL argument");
  CHIP_CHECK_ARG(mu_in != nullptr || mu_out != nullptr, "no attenuation volumes were given");
  CHIP_CHECK_ARG(mode == CHIP_BP_SKIMAGE_SYM || mode == CHIP_BP_ADJOINT,
                 "mode must be CHIP_BP_SKIMAGE_SYM or CHIP_BP_ADJOINT (CHIP_BP_SKIMAGE uses no maps: call chip_batch_bp)");
  BatchPlan& b = plan->b;
  hipStream_t st = as_stream(stream);
  std::vector<int> chunk;
  int rc = batch_otf_prepare(b, mu_in != nullptr, mu_out != nullptr, st, chunk);
  if (rc != CHIP_OK) return rc;
  for (const BatchSlice& sl : b.h_slices) {
    rc = otf_bp(b.size_plans[sl.size_idx]->p, sino + sl.sino_off, mu_in ? mu_in + sl.vol_off : nullptr,
                mu_out ? mu_out + sl.vol_off : nullptr, vol + sl.vol_off, 1, mode, chunk[sl.size_idx], b.d_scan,
                b.d_mblock, st);
    if (rc != CHIP_OK) return rc;
  }
  return CHIP_OK;
}

int chip_batch_otf_fp_row_bp(chip_batch_plan* plan, const float* x, const float* mu_in, const float* mu_out,
                             int row_op, float* p, const float* sigma, const float* b, float* ax, float* vol,
                             int bp_mode, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && x != nullptr && p != nullptr && b != nullptr && ax != nullptr && vol != nullptr,
                 "NULL argument");
  CHIP_CHECK_ARG(mu_in != nullptr || mu_out != nullptr, "no attenuation volumes were given");
  CHIP_CHECK_ARG(row_op >= CHIP_ROW_PDHG_L2 && row_op <= CHIP_ROW_MLEM, "unknown row_op");
  CHIP_CHECK_ARG(sigma != nullptr || row_op == CHIP_ROW_SIRT || row_op == CHIP_ROW_MLEM,
                 "the PDHG row ops need sigma");
  CHIP_CHECK_ARG(bp_mode == CHIP_BP_SKIMAGE_SYM || bp_mode == CHIP_BP_ADJOINT,
                 "bp_mode must be CHIP_BP_SKIMAGE_SYM or CHIP_BP_ADJOINT");
  CHIP_CHECK_ARG(p != ax, "p and ax must be different buffers");
  BatchPlan& bp = plan->b;
  hipStream_t st = as_stream(stream);
  std::vector<int> chunk;
  int rc = batch_otf_prepare(bp, mu_in != nullptr, mu_out != nullptr, st, chunk);
  if (rc != CHIP_OK) return rc;
  for (const BatchSlice& sl : bp.h_slices) {
    const int64_t so = sl.sino_off;
    rc = otf_fp_row_bp(bp.size_plans[sl.size_idx]->p, x + sl.vol_off, mu_in ? mu_in + sl.vol_off : nullptr,
                       mu_out ? mu_out + sl.vol_off : nullptr, 1, row_op, p + so, sigma ? sigma + so : nullptr,
                       b + so, nullptr, nullptr, ax + so, vol + sl.vol_off, bp_mode, chunk[sl.size_idx], bp.d_scan,
                       bp.d_mblock, st);
    if (rc != CHIP_OK) return rc;
  }
  return CHIP_OK;
}

int chip_batch_tv_dual(const chip_batch_plan* plan, float* q, const float* x, float lam, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && q != nullptr && x != nullptr, "NULL argument");
  const BatchPlan& b = plan->b;
  hipLaunchKernelGGL(tv_dual_batch_kernel, ew_grid(b.h_vol_start.back()), dim3(kEw), 0, as_stream(stream), q, x,
                     b.d_slices, b.d_vol_start, b.n_slices, lam);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_batch_pdhg_primal(const chip_batch_plan* plan, float* x, float* xbar, const float* atp, const float* q,
                           const float* tau, int has_lower, float lower, int has_upper, float upper,
                           const float* x_mask, void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && x && xbar && atp && tau, "NULL argument");
  const BatchPlan& b = plan->b;
  hipLaunchKernelGGL(pdhg_primal_batch_kernel, ew_grid(b.h_vol_start.back()), dim3(kEw), 0, as_stream(stream), x,
                     xbar, atp, q, tau, has_lower, lower, has_upper, upper, x_mask, b.d_slices, b.d_vol_start,
                     b.n_slices);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

int chip_batch_sirt_update(const chip_batch_plan* plan, float* x, const float* upd, const float* q, const float* tau,
                           int has_lower, float lower, int has_upper, float upper, const float* x_mask,
                           void* stream) {
  clear_error();
  CHIP_CHECK_ARG(plan != nullptr && x && upd && tau, "NULL argument");
  const BatchPlan& b = plan->b;
  hipLaunchKernelGGL(sirt_update_batch_kernel, ew_grid(b.h_vol_start.back()), dim3(kEw), 0, as_stream(stream), x,
                     upd, q, tau, has_lower, lower, has_upper, upper, x_mask, b.d_slices, b.d_vol_start, b.n_slices);
  CHIP_LAUNCH_CHECK();
  return CHIP_OK;
}

}  // extern "C"
