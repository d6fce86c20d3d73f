// Shared host/device definitions of libcorrct_hip (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/corrct_hip.h"

namespace chip {

// ---------------------------------------------------------------------------
// error handling: per-thread last-error string, status codes
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
void clear_error();

#define CHIP_CHECK_ARG(cond, msg)                 \
  do {                                            \
    if (!(cond)) {                                \
      ::chip::set_error(std::string(__func__) + ": " + (msg)); \
      return CHIP_EINVAL;                         \
    }                                             \
  } while (0)

#define CHIP_HIP(call)                                                                    \
  do {                                                                                    \
    hipError_t err_ = (call);                                                             \
    if (err_ != hipSuccess) {                                                             \
      ::chip::set_error(std::string(__func__) + ": " #call " -> " + hipGetErrorString(err_)); \
      return CHIP_EHIP;                                                                   \
    }                                                                                     \
  } while (0)

#define CHIP_LAUNCH_CHECK() CHIP_HIP(hipGetLastError())

// ---------------------------------------------------------------------------
// per-rotation affine inverse map of an order-1 warp (float32, skimage order)
//   X = (m00 * x + m01 * y) + m02,  Y = (m10 * x + m11 * y) + m12
// ---------------------------------------------------------------------------
struct Affine {
  float m00, m01, m02, m10, m11, m12;
};

// Host-side geometry plan. Device tables are allocated at plan creation.
constexpr int kBpMaxParts = 8;
constexpr int kBpPartsMaxN = 1024;

// Where the part sums of a split single-slice adjoint bp go: `part` holds
// cap_elems floats (parts x N^2 must fit); tune = 0 never splits into parts,
// 2..kBpMaxParts pins the count (the same bits on every device), -1 picks it
// from the kernel's occupancy on this device.
struct BpScratch {
  float* part = nullptr;
  int64_t cap_elems = 0;
  int tune = -1;
};

// Lazily allocated part scratch of a plan and the ordering of its users: the
// last stream that used it and an event recorded after that use, so a call
// on another stream waits for it instead of overwriting partial sums in flight.
struct BpPartState {
  std::mutex mu;
  hipEvent_t done = nullptr;
  hipStream_t last = nullptr;
  bool used = false;
};

// is `st` being captured into a graph (or cannot tell)?
inline bool stream_capturing(hipStream_t st) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone;
}

// Where the slices of an image-tile fp launch live: slot b at the element
// offsets offs[3 b], offs[3 b + 1], offs[3 b + 2] of its volume, maps and
// sinogram, or (offs NULL) at b times the strides.
struct TileSlots {
  const int64_t* offs = nullptr;
  int64_t vol_stride = 0, maps_stride = 0, sino_stride = 0;
  __device__ __forceinline__ int64_t vol(int b) const { return offs ? offs[3 * b] : b * vol_stride; }
  __device__ __forceinline__ int64_t maps(int b) const { return offs ? offs[3 * b + 1] : b * maps_stride; }
  __device__ __forceinline__ int64_t sino(int b) const { return offs ? offs[3 * b + 2] : b * sino_stride; }
};

// image-tile fp launch shape: kernel kind (0: map tiles fetched one angle ahead
// into registers, two samples per lane and trip; 1: ahead, one; 2: at staging,
// two; 3: at staging, one; -1: 1 for two-detector passes, 3 for one-detector
// passes), angles per workgroup, two slices per workgroup
struct FpTileShape {
  int kind = -1;
  int group = 8;
  int pairs = 1;
};

struct Plan {
  int n = 0;          // slice side N (square), also detector pixels
  int n_angles = 0;   // A
  int n_det = 0;      // D
  int pad = 0;        // attenuation pad width e
  int np = 0;         // padded side N + 2e
  // fp geometry per angle: (cos, sin, m02, m12) of the radon inverse map
  std::vector<float> h_fp_geom;  // [A][4]
  std::vector<float> h_w;        // [A][D] = w_det[d] * w_ang[a][d]
  std::vector<Affine> h_rot_in_fwd, h_rot_in_back;    // [A]
  std::vector<Affine> h_rot_out_fwd, h_rot_out_back;  // [A][D][n_sub]
  // outgoing-beam sub-angles per detector (detector surface sampling): the
  // out-beam maps average exp(-.) over n_sub_det[d] <= n_sub angles
  int n_sub = 1;
  int* d_nsub = nullptr;  // [D]
  float4* d_fp_geom = nullptr;
  float2* d_ytab = nullptr;  // [A][N] (fl(sin * y), fl(cos * y)) row terms of the fp map
  float* d_w = nullptr;
  Affine* d_rot = nullptr;  // [A] in-fwd, [A] in-back, [A*D*n_sub] out-fwd, [A*D*n_sub] out-back
  // scratch for the attenuation builder (rotated-frame cumulative sums)
  float* d_scratch = nullptr;
  size_t scratch_bytes = 0;
  // maps of one angle block for the on-the-fly projectors
  float* d_mblock = nullptr;
  size_t mblock_bytes = 0;
  // per-part sums of a single-slice adjoint bp whose angle range is split over
  // more workgroups than the tiles alone give ([kBpMaxParts][N][N], N <= 1024),
  // allocated by the first call that can use it (never under stream capture)
  mutable float* d_bp_part = nullptr;
  std::unique_ptr<BpPartState> bp_state{new BpPartState};
  // chip_plan_tune "bp_parts": -1 = from occupancy, 0 = never, 2..8 = fixed
  int bp_parts_tune = -1;
  // map builder blends: 0 = float32 (default), 1 = scikit-image's exact float64 blends
  // (chip_plan_tune "maps_exact")
  int maps_exact = 0;
  // bound on d_scratch + d_mblock (chip_plan_tune "otf_budget_mib"); larger angle blocks
  // keep more of the map builder's waves in flight (22.4 against 24.3 ms for 4 against
  // 1 GiB at 1024^2, 1440 angles)
  size_t otf_budget = size_t(4) << 30;
  int device = 0;
  // kernel selection (chip_plan_tune): fp 0 = gather, 1 = LDS-staged,
  // 2 = LDS-staged with register prefetch, 3 = streamed per-detector tiles,
  // 4 = 3 staging only the footprint of interior chunks, 5 = 3 with 32-bin tiles,
  // 6 = image-space bands (2..6 fall back to 1 when N % 4 != 0), 10 = image tiles
  // looping over angle groups (fp_tile.hip; 5 where a slice has more than
  // kFpTileMaxTiles tiles or the buffers are not 16-byte aligned)
  int fp_variant = 10;
  // image-tile fp shape (chip_plan_tune "fp_tile_kind", "fp_tile_group", "fp_tile_pairs")
  FpTileShape fp_tile;
  // the image-tile fp's bin table ([A][tiles] records, geometry only) and its
  // partial sums ([batch][angles][tiles][bins][<= 2]), allocated by the first
  // whole-sinogram fp that needs them (never under stream capture)
  mutable int* d_fp_tab = nullptr;
  mutable size_t fp_tab_bytes = 0;
  mutable bool fp_tab_ok = false;  // built for the current radon table
  mutable float* d_fp_tile = nullptr;
  mutable size_t fp_tile_bytes = 0;
  std::unique_ptr<BpPartState> fp_state{new BpPartState};  // users of d_fp_tile across streams
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------------------
// batches of slices of different sizes (one persistent launch per pass)
// ---------------------------------------------------------------------------
// detector bins per workgroup of the batched fp (the kernel's tile width)
constexpr int kFpBatchBins = 32;
// image-tile fp: cells per tile row / column, bin slots per (tile, angle) and
// ints per (angle, tile) record of its bin table
constexpr int kFpTileCols = 64;
constexpr int kFpTileRows = 32;
constexpr int kFpTileBins = 80;
constexpr int kFpTileTab = kFpTileBins + 4;
// tiles per slice the image-tile fp handles (its per-angle tile sort runs in LDS)
constexpr int kFpTileMaxTiles = 8192;

struct BatchSlice {
  int n;         // slice side
  int n_ublk;    // 64-bin blocks of the fp
  int size_idx;  // index of the slice's size: geometry at geom + size_idx * A
  int pad;
  int64_t vol_off;   // element offset of [n][n] in volume-shaped buffers
  int64_t sino_off;  // element offset of [D][A][n] in sinogram-shaped buffers
  int64_t maps_off;  // element offset of [A][D][n][n] in the maps buffer
};

struct BatchPlan {
  int n_slices = 0, n_angles = 0, n_det = 0;
  int n_cu = 256;
  int fp_wg_per_cu = 5;  // persistent grid of the fp: fp_wg_per_cu x CUs workgroups
  int bp_wg_per_cu = 4;  // same for the bp
  int fp_trim = 0;       // stage only the footprint of interior fp chunks
  int fp_direct = 0;     // stage the fp chunks with direct-to-LDS loads (fp_kernel 7 / 8 of a plan)
  int fp_patch = 0;      // gather 8 x 8 sample patches per wave instruction (fp_kernel 9 of a plan)
  std::vector<chip_plan*> size_plans;  // one single-slice plan per distinct size
  std::vector<BatchSlice> h_slices;
  // prefix sums (n_slices + 1 entries): fp items, adjoint-bp tiles, skimage-bp
  // tiles, pixels, sinogram elements, map elements
  std::vector<int64_t> h_item_start_fp, h_item_start_bpa, h_item_start_bps, h_vol_start, h_sino_start,
      h_maps_start;
  BatchSlice* d_slices = nullptr;
  int64_t* d_item_start_fp = nullptr;
  int64_t* d_item_start_bpa = nullptr;
  int64_t* d_item_start_bps = nullptr;
  int64_t* d_vol_start = nullptr;
  float4* d_geom = nullptr;   // [size][A]
  const float* d_w = nullptr; // [A][D], owned by size_plans[0]
  // map-builder scratch shared by all sizes: prefix sums and one maps block
  // (chip_batch_att_maps, chip_batch_fp_otf / _bp_otf), grown on first use
  float* d_scan = nullptr;
  size_t scan_bytes = 0;
  float* d_mblock = nullptr;
  size_t mblock_bytes = 0;
  size_t otf_budget = size_t(4) << 30;  // chip_batch_tune "otf_budget_mib"
  int device = 0;
  // image-tile fp (chip_batch_tune "fp_tile" and its shape keys): one launch per
  // slice size over that size's slots (tile_slots: (vol, maps, sino) offsets of
  // each size's slices in plan order, size group g at tile_group_start[g])
  int fp_tile = 1;
  FpTileShape fp_tile_shape;
  // exact-adjoint bp: two slices of a size per workgroup (chip_batch_tune "bp_pairs"),
  // the angle range split into parts ("bp_pair_parts": 0 = from occupancy, 1..8 pinned)
  // summed in order from part scratch (parts x the packed volume, grown on first use)
  int bp_pairs = 1;
  int bp_quads = 1;  // with bp_pairs: four slices of a size per workgroup first ("bp_quads")
  int bp_pair_parts = 0;
  float* d_bp_part = nullptr;
  size_t bp_part_bytes = 0;
  std::unique_ptr<BpPartState> bp_state{new BpPartState};  // users of d_bp_part across streams
  std::vector<int> tile_group_start;  // [sizes + 1]
  int64_t* d_tile_slots = nullptr;
  std::vector<int> h_tile_slot_slice;  // slice of each slot
  float* d_fp_part = nullptr;
  size_t fp_part_bytes = 0;
  std::unique_ptr<BpPartState> fp_state{new BpPartState};
};

// largest s in [0, n_slices) with start[s] <= k (start is non-decreasing)
__device__ __forceinline__ int find_slice(const int64_t* __restrict__ start, int n_slices, int64_t k) {
  int lo = 0, hi = n_slices - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

void launch_fp_stream_batch(const BatchPlan& bp, const float* vol, const float* maps, float* sino, hipStream_t st);
// the batched exact-adjoint bp; with bp_pairs, bp_pair_parts_ready first (it picks the part
// count and grows the part scratch; an error under stream capture when the scratch is missing)
int bp_pair_parts_ready(const BatchPlan& bp, bool maps, hipStream_t st, int* parts);
void launch_bp_adjoint_batch(const BatchPlan& bp, const float* sino, const float* maps, float* vol, hipStream_t st,
                             int parts);
void launch_bp_skimage_batch(const BatchPlan& bp, const float* sino, const float* maps, float* vol, int sym,
                             hipStream_t st);

// fp over angles [angle0, angle0 + n_ang) and detectors [det0, det0 + nd),
// nd in {1, 2}; maps[0] holds angle maps_a0. Needs N % 4 == 0 and 16-byte
// aligned vol / maps (fp_stream.hip)
void launch_fp_stream(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch,
                      int det0, int nd, int angle0, int n_ang, int maps_a0, hipStream_t st);

// the same fp in image-space bands of rows or columns, staging each map value once
// (fp_band.hip); same preconditions as launch_fp_stream
void launch_fp_band(const Plan& p, const float* vol, const float* maps, int64_t stride, float* sino, int batch,
                    int det0, int nd, int angle0, int n_ang, int maps_a0, hipStream_t st);

// the same fp by image tiles, each workgroup looping over a group of angles
// (fp_tile.hip): partial sums per (slot, angle, tile) in `part` (at least
// fp_tile_part_bytes), then a fixed-order sum over the tiles of each line;
// every detector, two per pass; same preconditions as launch_fp_stream.
// fp_tile_table builds the plan's bin table once per radon table (an error
// under stream capture when it is missing).
int fp_tile_count(int n);
int fp_tile_table(const Plan& p, hipStream_t st);
size_t fp_tile_part_bytes(const Plan& p, int slots, int n_ang);
void launch_fp_tile(const Plan& p, const float* vol, const float* maps, const TileSlots& sl, float* sino, int slots,
                    int angle0, int n_ang, int maps_a0, float* part, const FpTileShape& shape, hipStream_t st);

// One call's use of scratch shared across streams (the state's mutex held while
// it lives): waits on `st` for the previous user on another stream, then records
// this use (no event work while `st` is being captured).
// chip_plan_tune / chip_batch_tune keys "fp_tile_kind", "fp_tile_group", "fp_tile_pairs"
int tune_tile_shape(FpTileShape& sh, const std::string& k, int value);

class ScratchUse {
 public:
  ScratchUse(BpPartState& s, hipStream_t st);
  ~ScratchUse();
  ScratchUse(const ScratchUse&) = delete;
  ScratchUse& operator=(const ScratchUse&) = delete;

 private:
  BpPartState& s_;
  hipStream_t st_;
  std::unique_lock<std::mutex> lk_;
  bool record_ = false;
};

// exact adjoint of the fp over angles [angle0, angle0 + n_ang), all detectors;
// maps[0] holds angle maps_a0; accumulate = 1 adds to vol (bp_adjoint.hip)
void launch_bp_skimage_tiles_batch(const BatchPlan& bp, const float* sino, const float* maps, float* vol, int sym,
                                   hipStream_t st);
void launch_bp_skimage_tiles(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol,
                             int batch, int sym, int angle0, int n_ang, int maps_a0, int accumulate,
                             hipStream_t st);
void launch_bp_adjoint_tiles(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol,
                             int batch, int angle0, int n_ang, int maps_a0, int accumulate, hipStream_t st,
                             const BpScratch& scr = BpScratch{});

// would a bp of `batch` slices in `mode` over n_ang angles split into parts?
bool bp_may_split(const Plan& p, int batch, int mode, int n_ang);
// the part count a split single-slice adjoint bp of all angles uses with room for
// max_parts parts (0 = no split into parts) (bp_adjoint.hip)
int bp_adjoint_parts(const Plan& p, bool maps, int max_parts);
// chip_bp's launches for whole sinograms with the part scratch `scr` (projector.hip)
void launch_bp_sinogram(const Plan& p, const float* sino, const float* maps, int64_t stride, float* vol, int batch,
                        int mode, hipStream_t st, const BpScratch& scr);

// The plan's part scratch for one call (bp_adjoint.hip). While it lives it holds
// the plan's BpPartState: on entry it allocates the scratch when the call may
// split into parts and `st` is not being captured, and makes `st` wait for the
// previous user of the scratch on another stream; on exit it records this use.
class BpPartGuard {
 public:
  BpPartGuard(const Plan& p, int batch, int mode, int n_ang, hipStream_t st);
  ~BpPartGuard();
  BpPartGuard(const BpPartGuard&) = delete;
  BpPartGuard& operator=(const BpPartGuard&) = delete;
  const BpScratch& scratch() const { return scr_; }

 private:
  const Plan& p_;
  hipStream_t st_;
  std::unique_lock<std::mutex> lk_;
  BpScratch scr_;
  bool record_ = false;
};

// fp / bp (mode CHIP_BP_SKIMAGE_SYM or CHIP_BP_ADJOINT) over an angle range
// with a maps block whose first angle is maps_a0 (projector.hip); the fp runs
// the image-tile kernel when the plan selects it and fp_range_ready has set it
// up for ranges of up to `chunk` angles (the callers hold the plan's fp_state)
bool fp_range_tile(const Plan& p);
int fp_range_ready(const Plan& p, int chunk, hipStream_t st);
void launch_fp_range(const Plan& p, const float* vol, const float* maps, float* sino, int angle0, int n_ang,
                     int maps_a0, hipStream_t st);
void launch_bp_range(const Plan& p, const float* sino, const float* maps, float* vol, int mode, int angle0,
                     int n_ang, int maps_a0, int accumulate, hipStream_t st, const BpScratch& scr = BpScratch{});

constexpr int kMaxDetPerPass = 4;

// ---------------------------------------------------------------------------
// attenuation maps through caller-owned scratch (attenuation.hip)
// ---------------------------------------------------------------------------
// prefix-sum scratch bound of the precomputed-maps builder
constexpr size_t kMapsScanBudget = size_t(1) << 30;

// Grows the device buffer `buf` of `bytes` bytes to at least `need` (draining
// `st` before it frees the old one).
int ensure_buffer(float*& buf, size_t& bytes, size_t need, hipStream_t st, const char* what);
// Angles per map block of a slice of plan p within `budget` bytes of scratch
// (prefix sums, plus the maps block itself when with_block), and the bytes of
// each part at that block size.
int maps_block_angles(const Plan& p, bool in, bool out, bool with_block, size_t budget, size_t* scan_bytes,
                      size_t* block_bytes);
// Maps [A][D][N][N] of `batch` slices, `chunk` angles at a time through `scan`.
int att_maps_slices(const Plan& p, const float* mu_in, const float* mu_out, float* maps, int batch, int chunk,
                    float* scan, hipStream_t st);
// fp / bp (mode CHIP_BP_SKIMAGE_SYM or CHIP_BP_ADJOINT) of `batch` slices with
// their maps rebuilt per block of `chunk` angles into `block` through `scan`.
int otf_fp(const Plan& p, const float* vol, const float* mu_in, const float* mu_out, float* sino, int batch,
           int chunk, float* scan, float* block, hipStream_t st);
int otf_bp(const Plan& p, const float* sino, const float* mu_in, const float* mu_out, float* vol, int batch,
           int mode, int chunk, float* scan, float* block, hipStream_t st);
// per angle block: maps, fp of x into ax, the row op (enum chip_row_op) on p's rows, bp of p's rows
int otf_fp_row_bp(const Plan& p, const float* x, const float* mu_in, const float* mu_out, int batch, int op,
                  float* pd, const float* sigma, const float* b, const float* bg, const float* mask, float* ax,
                  float* vol, int mode, int chunk, float* scan, float* block, hipStream_t st);

}  // namespace chip

struct chip_plan {
  chip::Plan p;
};
