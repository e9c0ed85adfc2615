// Host-side launchers of the HIP kernels of libmrfmap_hip (internal).
#pragma once

#include "mrf_internal.h"

namespace mrf {

// Storage of the per-incidence messages: voxel->ray ln n (input side) and
// ray->voxel p*l (output side).  The mixed types keep ln n in FP32.
enum MsgType { kMsgF32 = 0, kMsgF16 = 1, kMsgBF16 = 2, kMsgMixedF16 = 3, kMsgMixedBF16 = 4 };
inline int msg_in_type(int t) { return t == kMsgF16 ? kMsgF16 : (t == kMsgBF16 ? kMsgBF16 : kMsgF32); }
inline int msg_out_type(int t) {
  return (t == kMsgF16 || t == kMsgMixedF16) ? kMsgF16 : ((t == kMsgBF16 || t == kMsgMixedBF16) ? kMsgBF16 : kMsgF32);
}
inline size_t msg_bytes(int type) { return type == kMsgF32 ? 4 : 2; }

// Schedules of the ray-message kernel: lanes per ray (L) x steps of loads in
// flight (U).  Picked per workload by the autotuner in api.hip.
enum RaySched { kSchedL1U4 = 0, kSchedL1U8 = 1, kSchedL2U4 = 2, kSchedL2U8 = 3, kSchedL4U2 = 4, kSchedL4U4 = 5,
                kNumSched = 6 };

// traverse.hip
void launch_mark_rays(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                      const float* bia, uint8_t* bflags, hipStream_t s);
void launch_mark_endpoints(const CamGeom& c, const GridGeom& g, const float* depth, uint8_t* bflags,
                           hipStream_t s);
void launch_count(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                  const float* bia, const int32_t* slot, uint32_t* counts, uint32_t* has, hipStream_t s);
// slice (8x8 pixel tile) lengths: max records over the tile's 64 rays
void launch_slice_len(const uint32_t* counts, uint32_t ntiles, uint32_t* slice_len, uint32_t* slice_rec,
                      hipStream_t s);
// sliced-ELL emission: record k of lane l of slice t at slice_off[t] + k*64 + l,
// padded to the slice length with {kGapVox, 0}; lanes[t*64+l] describes the ray
void launch_emit(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig, const float* bia,
                 const int32_t* slot, const uint32_t* slice_off, const uint32_t* slice_len, uint2* sell,
                 RayRec* lanes, hipStream_t s);

// scan.hip: exclusive prefix sum of n uint32 values; total is written to
// *d_total (device, uint64).  tmp must hold scan_tmp_words(n) uint64 words.
size_t scan_tmp_words(size_t n);
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint64_t* tmp, uint64_t* d_total,
                           hipStream_t s);

int num_tiled_pixels(const CamGeom& c);

// geometry.hip: stable sort of a camera's incidences by voxel into
// (camera, voxel) segments; records are rewritten to sorted positions.
struct SegmentBuild {
  uint2* rec;          // incidence records (voxel -> sorted position in place)
  uint32_t nrec;
  uint32_t gap_key;    // key of gap records (> every voxel id)
  int key_bits;
  uint32_t *keys_in, *keys_out, *vals_in, *vals_out, *heads, *hscan;
  void* temp;
  size_t temp_bytes;
  uint64_t* scan_tmp;
  uint64_t* d_total;   // receives the number of pairs
};
size_t segment_sort_temp_bytes(uint32_t n, int bits);
hipError_t build_segments(SegmentBuild& b, hipStream_t s);
// pair offsets/voxels, summed chord length C per pair (needs the head
// bitmask of launch_pack_heads), then records rewritten to sorted positions
// and rec_pair[i] set to the pair of live record i
void finish_segments(SegmentBuild& b, uint32_t npairs, uint32_t nlive, uint32_t* pair_off, uint32_t* pair_vox,
                     float* pair_c, const uint64_t* head_bits, const uint32_t* tile_h0, uint32_t* rec_pair,
                     hipStream_t s);
// records with their voxel ids restored (for export)
void launch_export_vox(const uint2* rec, const uint32_t* rec_pair, uint32_t n, const uint32_t* pair_vox, uint2* out,
                       hipStream_t s);
// head bitmask (bit j = position j starts a pair) and heads before each tile
void launch_pack_heads(const SegmentBuild& b, uint32_t nlive, uint64_t* bits, uint32_t* tile_h0, hipStream_t s);

// bp.hip: one camera pass = per-pair voxel->ray messages -> ray messages ->
// segmented epilogue.
struct CamPass {
  const RayRec* lanes;     // one per lane of every slice
  const uint32_t* slice_len;
  uint32_t nslices;
  const uint2* inc;        // sliced records {sorted position | kGapVox, chord length}
  const uint32_t* inc_pair;  // pair index of each live sliced record
  const uint32_t* pair_vox;
  const uint32_t* pair_off;  // npairs + 1
  const float* pair_c;
  uint32_t npairs;
  uint32_t nlive;          // positions (live incidences)
  const uint64_t* heads;   // head bitmask
  const uint32_t* tile_h0; // heads before each kTile tile
  void* apair;             // per-pair voxel -> ray message ln n (input side type)
  float* S;                // per-pair scratch: segment sums
  uint32_t *pf_q, *pl_q;   // per-tile scratch of the segment sums
  float *pf_v, *pl_v, *mid;
  float* B;
  float* M;                // this camera's messages M_c (FP32, pool sized)
  void* Xo;                // segment array of ray->voxel messages p*l (output side type)
  uint8_t* upd;
  const float* sig;        // sensor-model lookup tables
  const float* bia;
  int lookup_n;
  int msg_type;
  int ray_sched;           // RaySched
};
void launch_pair_msgs(const CamPass& a, hipStream_t s);
// deterministic segment sums of an FP32 position array into S[pair]
void launch_segment_sum_f32(const uint64_t* heads, const uint32_t* tile_h0, uint32_t nlive, const float* X, float* S,
                            uint32_t* pf_q, float* pf_v, uint32_t* pl_q, float* pl_v, float* mid,
                            const uint32_t* pair_off, hipStream_t s);
void launch_ray_messages(const CamPass& a, hipStream_t s);
void launch_epilogue(const CamPass& a, hipStream_t s);
void launch_marginals(const float* B, float* Aprev, size_t nvox, hipStream_t s);
void launch_residual(const float* B, float* Aprev, size_t nvox, unsigned long long* d_max, hipStream_t s);
void launch_alphas(const float* B, const uint8_t* upd, float* out, size_t nvox, hipStream_t s);
void launch_fill_f32(float* p, float v, size_t n, hipStream_t s);
void launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s);
void launch_zero_words(void* p, size_t bytes, hipStream_t s);

}  // namespace mrf
