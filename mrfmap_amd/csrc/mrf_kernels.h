// Host-side launchers of the HIP kernels of libmrfmap_hip (internal).
#pragma once

#include "mrf_internal.h"

namespace mrf {

// Storage of the per-incidence messages: voxel->ray log2 n (input side) and
// ray->voxel p*l (output side).  The mixed types keep log2 n in FP32.
enum MsgType { kMsgF32 = 0, kMsgF16 = 1, kMsgBF16 = 2, kMsgMixedF16 = 3, kMsgMixedBF16 = 4 };
inline int msg_in_type(int t) { return t == kMsgF16 ? kMsgF16 : (t == kMsgBF16 ? kMsgBF16 : kMsgF32); }
inline int msg_out_type(int t) {
  return (t == kMsgF16 || t == kMsgMixedF16) ? kMsgF16 : ((t == kMsgBF16 || t == kMsgMixedBF16) ? kMsgBF16 : kMsgF32);
}
inline size_t msg_bytes(int type) { return type == kMsgF32 ? 4 : 2; }

// Schedules of the ray-message kernel: records loaded ahead of their
// arithmetic in the front loop (F) and in the window passes (W).  Picked per
// workload by the autotuner in api.hip.
enum RaySched { kSchedF8W4 = 0, kSchedF8W8 = 1, kSchedF16W4 = 2, kSchedF16W8 = 3, kSchedF4W4 = 4, kSchedF16W16 = 5,
                kNumSched = 6 };

// traverse.hip
// full-ray activation (bricks marked in bflags); bound: per tiled pixel, a
// bound of the ray's record count (may be null)
void launch_mark_rays(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                      const float* bia, uint8_t* bflags, uint32_t* bound, hipStream_t s);
void launch_mark_endpoints(const CamGeom& c, const GridGeom& g, const float* depth, uint8_t* bflags,
                           hipStream_t s);
// records per tiled pixel (endpoint activation: gap records included)
void launch_count(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                  const float* bia, const int32_t* slot, uint32_t* counts, hipStream_t s);
// slice (8x8 pixel tile) lengths: max records over the tile's 64 rays
void launch_slice_len(const uint32_t* counts, uint32_t ntiles, uint32_t* slice_len, uint32_t* slice_rec,
                      hipStream_t s);
// sliced-ELL emission: record k of lane l of slice t at slice_off[t] + k*64 + l
// for k below the ray's record count (nothing past it is written), split
// into front and window: front[t*64+l].nf front records, then
// lanes[t*64+l].nw window records (the ray's description, the Kahan state of
// the front's length sum); cnt_f / cnt_w the same per tiled pixel.  A ray
// with more records than its slice's length slice_len[t] (sized by a count
// or a bound of it) sets *overflow.
void launch_emit(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig, const float* bia,
                 const int32_t* slot, const uint32_t* slice_off, const uint32_t* slice_len, uint2* sell,
                 FrontRec* front, RayRec* lanes, uint32_t* cnt_f, uint32_t* cnt_w, uint32_t* overflow,
                 hipStream_t s);

// map evaluation of one camera's pixels (log-likelihood, expected depth,
// accuracy and simple accuracy images in row-major pixel order, any may be
// null); part gets eval_blocks(c) FP64 partial sums of the log-likelihoods,
// counts (7, zeroed by the caller) the accuracy classes -1, 0, 1 and the
// simple accuracy classes -1, 0, 1, 2
unsigned eval_blocks(const CamGeom& c);
void launch_eval(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig, const float* bia,
                 const int32_t* slot, const float* B, const uint8_t* upd, float acc_thresh, float* ll, float* ed,
                 float* acc, float* sacc, double* part, unsigned long long* counts, hipStream_t s);

// scan.hip: exclusive prefix sum of n uint32 values; total is written to
// *d_total (device, uint64).  tmp must hold scan_tmp_words(n) uint64 words.
// exclusive scan of 0/1 flags (n < 2^31), total into *d_total; temp holds
// flag_scan_temp_bytes(n) bytes
size_t flag_scan_temp_bytes(size_t n);
hipError_t launch_flag_scan(const uint32_t* in, uint32_t* out, size_t n, void* temp, size_t temp_bytes,
                            uint64_t* d_total, hipStream_t s);
// exclusive scan of n uint64 values (n < 2^31; sums below 2^64), total into
// *d_total; temp holds u64_scan_temp_bytes(n) bytes
size_t u64_scan_temp_bytes(size_t n);
hipError_t launch_u64_scan(const uint64_t* in, uint64_t* out, size_t n, void* temp, size_t temp_bytes,
                           uint64_t* d_total, hipStream_t s);

size_t scan_tmp_words(size_t n);
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint64_t* tmp, uint64_t* d_total,
                           hipStream_t s);

int num_tiled_pixels(const CamGeom& c);

// geometry.hip: (camera, voxel) pairs of a camera's incidences and the
// window segment layout of the BP iteration.
// per brick of the pool: the voxels with any / wany set, packed as
// count_any | count_wany << 32 (the flags are bytes of 0 or 1)
void launch_brick_flag_counts(const uint8_t* any, const uint8_t* wany, uint32_t nbricks, uint64_t* counts,
                              hipStream_t s);
// from the exclusive scan boff of those counts: pid[v] and pair_vox[pid] of
// every touched voxel v (window pairs first, each group in voxel order;
// other entries of pid are left as they are); clears any and wany behind
void launch_pair_table(uint8_t* any, uint8_t* wany, const uint64_t* boff, uint32_t nbricks, uint32_t nwp,
                       uint32_t* pid, uint32_t* pair_vox, hipStream_t s);
// per pair: the exact fixed-point sums of its front records' chord lengths
// (acc_f, 2 words per pair) and of the parts below 2^-24 of its front
// records' lengths (acc_r).  Rewrites every live record's voxel to its pair
// (pid), and writes the window sort input: wkeys[j] = pair of window record
// j (sentinel for a gap or padding record), wvals[j] = j.
void launch_pair_lengths(uint2* rec, uint64_t nrec, uint64_t base_w, const uint32_t* pid, uint32_t* acc_f,
                         uint32_t* acc_r, uint32_t np, uint32_t* wkeys, uint32_t* wvals, uint32_t sentinel,
                         hipStream_t s);
// after the window sort (sorted pair keys and win_idx of the nwin live
// window records): the window part of every window pair (acc_w; the parts
// below 2^-24 into acc_r), then per pair C (summed chord length, exact, FP32
// with the exact sum's gate) and F_EPS x the chord length of its front
// records
void launch_pair_finish(const uint2* recw, const uint32_t* keys, const uint32_t* win_idx, uint32_t nwin,
                        uint32_t nwp, uint32_t* acc_w, const uint32_t* acc_f, uint32_t* acc_r, uint32_t np,
                        float* pair_c, float* pair_sf, hipStream_t s);
size_t window_sort_temp_bytes(uint32_t n, int bits);
// window region records stably sorted by pair (gap / padding records, with
// the sentinel key, last): keys_out = pairs, vals_out = record index in the
// region
hipError_t sort_window(uint32_t nwr, int bits, uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_in,
                       uint32_t* vals_out, void* temp, size_t temp_bytes, hipStream_t s);
void launch_window_heads(const uint32_t* keys, uint32_t nwin, uint32_t* whead, hipStream_t s);
// wseg_off[pair] = first window position of the pair, wseg_off[nwp] = nwin
void launch_window_offsets(const uint32_t* keys, const uint32_t* whead, uint32_t nwin, uint32_t nwp,
                           uint32_t* wseg_off, hipStream_t s);
// flag bitmask (bit j = flags[j] != 0) and, per kTile tile, fscan at its start
void launch_pack_flags(const uint32_t* flags, const uint32_t* fscan, uint32_t n, uint64_t* bits, uint32_t* tile_h0,
                       hipStream_t s);
// records with their voxel ids restored (for export)
void launch_export_vox(const uint2* rec, uint32_t n, const uint32_t* pair_vox, uint2* out, hipStream_t s);

// bp.hip: one camera pass = per-pair voxel->ray messages -> ray messages ->
// window segment sums -> per-pair update.
struct CamPass {
  const FrontRec* front;   // one per lane of every slice
  const RayRec* lanes;     // one per lane of every slice
  const uint32_t* len_f;   // front records per lane of each slice
  const uint32_t* len_w;   // window records per lane of each slice
  uint32_t nslices;
  const uint2* inc;        // sliced records {pair | kGapVox, chord length}: front region, then window region
  const uint32_t* win_idx; // per window position (voxel order): its record's slot in the window region
  uint32_t base_w;         // first record of the window region
  const uint32_t* chunk_off;   // nslices + 1: first front chunk of each slice
  const uint4* chunk_desc;     // per front chunk: {chunk, slice, first row, end row}
  const uint2* fspan;          // per lane of every slice: {first front record, front records to sum}
  uint32_t nchunks;
  float* partial;          // per chunk and lane: front visibility partial sum
  const uint32_t* pair_vox;
  const float* pair_c;     // summed chord length C per pair
  const float* pair_sf;    // summed constant front messages per pair
  uint32_t npairs;
  uint32_t nwp;            // pairs with window incidences (numbered first)
  uint32_t nwin;           // window positions (live window incidences)
  const uint32_t* wseg_off;  // nwp + 1 window segment offsets
  const uint64_t* heads;   // window segment head bitmask
  const uint32_t* tile_h0; // window heads before each kTile tile
  void* apair;             // per-pair voxel -> ray message log2 n (input side type), kept per camera
  const uint2* dyn;        // front-only pairs at voxels where some camera has window records: {pair, voxel}
  uint32_t ndyn;
  const uint32_t* msg_all; // device countdown: nonzero while every pair's log2 n must be recomputed
  const uint4* dyn_desc;        // descriptors of the front chunks whose log2 n keep changing
  uint32_t ndyn_chunks;
  const uint32_t* fdmask;       // per front chunk and lane: bits of its records whose log2 n keeps changing
  float* fstatic;               // per front chunk and lane: sum of the others (first iterations of a call)
  const uint2* dfront;          // per changing chunk: its changing records, compacted per lane (rows of 64)
  const uint2* dcr;             // per entry of dyn_desc: {first row, rows} of its block in dfront
  float* S;                // per window pair: window segment sum
  float *pf_v, *pl_v, *mid;  // per-tile parts of the segments that cross tiles
  float* B;
  float* M;                // this camera's messages M_c (FP32, one per pair)
  void* Xo;                // ray->voxel messages p*l in window-region order (output side type)
  uint8_t* upd;
  // device flag: nonzero while every pair's message must be refreshed (the
  // first iteration of a call, or after messages were moved); otherwise only
  // the window pairs are updated (a front-only pair's message is the constant
  // logit(clamp(sf / C)) and already in M; its update would only add FP32
  // rounding to B)
  const uint32_t* all_pairs;
  const float* sig;        // sensor-model lookup tables
  const float* bia;
  int lookup_n;
  int msg_type;
  int ray_sched;           // RaySched
  // pre-pass visibility restarted at brick entries until a tau survives
  // flush-to-zero (MRF_VIS_BRICK_LOCAL): per window row of every slice the
  // lanes whose record enters a brick, and per lane the front records that
  // share the first window record's brick
  bool brick_local;
  const uint64_t* bstart;
  const uint8_t* tail;
};
// copy into the final front/window regions; window region starts at base_w.
// any / wany: pool-sized byte flags of voxels touched by a live / window
// record (set, never cleared); counts[0] / counts[1] += the live / live
// window records, counts[2] += all records (gap records included),
// counts[3] += rays with records
void launch_reslice(FrontRec* front, RayRec* lanes, uint32_t npix, const uint2* sell, const uint32_t* off_f,
                    const uint32_t* len_f, const uint32_t* off_w, const uint32_t* len_w, uint32_t base_w, uint2* inc,
                    uint8_t* any, uint8_t* wany, unsigned long long* counts, hipStream_t s);
// brick entries of every window record (records still hold voxel ids):
// bstart[(region row of the slice) + k] bit l = record k of lane l enters a
// brick (the ray's first record, the record after a gap, or a voxel in
// another brick than the previous record's); tail[lane] = front records at
// the end of the front that lie in the first window record's brick
void launch_brick_starts(const FrontRec* front, const RayRec* lanes, const uint32_t* len_w, uint32_t nslices,
                         const uint2* inc, uint32_t base_w, uint64_t* bstart, uint8_t* tail, hipStream_t s);
// front chunks per slice (cnt has nslices + 1 entries, the last 0), the
// descriptor of every chunk from their offsets, and the per-lane front spans
void launch_chunk_counts(const uint32_t* len_f, const uint32_t* len_w, uint32_t nslices, uint32_t* cnt, hipStream_t s);
void launch_chunk_desc(const uint32_t* chunk_off, const uint32_t* len_f, uint32_t nslices, uint4* desc, hipStream_t s);
void launch_front_span(const FrontRec* front, const RayRec* lanes, uint32_t n, uint2* span, hipStream_t s);
// out[j] = desc[idx[j]]
void launch_gather_desc(const uint4* desc, const uint32_t* idx, uint32_t n, uint4* out, hipStream_t s);
void launch_pair_msgs(const CamPass& a, hipStream_t s);
// part: 0 both kernels, 1 the front sums only, 2 the window kernel only
void launch_ray_messages(const CamPass& a, hipStream_t s, int part = 0);
// per front chunk of the camera: 0 if all its pairs are front-only pairs whose pflag (front-only
// pair q at pflag[q - nwp]: at a voxel of W) is 0 (static; its records are added to *nstatic), 1
// otherwise; per chunk and lane the mask of its changing records (dmask), per chunk the rows of
// its compact block (the most changing records of a lane)
void launch_chunk_static(const CamPass& a, const uint32_t* pflag, uint32_t* dyn_flag, uint32_t* dmask, uint32_t* rows,
                         unsigned long long* nstatic, hipStream_t s);
// rows of the compact block of every changing chunk (dyn_desc order)
void launch_dyn_rows(const uint4* dyn_desc, uint32_t n, const uint32_t* rows, uint32_t* out, hipStream_t s);
// the compact blocks of the changing chunks (a.dyn_desc, a.fdmask) at row_off, and dcr
void launch_fill_dfront(const CamPass& a, const uint32_t* rows, const uint32_t* row_off, uint2* dfront, uint2* dcr,
                        hipStream_t s);
// part: 0 both kernels, 1 the segment sums only, 2 the per-pair update only
void launch_epilogue(const CamPass& a, hipStream_t s, int part = 0);
void launch_marginals(const float* B, float* Aprev, size_t nvox, hipStream_t s);
void launch_residual(const float* B, float* Aprev, size_t nvox, unsigned long long* d_max, hipStream_t s);
void launch_ref_error(const float* A, const uint8_t* upd, const float* ref, size_t n, unsigned long long* d_max,
                      hipStream_t s);
void launch_alphas(const float* B, const uint8_t* upd, float* out, size_t nvox, hipStream_t s);
// load_map: B[vox[i]] = logodds[i], upd = 1 (distinct voxels)
void launch_set_voxels(const uint32_t* vox, const float* logodds, uint32_t n, float* B, uint8_t* upd, hipStream_t s);
// set_pose: summed chord lengths of a camera's pairs into cum; then voxels
// with cum < 0.05 back to the prior, and the moved camera's messages there
// cleared
void launch_cum_add(const uint32_t* pair_vox, const float* pair_c, uint32_t npairs, float* cum, hipStream_t s);
void launch_reset_empty(const float* cum, size_t n, float prior_logodds, float* B, uint8_t* upd, hipStream_t s);
void launch_reset_pair_msgs(const uint32_t* pair_vox, uint32_t npairs, const float* cum, float* M, hipStream_t s);
// camera messages per pair <-> per voxel (across a re-traversal)
void launch_scatter_pairs(const uint32_t* pair_vox, const float* M, uint32_t npairs, float* dense, hipStream_t s);
void launch_gather_pairs(const uint32_t* pair_vox, const float* dense, uint32_t npairs, float* M, hipStream_t s);
void launch_fill_f32(float* p, float v, size_t n, hipStream_t s);
void launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s);
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void launch_countdown(uint32_t* p, hipStream_t s);  // *p -= 1 unless 0
void launch_mark_window_voxels(const uint32_t* pair_vox, uint32_t nwp, uint32_t* W, hipStream_t s);
void launch_dyn_flags(const uint32_t* pair_vox, uint32_t n, const uint32_t* W, uint32_t* flags, hipStream_t s);
// the flagged indices q0 + k at dyn[pos[k]]
void launch_dyn_compact(const uint32_t* flags, const uint32_t* pos, uint32_t n, uint32_t q0, uint32_t* dyn,
                        hipStream_t s);
// the flagged pairs q0 + k as {pair, its voxel} at dyn[pos[k]]
void launch_dyn_pairs(const uint32_t* flags, const uint32_t* pos, uint32_t n, uint32_t q0, const uint32_t* pair_vox,
                      uint2* dyn, hipStream_t s);
void launch_zero_words(void* p, size_t bytes, hipStream_t s);

// ---- learned warm start (warm.hip): features, MLP, warm-start update
constexpr int kMlpIn = 32, kMlpHid = 64;
constexpr int kMlpW1 = 0;                          // [64][32]
constexpr int kMlpB1 = kMlpW1 + kMlpHid * kMlpIn;  // [64]
constexpr int kMlpW2 = kMlpB1 + kMlpHid;           // [64][64]
constexpr int kMlpB2 = kMlpW2 + kMlpHid * kMlpHid; // [64]
constexpr int kMlpW3 = kMlpB2 + kMlpHid;           // [64]
constexpr int kMlpB3 = kMlpW3 + kMlpHid;           // [1]
constexpr int kMlpParams = kMlpB3 + 1;             // 6337
uint32_t mlp_grad_waves(uint32_t cnt);
void launch_mlp_fwd(const float* P, const float* X, uint32_t n, float* y, hipStream_t s);
void launch_mlp_grad(const float* P, const float* X, const float* T, const uint32_t* sel, uint32_t off,
                     uint32_t cnt, float scale, float* partial, double* lpart, hipStream_t s);
void launch_mlp_reduce(const float* partial, const double* lpart, uint32_t nw, float* grad, double* loss,
                       double scale, hipStream_t s);
void launch_adam(float* P, const float* G, float* m, float* v, float lr, float b1, float b2, float eps, float c1,
                 float c2, hipStream_t s);
void launch_feat_hits(const CamGeom& c, const GridGeom& g, const float* depth, const int32_t* slot, uint32_t* hits,
                      hipStream_t s);
void launch_feat_pairs(const uint32_t* pair_vox, const float* pair_c, const float* M, uint32_t npairs, float* pass,
                       float* msum, float* mmax, uint32_t* nc, hipStream_t s);
void launch_feat_flags(const uint32_t* nc, size_t n, uint32_t* flag, hipStream_t s);
void launch_features(const GridGeom& g, const int32_t* slot, const int32_t* slot_lin, uint32_t nslots,
                     const uint32_t* hits, const float* pass, const float* msum, const float* mmax, const uint32_t* nc,
                     const float* B, float prior, const uint32_t* row, float* X, uint32_t* vox, hipStream_t s);
void launch_warm_apply(const uint32_t* vox, const float* pred, uint32_t n, const uint32_t* nc, float* B,
                       float* delta, hipStream_t s);
void launch_warm_pairs(const uint32_t* pair_vox, const float* pair_c, uint32_t npairs, const float* delta, float* M,
                       hipStream_t s);

}  // namespace mrf
