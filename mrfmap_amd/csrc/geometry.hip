// Per-camera pair and segment layout of the incidence lists (built once per
// traversal, not per BP iteration).
//
// The camera's incidences are stably sorted by voxel (hipCUB radix sort of
// (voxel, incidence index)); each run of equal voxels is one (camera, voxel)
// pair.  Per pair, two deterministic segment sums over the sorted order give
// the summed chord length C (what render_cumulative_length_kernel
// accumulates with float atomics every camera pass,
// src/cuda/gvdbkernels.cu:519-571, 1276-1302) and the sum of the constant
// messages F_EPS * l of its front incidences (those where the depth
// likelihood is exactly 0).  The window incidences, in sorted order, form
// the window segments of the BP iteration: the ray kernel writes each window
// message into the incidence's own slot of the window region and each pair
// sums its segment, gathered through win_idx, in a fixed order, so the
// iteration needs no atomics and is bit-reproducible.  Pairs are renumbered with the pairs that have window
// incidences first (both groups in voxel order), so that window segment m
// belongs to pair m, and every record is rewritten to hold its pair.
#include <hipcub/hipcub.hpp>

#include "mrf_kernels.h"

namespace mrf {

namespace {

__global__ void k_keys(const uint2* __restrict__ rec, uint32_t n, uint32_t gap_key, uint32_t* __restrict__ keys,
                       uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = rec[i].x;
  keys[i] = v == kGapVox ? gap_key : v;
  vals[i] = i;
}

__global__ void k_heads(const uint32_t* __restrict__ keys, uint32_t n, uint32_t gap_key,
                        uint32_t* __restrict__ heads) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t k = keys[j];
  heads[j] = (k != gap_key && (j == 0 || keys[j - 1] != k)) ? 1u : 0u;
}

__global__ void k_pairs(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ heads,
                        const uint32_t* __restrict__ hscan, uint32_t n, uint32_t* __restrict__ pair_off,
                        uint32_t* __restrict__ pair_vox) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n || !heads[j]) return;
  const uint32_t q = hscan[j];
  pair_off[q] = j;
  pair_vox[q] = keys[j];
}

__global__ void k_set_last(uint32_t* p, uint32_t v) { *p = v; }

__global__ void k_sorted_vals(const uint32_t* __restrict__ vals, uint32_t nlive, const uint2* __restrict__ rec,
                              uint32_t base_w, float* __restrict__ len, float* __restrict__ fmsg,
                              uint32_t* __restrict__ win) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlive) return;
  const uint32_t i = vals[j];
  const float l = __uint_as_float(rec[i].y);
  const bool w = i >= base_w;
  len[j] = l;
  fmsg[j] = w ? 0.0f : kFEps * l;
  win[j] = w ? 1u : 0u;
}

__global__ void k_pack_flags(const uint32_t* __restrict__ flags, uint32_t n, uint64_t* __restrict__ bits) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;  // blockDim is a multiple of 64
  const bool h = j < n && flags[j] != 0u;
  const unsigned long long word = __ballot(h);
  if ((threadIdx.x & 63) == 0 && j < n) bits[j >> 6] = word;
}

__global__ void k_tile_h0(const uint32_t* __restrict__ fscan, uint32_t n, uint32_t* __restrict__ tile_h0) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w * kTile < n) tile_h0[w] = fscan[w * kTile];
}

__global__ void k_window_pairs(const uint32_t* __restrict__ win, const uint32_t* __restrict__ wpos,
                               const uint32_t* __restrict__ heads, const uint32_t* __restrict__ hscan, uint32_t nlive,
                               uint32_t* __restrict__ wq) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlive || !win[j]) return;
  wq[wpos[j]] = hscan[j] + heads[j] - 1u;
}

__global__ void k_window_heads(const uint32_t* __restrict__ wq, uint32_t nwin, uint32_t* __restrict__ whead) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nwin) return;
  whead[m] = (m == 0 || wq[m - 1] != wq[m]) ? 1u : 0u;
}

__global__ void k_mark_window_pairs(const uint32_t* __restrict__ whead, const uint32_t* __restrict__ hw,
                                    const uint32_t* __restrict__ wq, uint32_t nwin, uint32_t* __restrict__ mark,
                                    uint32_t* __restrict__ wseg_off) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nwin || !whead[m]) return;
  mark[wq[m]] = 1u;
  wseg_off[hw[m]] = m;
}

__global__ void k_permute_pairs(const uint32_t* __restrict__ mark, const uint32_t* __restrict__ rank, uint32_t npairs,
                                uint32_t nwp, const uint32_t* __restrict__ vox_old, const float* __restrict__ c_old,
                                const float* __restrict__ sf_old, uint32_t* __restrict__ perm,
                                uint32_t* __restrict__ pair_vox, float* __restrict__ pair_c,
                                float* __restrict__ pair_sf) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  const uint32_t p = mark[q] ? rank[q] : nwp + (q - rank[q]);
  perm[q] = p;
  pair_vox[p] = vox_old[q];
  pair_c[p] = c_old[q];
  pair_sf[p] = sf_old[q];
}

__global__ void k_set_records(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ heads,
                              const uint32_t* __restrict__ hscan, const uint32_t* __restrict__ win,
                              const uint32_t* __restrict__ wpos, const uint32_t* __restrict__ perm, uint32_t nlive,
                              uint32_t base_w, uint2* __restrict__ rec, uint32_t* __restrict__ win_idx) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlive) return;
  const uint32_t i = vals[j];
  rec[i].x = perm[hscan[j] + heads[j] - 1u];
  if (win[j]) win_idx[wpos[j]] = i - base_w;
}

// record -> voxel (for export)
__global__ void k_export_vox(const uint2* __restrict__ rec, uint32_t n, const uint32_t* __restrict__ pair_vox,
                             uint2* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint2 e = rec[i];
  if (e.x != kGapVox) e.x = pair_vox[e.x];
  out[i] = e;
}

inline unsigned blocks(size_t n, int t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

size_t segment_sort_temp_bytes(uint32_t n, int bits) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, bits);
  return bytes;
}

hipError_t build_segments(SegmentBuild& b, hipStream_t s) {
  const uint32_t n = b.nrec;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_keys, dim3(blocks(n, 256)), dim3(256), 0, s, b.rec, n, b.gap_key, b.keys_in, b.vals_in);
  size_t tb = b.temp_bytes;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(b.temp, tb, b.keys_in, b.keys_out, b.vals_in, b.vals_out,
                                                    (int)n, 0, b.key_bits, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_heads, dim3(blocks(n, 256)), dim3(256), 0, s, b.keys_out, n, b.gap_key, b.heads);
  launch_exclusive_scan(b.heads, b.hscan, n, b.scan_tmp, b.d_total, s);
  return hipGetLastError();
}

void launch_pairs(const SegmentBuild& b, uint32_t npairs, uint32_t nlive, uint32_t* pair_off, uint32_t* pair_vox,
                  hipStream_t s) {
  if (b.nrec == 0) return;
  hipLaunchKernelGGL(k_pairs, dim3(blocks(b.nrec, 256)), dim3(256), 0, s, b.keys_out, b.heads, b.hscan, b.nrec,
                     pair_off, pair_vox);
  hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, pair_off + npairs, nlive);
}

void launch_sorted_vals(const SegmentBuild& b, uint32_t nlive, uint32_t base_w, float* len, float* fmsg,
                        uint32_t* win, hipStream_t s) {
  if (nlive == 0) return;
  hipLaunchKernelGGL(k_sorted_vals, dim3(blocks(nlive, 256)), dim3(256), 0, s, b.vals_out, nlive, b.rec, base_w, len,
                     fmsg, win);
}

void launch_pack_flags(const uint32_t* flags, const uint32_t* fscan, uint32_t n, uint64_t* bits, uint32_t* tile_h0,
                       hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_pack_flags, dim3(blocks(n, 256)), dim3(256), 0, s, flags, n, bits);
  const uint32_t ntiles = (n + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_tile_h0, dim3(blocks(ntiles, 256)), dim3(256), 0, s, fscan, n, tile_h0);
}

void launch_window_pairs(const SegmentBuild& b, const uint32_t* win, const uint32_t* wpos, uint32_t nlive,
                         uint32_t* wq, hipStream_t s) {
  if (nlive == 0) return;
  hipLaunchKernelGGL(k_window_pairs, dim3(blocks(nlive, 256)), dim3(256), 0, s, win, wpos, b.heads, b.hscan, nlive,
                     wq);
}

void launch_window_heads(const uint32_t* wq, uint32_t nwin, uint32_t* whead, hipStream_t s) {
  if (nwin == 0) return;
  hipLaunchKernelGGL(k_window_heads, dim3(blocks(nwin, 256)), dim3(256), 0, s, wq, nwin, whead);
}

void launch_mark_window_pairs(const uint32_t* whead, const uint32_t* hw, const uint32_t* wq, uint32_t nwin,
                              uint32_t nwp, uint32_t* mark, uint32_t* wseg_off, hipStream_t s) {
  if (nwin) hipLaunchKernelGGL(k_mark_window_pairs, dim3(blocks(nwin, 256)), dim3(256), 0, s, whead, hw, wq, nwin,
                               mark, wseg_off);
  hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, wseg_off + nwp, nwin);
}

void launch_permute_pairs(const uint32_t* mark, const uint32_t* rank, uint32_t npairs, uint32_t nwp,
                          const uint32_t* vox_old, const float* c_old, const float* sf_old, uint32_t* perm,
                          uint32_t* pair_vox, float* pair_c, float* pair_sf, hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_permute_pairs, dim3(blocks(npairs, 256)), dim3(256), 0, s, mark, rank, npairs, nwp, vox_old,
                     c_old, sf_old, perm, pair_vox, pair_c, pair_sf);
}

void launch_set_records(const SegmentBuild& b, const uint32_t* win, const uint32_t* wpos, const uint32_t* perm,
                        uint32_t nlive, uint32_t base_w, uint32_t* win_idx, hipStream_t s) {
  if (nlive == 0) return;
  hipLaunchKernelGGL(k_set_records, dim3(blocks(nlive, 256)), dim3(256), 0, s, b.vals_out, b.heads, b.hscan, win, wpos,
                     perm, nlive, base_w, b.rec, win_idx);
}

void launch_export_vox(const uint2* rec, uint32_t n, const uint32_t* pair_vox, uint2* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_export_vox, dim3(blocks(n, 256)), dim3(256), 0, s, rec, n, pair_vox, out);
}

}  // namespace mrf
