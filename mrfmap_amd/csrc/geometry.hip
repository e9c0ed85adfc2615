// Per-camera segment layout of the incidence lists (built once per
// traversal, not per BP iteration).
//
// The camera's incidences are stably sorted by voxel (hipCUB radix sort of
// (voxel, incidence index)).  Each run of equal voxels is one
// (camera, voxel) pair with a segment [pair_off[q], pair_off[q+1]) of the
// sorted order; every incidence record is rewritten to hold its sorted
// position instead of its voxel, and a parallel array receives its pair
// index.  The BP iteration reads the voxel->ray message of a record from the
// compact per-pair array (a few tens of MB per camera, cache resident) and
// writes the ray->voxel message into the record's slot of the segment array.  The BP iteration then needs no atomics:
// messages are written into the incidence's own slot of the segment array
// and each pair sums its segment in a fixed order, so results are
// bit-reproducible.  The pair's summed chord length C (what
// render_cumulative_length_kernel accumulates with float atomics every camera
// pass, src/cuda/gvdbkernels.cu:519-571, 1276-1302) is summed here once with
// the same deterministic segment sums.
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

__global__ void k_sorted_len(const uint32_t* __restrict__ vals, uint32_t nlive, const uint2* __restrict__ rec,
                             float* __restrict__ len) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlive) return;
  len[j] = __uint_as_float(rec[vals[j]].y);
}

// record -> its sorted position, and its pair: heads before position j plus
// j's own head, minus one
__global__ void k_set_pos(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ heads,
                          const uint32_t* __restrict__ hscan, uint32_t nlive, uint2* __restrict__ rec,
                          uint32_t* __restrict__ pair) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nlive) return;
  const uint32_t i = vals[j];
  rec[i].x = j;
  pair[i] = hscan[j] + heads[j] - 1u;
}

__global__ void k_set_last(uint32_t* p, uint32_t v) { *p = v; }

// record -> voxel (for export)
__global__ void k_export_vox(const uint2* __restrict__ rec, const uint32_t* __restrict__ pair, uint32_t n,
                             const uint32_t* __restrict__ pair_vox, uint2* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint2 e = rec[i];
  if (e.x != kGapVox) e.x = pair_vox[pair[i]];
  out[i] = e;
}

__global__ void k_pack_heads(const uint32_t* __restrict__ heads, uint32_t nlive, uint64_t* __restrict__ bits) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;  // blockDim is a multiple of 64
  const bool h = j < nlive && heads[j] != 0u;
  const unsigned long long word = __ballot(h);
  if ((threadIdx.x & 63) == 0 && j < nlive) bits[j >> 6] = word;
}

__global__ void k_tile_h0(const uint32_t* __restrict__ hscan, uint32_t nlive, uint32_t* __restrict__ tile_h0) {
  const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w * kTile < nlive) tile_h0[w] = hscan[w * kTile];
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

void finish_segments(SegmentBuild& b, uint32_t npairs, uint32_t nlive, uint32_t* pair_off, uint32_t* pair_vox,
                     float* pair_c, const uint64_t* head_bits, const uint32_t* tile_h0, uint32_t* rec_pair,
                     hipStream_t s) {
  const uint32_t n = b.nrec;
  if (n == 0) return;
  hipLaunchKernelGGL(k_pairs, dim3(blocks(n, 256)), dim3(256), 0, s, b.keys_out, b.heads, b.hscan, n, pair_off,
                     pair_vox);
  hipLaunchKernelGGL(k_set_last, dim3(1), dim3(1), 0, s, pair_off + npairs, nlive);
  if (nlive) {
    // chord lengths in sorted order (into keys_in, free after the sort), then
    // their deterministic segment sums; per-tile scratch carved from vals_in
    float* len = (float*)b.keys_in;
    hipLaunchKernelGGL(k_sorted_len, dim3(blocks(nlive, 256)), dim3(256), 0, s, b.vals_out, nlive, b.rec, len);
    const uint32_t nt = (nlive + kTile - 1) / kTile;
    uint32_t* scratch = b.vals_in;
    launch_segment_sum_f32(head_bits, tile_h0, nlive, len, pair_c, scratch, (float*)(scratch + nt), scratch + 2 * nt,
                           (float*)(scratch + 3 * nt), (float*)(scratch + 4 * nt), pair_off, s);
    hipLaunchKernelGGL(k_set_pos, dim3(blocks(nlive, 256)), dim3(256), 0, s, b.vals_out, b.heads, b.hscan, nlive,
                       b.rec, rec_pair);
  }
}

void launch_pack_heads(const SegmentBuild& b, uint32_t nlive, uint64_t* bits, uint32_t* tile_h0, hipStream_t s) {
  if (nlive == 0) return;
  hipLaunchKernelGGL(k_pack_heads, dim3(blocks(nlive, 256)), dim3(256), 0, s, b.heads, nlive, bits);
  const uint32_t ntiles = (nlive + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_tile_h0, dim3(blocks(ntiles, 256)), dim3(256), 0, s, b.hscan, nlive, tile_h0);
}

void launch_export_vox(const uint2* rec, const uint32_t* rec_pair, uint32_t n, const uint32_t* pair_vox, uint2* out,
                       hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_export_vox, dim3(blocks(n, 256)), dim3(256), 0, s, rec, rec_pair, n, pair_vox, out);
}

}  // namespace mrf
