// Per-camera pair and window-segment layout of the incidence lists (built
// once per traversal, not per BP iteration).
//
// Pairs.  Every voxel the camera's live records touch is one (camera, voxel)
// pair.  Two byte flags over the brick pool -- touched by any record, by a
// window record -- number the pairs by their ranks (counts per brick, a scan
// over bricks, ranks inside the brick): the pairs with window records first,
// then the others, each group in voxel order, so that window segment m
// belongs to pair m.
//
// Per pair, the summed chord length C (what render_cumulative_length_kernel
// accumulates with float atomics in every camera pass,
// src/cuda/gvdbkernels.cu:519-571, 1276-1302) and the chord length of its
// front records (whose messages are the constant F_EPS * l) are summed in
// 2^-24 fixed point as 64-bit values kept in two 32-bit words: a 32-bit
// atomic add to the low word, plus one to the high word when it wraps.  The
// sums are exact, so they do not depend on the order in which records
// arrive.  The part of each length below 2^-24 is summed the same way in
// units of 2^-48, so C is exact for every chord of at least 2^-25 and the
// epilogue gate C > 0.1f (gvdbkernels.cu:408) falls as it does on the exact
// sum, however close to 0.1f C lies.
//
// Window segments.  The window records are stably radix-sorted by pair, so
// each pair's records stay in record order; the BP iteration sums each
// segment in that order, gathered through win_idx (no atomics,
// bit-reproducible).  The pass that sums the front part of C also rewrites
// every live record to hold its pair instead of its voxel and writes the
// sort's input.
#include <hipcub/hipcub.hpp>

#include "mrf_kernels.h"

namespace mrf {

namespace {

constexpr double kFixLen = 16777216.0;  // 2^24: chord lengths (< 2) in fixed point, a 64-record sum < 2^31

// Pair numbering by brick (one wavefront per brick of the pool, 8 voxels
// per lane, the byte flags read 8 at a time): a voxel's rank among the
// touched voxels in voxel order is its brick's offset plus the touched
// voxels before it in the brick.

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t x) {
  for (int o = 32; o > 0; o >>= 1) x += (uint32_t)__shfl_xor((int)x, o, 64);
  return x;
}

__global__ __launch_bounds__(256) void k_brick_flag_counts(const uint8_t* __restrict__ any,
                                                           const uint8_t* __restrict__ wany, uint32_t nbricks,
                                                           uint64_t* __restrict__ counts) {
  const uint32_t b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= nbricks) return;  // uniform per wavefront
  const uint32_t l = threadIdx.x & 63u;
  const uint64_t a = reinterpret_cast<const uint64_t*>(any)[(size_t)b * 64 + l];
  const uint64_t w = reinterpret_cast<const uint64_t*>(wany)[(size_t)b * 64 + l];
  const uint32_t na = wave_sum_u32((uint32_t)__popcll(a)), nw = wave_sum_u32((uint32_t)__popcll(w));
  if (l == 0) counts[b] = (uint64_t)na | ((uint64_t)nw << 32);
}

// pid[v] and pair_vox[pid] of every touched voxel (window pairs first, the
// ranks of wany; then the others, nwp + the ranks of any minus those of
// wany); the flags are cleared behind (all zero again for the next camera)
__global__ __launch_bounds__(256) void k_pair_table(uint8_t* __restrict__ any, uint8_t* __restrict__ wany,
                                                    const uint64_t* __restrict__ boff, uint32_t nbricks,
                                                    uint32_t nwp, uint32_t* __restrict__ pid,
                                                    uint32_t* __restrict__ pair_vox) {
  const uint32_t b = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (b >= nbricks) return;  // uniform per wavefront
  const uint32_t l = threadIdx.x & 63u;
  uint64_t* A = reinterpret_cast<uint64_t*>(any) + (size_t)b * 64 + l;
  uint64_t* Wf = reinterpret_cast<uint64_t*>(wany) + (size_t)b * 64 + l;
  const uint64_t a = *A, w = *Wf;
  if (__ballot(a != 0) == 0) return;  // no touched voxel in the brick (uniform)
  // exclusive prefix of the lanes' counts
  const uint32_t na = (uint32_t)__popcll(a), nw = (uint32_t)__popcll(w);
  uint32_t ia = na, iw = nw;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t ua = (uint32_t)__shfl_up((int)ia, o, 64), uw = (uint32_t)__shfl_up((int)iw, o, 64);
    if (l >= (uint32_t)o) {
      ia += ua;
      iw += uw;
    }
  }
  const uint64_t o = boff[b];
  uint32_t ra = (uint32_t)o + ia - na, rw = (uint32_t)(o >> 32) + iw - nw;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const bool fa = ((a >> (8 * k)) & 1u) != 0, fw = ((w >> (8 * k)) & 1u) != 0;
    if (fa) {
      const uint32_t v = b * (uint32_t)kBrickVox + l * 8u + (uint32_t)k;
      const uint32_t p = fw ? rw : nwp + (ra - rw);
      pid[v] = p;
      pair_vox[p] = v;
    }
    ra += fa ? 1u : 0u;
    rw += fw ? 1u : 0u;
  }
  if (a) *A = 0;
  if (w) *Wf = 0;
}

// value of lane ^ J; J < 32 stays inside the 32-lane half (ds_swizzle, bitmask mode)
template <int J>
__device__ __forceinline__ uint32_t xor_lane(uint32_t v) {
  if constexpr (J < 32)
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (J << 10) | 0x1F);
  else
    return (uint32_t)__shfl_xor((int)v, J, 64);
}

// one bitonic step: lanes l and l ^ J order (key, val) ascending inside
// blocks of K lanes whose index bit K is 0, descending otherwise
template <int K, int J>
__device__ __forceinline__ void bitonic_step(int lane, uint32_t& key, uint32_t& val, uint32_t& val2) {
  const uint32_t ok = xor_lane<J>(key), ov = xor_lane<J>(val), ov2 = xor_lane<J>(val2);
  const bool up = (lane & K) == 0;
  const bool lower = (lane & J) == 0;
  const bool take = (lower == up) ? (ok < key) : (ok > key);
  key = take ? ok : key;
  val = take ? ov : val;
  val2 = take ? ov2 : val2;
}

template <int K, int J>
__device__ __forceinline__ void bitonic_merge(int lane, uint32_t& key, uint32_t& val, uint32_t& val2) {
  bitonic_step<K, J>(lane, key, val, val2);
  if constexpr (J > 1) bitonic_merge<K, J / 2>(lane, key, val, val2);
}

// sorts key ascending over the K lanes, val and val2 travel with it
template <int K>
__device__ __forceinline__ void bitonic_sort(int lane, uint32_t& key, uint32_t& val, uint32_t& val2) {
  if constexpr (K > 2) bitonic_sort<K / 2>(lane, key, val, val2);
  bitonic_merge<K, K / 2>(lane, key, val, val2);
}

// Sort of one 32-bit value per lane over the 64 lanes, ascending; every
// compare-exchange keeps the minimum in the lower lane.  Merging two sorted
// halves of K lanes starts with the mirror partner l ^ (K - 1), then the
// half-cleaners l ^ J, J = K/4 .. 1.
template <int J>
__device__ __forceinline__ void cx_min_low(int lane, int low_bit, uint32_t& v) {
  const uint32_t o = xor_lane<J>(v);
  v = (lane & low_bit) == 0 ? min(v, o) : max(v, o);
}

template <int K, int J>
__device__ __forceinline__ void half_clean(int lane, uint32_t& v) {
  cx_min_low<J>(lane, J, v);
  if constexpr (J > 1) half_clean<K, J / 2>(lane, v);
}

template <int K>
__device__ __forceinline__ void sort_u32(int lane, uint32_t& v) {
  if constexpr (K > 2) sort_u32<K / 2>(lane, v);
  cx_min_low<K - 1>(lane, K / 2, v);
  if constexpr (K > 2) half_clean<K, K / 4>(lane, v);
}

template <int CTRL, int ROWS>
__device__ __forceinline__ uint32_t dpp_u(uint32_t x) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROWS, 0xF, true);
}

// (x, f) <- (x_left, f_left) (+) (x, f): add the left value unless a run
// starts at or before this lane
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_step_u(uint32_t& x, uint32_t& x2, uint32_t& f) {
  const uint32_t y = dpp_u<CTRL, ROWS>(x);
  const uint32_t y2 = dpp_u<CTRL, ROWS>(x2);
  const uint32_t g = dpp_u<CTRL, ROWS>(f);
  x = f ? x : x + y;
  x2 = f ? x2 : x2 + y2;
  f |= g;
}

// Inclusive segmented sum over the 64 lanes (all active): row_shr 1/2/4/8
// inside each row of 16, then row_bcast 15 and 31.
__device__ __forceinline__ void wave_seg_sum_u(uint32_t& x, uint32_t& x2, uint32_t f) {
  seg_step_u<0x111, 0xF>(x, x2, f);
  seg_step_u<0x112, 0xF>(x, x2, f);
  seg_step_u<0x114, 0xF>(x, x2, f);
  seg_step_u<0x118, 0xF>(x, x2, f);
  seg_step_u<0x142, 0xA>(x, x2, f);
  seg_step_u<0x143, 0xC>(x, x2, f);
}

// A wave takes 64 consecutive records -- one step of the 64 rays of a pixel
// tile, whose voxels repeat across neighbouring rays -- rewrites each live
// record's voxel to its pair, sorts the 64 (pair, fixed-point length) by
// pair across its lanes (bitonic network), sums each run of equal pairs
// (segmented scan) and issues one atomic per run instead of one per record.
// Regions are whole rows of 64 records, so a row is either all front or all
// window.  Front runs go to the front sum acc_f and their parts below 2^-24
// to acc_r; a window row only has its records rewritten: the window part of
// C is summed after the window sort, where each pair's records are already
// contiguous (k_window_lengths).  C is the window part plus acc_f plus acc_r.
// PACKED (fewer than 2^26 - 1 pairs): the sort moves one word per lane, the
// pair and the lane it came from (pair 2^6 + lane, unique), and each lane then
// fetches its lengths from that lane; else (pair, q, r) move together.
// The same pass writes the window region's sort input: per window record j
// (index base_w + j) key its pair (sentinel for a gap or padding record,
// sorting last) and value j.
template <bool PACKED>
__global__ void __launch_bounds__(256) k_pair_lengths(uint2* __restrict__ rec, uint64_t nrec, uint64_t base_w,
                                                      const uint32_t* __restrict__ pid,
                                                      uint32_t* __restrict__ acc_f, uint32_t* __restrict__ acc_r,
                                                      uint32_t* __restrict__ wkeys, uint32_t* __restrict__ wvals,
                                                      uint32_t sentinel) {
  // batches of kR rows per wave (nrec is a multiple of 64): the batch's
  // records and then their pairs are loaded together, with no branches (a
  // row past the end reads row 0 as gap records); the rows are sorted and
  // summed; then the runs' exact adds {low word += q, high word += 1 if it
  // wrapped} are issued, and their carries taken once the old values of all
  // the batch's adds are back
  constexpr int kR = 8;
  const int lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t nrows = nrec / 64;
  for (uint64_t row0 = wave * kR; row0 < nrows; row0 += nwaves * kR) {
    uint2 e[kR];
    uint32_t key[kR], q[kR], r[kR];
#pragma unroll
    for (int j = 0; j < kR; ++j) {
      const bool in = row0 + j < nrows;
      e[j] = rec[(in ? row0 + j : 0) * 64 + lane];
      if (!in) e[j].x = kGapVox;
    }
#pragma unroll
    for (int j = 0; j < kR; ++j) key[j] = pid[e[j].x != kGapVox ? e[j].x : 0u];
#pragma unroll
    for (int j = 0; j < kR; ++j) {
      q[j] = r[j] = 0;
      // every record of the row is written back (gap records unchanged), so
      // the row's lines are written whole
      if (row0 + j < nrows)  // uniform per row
        rec[(row0 + j) * 64 + lane] = e[j].x != kGapVox ? make_uint2(key[j], e[j].y) : e[j];
      if (e[j].x != kGapVox) {
        // l 2^24 is exact in double; q its integer part (< 2^25), r the rest
        // in units of 2^-48 (< 2^24)
        const double x = (double)__uint_as_float(e[j].y) * kFixLen;
        q[j] = (uint32_t)x;
        r[j] = (uint32_t)((x - (double)q[j]) * kFixLen);
      } else {
        key[j] = 0xFFFFFFFFu;
      }
      if ((row0 + j) * 64 >= base_w) {  // a window row (uniform): the sort input only
        if (row0 + j < nrows) {
          const uint64_t jw = (row0 + j) * 64 - base_w + lane;
          wkeys[jw] = key[j] == 0xFFFFFFFFu ? sentinel : key[j];
          wvals[jw] = (uint32_t)jw;
        }
        key[j] = 0xFFFFFFFFu;
        continue;
      }
      if constexpr (PACKED) {
        uint32_t v = key[j] == 0xFFFFFFFFu ? 0xFFFFFFFFu : (key[j] << 6) | (uint32_t)lane;
        sort_u32<64>(lane, v);
        const int src = (int)(v & 63u);
        key[j] = v == 0xFFFFFFFFu ? 0xFFFFFFFFu : v >> 6;
        q[j] = (uint32_t)__shfl((int)q[j], src, 64);
        r[j] = (uint32_t)__shfl((int)r[j], src, 64);
      } else {
        bitonic_sort<64>(lane, key[j], q[j], r[j]);
      }
      const uint32_t prev = (uint32_t)__shfl_up((int)key[j], 1, 64);
      const uint32_t next = (uint32_t)__shfl_down((int)key[j], 1, 64);
      wave_seg_sum_u(q[j], r[j], lane == 0 || prev != key[j] ? 1u : 0u);
      // one add per run, by its last lane
      if (!(lane == 63 || next != key[j])) key[j] = 0xFFFFFFFFu;
    }
    uint32_t po[kR], por[kR];
#pragma unroll
    for (int j = 0; j < kR; ++j) {
      po[j] = por[j] = 0;
      if (key[j] != 0xFFFFFFFFu) {
        po[j] = atomicAdd(acc_f + 2 * (size_t)key[j], q[j]);
        if (r[j]) por[j] = atomicAdd(acc_r + 2 * (size_t)key[j], r[j]);
      }
    }
#pragma unroll
    for (int j = 0; j < kR; ++j) {
      if (key[j] == 0xFFFFFFFFu) continue;
      if (po[j] + q[j] < po[j]) atomicAdd(acc_f + 2 * (size_t)key[j] + 1, 1u);
      if (r[j] && por[j] + r[j] < por[j]) atomicAdd(acc_r + 2 * (size_t)key[j] + 1, 1u);
    }
  }
}

// The window part of C of every window pair, from the sorted window records
// (position m holds pair keys[m], record win_idx[m] of the window region): a
// wave takes 64 consecutive positions, computes the same exact fixed-point
// parts as k_pair_lengths, sums each run of equal pairs (segmented scan; the
// runs are already contiguous) and adds it with one exact add per run into
// acc_w (2^-24 units) and acc_r (below): one add per segment and wave it
// touches, instead of one per run of every window row.
__global__ __launch_bounds__(256) void k_window_lengths(const uint2* __restrict__ recw,
                                                        const uint32_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ win_idx, uint32_t nwin,
                                                        uint32_t* __restrict__ acc_w, uint32_t* __restrict__ acc_r) {
  const int lane = threadIdx.x & 63;
  const uint32_t m0 = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u;
  if (m0 >= nwin) return;  // uniform per wavefront
  const uint32_t m = m0 + (uint32_t)lane;
  const bool valid = m < nwin;
  uint32_t key = 0xFFFFFFFFu, q = 0, r = 0;
  if (valid) {
    key = keys[m];
    const double x = (double)__uint_as_float(recw[win_idx[m]].y) * kFixLen;
    q = (uint32_t)x;
    r = (uint32_t)((x - (double)q) * kFixLen);
  }
  const uint32_t prev = (uint32_t)__shfl_up((int)key, 1, 64);
  const uint32_t next = (uint32_t)__shfl_down((int)key, 1, 64);
  wave_seg_sum_u(q, r, lane == 0 || prev != key ? 1u : 0u);
  if ((lane == 63 || next != key) && key != 0xFFFFFFFFu) {
    uint32_t* pw = acc_w + 2 * (size_t)key;
    const uint32_t o = atomicAdd(pw, q);
    if (o + q < o) atomicAdd(pw + 1, 1u);
    if (r) {
      uint32_t* pr = acc_r + 2 * (size_t)key;
      const uint32_t orr = atomicAdd(pr, r);
      if (orr + r < orr) atomicAdd(pr + 1, 1u);
    }
  }
}

// C in FP32 with the gate of the exact sum (C > 0.1f, gvdbkernels.cu:408):
// where rounding to float would move C onto kCumLenMin from above, C takes
// the next float up, so that the float gate falls as on the exact sum
__global__ void k_pair_finish(const uint32_t* __restrict__ acc_w, const uint32_t* __restrict__ acc_f,
                              const uint32_t* __restrict__ acc_r, uint32_t np, uint32_t nwp,
                              float* __restrict__ pair_c, float* __restrict__ pair_sf) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= np) return;
  // integers below 2^53: the doubles and their sum are exact; a front-only
  // pair has no window part
  const double w = p < nwp ? (double)acc_w[2 * p + 1] * 4294967296.0 + (double)acc_w[2 * p] : 0.0;
  const double cr = (double)acc_r[2 * p + 1] * 4294967296.0 + (double)acc_r[2 * p];
  const double f = (double)acc_f[2 * p + 1] * 4294967296.0 + (double)acc_f[2 * p];
  const double cd = (w + f) / kFixLen + cr / (kFixLen * kFixLen);
  float cf = (float)cd;
  if (cd > (double)kCumLenGate) {
    if (!(cf > kCumLenMin)) cf = __uint_as_float(__float_as_uint(kCumLenMin) + 1u);
  } else if (cf > kCumLenMin) {
    cf = kCumLenMin;
  }
  pair_c[p] = cf;
  pair_sf[p] = (float)((double)kFEps * (f / kFixLen));
}

__global__ void k_window_heads(const uint32_t* __restrict__ keys, uint32_t nwin, uint32_t* __restrict__ whead) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nwin) return;
  whead[m] = (m == 0 || keys[m - 1] != keys[m]) ? 1u : 0u;
}

__global__ void k_window_offsets(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ whead, uint32_t nwin,
                                 uint32_t nwp, uint32_t* __restrict__ wseg_off) {
  const uint32_t m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < nwin && whead[m]) wseg_off[keys[m]] = m;
  if (m == 0) wseg_off[nwp] = nwin;
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
inline unsigned stride_blocks(uint64_t n) {
  const uint64_t b = (n + 255) / 256;
  return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

}  // namespace

size_t window_sort_temp_bytes(uint32_t n, int bits) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                           (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, bits);
  return bytes;
}

void launch_brick_flag_counts(const uint8_t* any, const uint8_t* wany, uint32_t nbricks, uint64_t* counts,
                              hipStream_t s) {
  if (nbricks == 0) return;
  hipLaunchKernelGGL(k_brick_flag_counts, dim3(blocks(nbricks, 4)), dim3(256), 0, s, any, wany, nbricks, counts);
}

void launch_pair_table(uint8_t* any, uint8_t* wany, const uint64_t* boff, uint32_t nbricks, uint32_t nwp,
                       uint32_t* pid, uint32_t* pair_vox, hipStream_t s) {
  if (nbricks == 0) return;
  hipLaunchKernelGGL(k_pair_table, dim3(blocks(nbricks, 4)), dim3(256), 0, s, any, wany, boff, nbricks, nwp, pid,
                     pair_vox);
}

void launch_pair_lengths(uint2* rec, uint64_t nrec, uint64_t base_w, const uint32_t* pid, uint32_t* acc_f,
                         uint32_t* acc_r, uint32_t np, uint32_t* wkeys, uint32_t* wvals, uint32_t sentinel,
                         hipStream_t s) {
  if (np == 0) return;
  launch_zero_words(acc_f, (size_t)np * 8, s);
  launch_zero_words(acc_r, (size_t)np * 8, s);
  if (np < (1u << 26) - 1u)
    hipLaunchKernelGGL(k_pair_lengths<true>, dim3(stride_blocks(nrec)), dim3(256), 0, s, rec, nrec, base_w, pid, acc_f,
                       acc_r, wkeys, wvals, sentinel);
  else
    hipLaunchKernelGGL(k_pair_lengths<false>, dim3(stride_blocks(nrec)), dim3(256), 0, s, rec, nrec, base_w, pid,
                       acc_f, acc_r, wkeys, wvals, sentinel);
}

void launch_pair_finish(const uint2* recw, const uint32_t* keys, const uint32_t* win_idx, uint32_t nwin,
                        uint32_t nwp, uint32_t* acc_w, const uint32_t* acc_f, uint32_t* acc_r, uint32_t np,
                        float* pair_c, float* pair_sf, hipStream_t s) {
  if (np == 0) return;
  launch_zero_words(acc_w, (size_t)nwp * 8, s);
  if (nwin)
    hipLaunchKernelGGL(k_window_lengths, dim3(blocks(nwin, 256)), dim3(256), 0, s, recw, keys, win_idx, nwin, acc_w,
                       acc_r);
  hipLaunchKernelGGL(k_pair_finish, dim3(blocks(np, 256)), dim3(256), 0, s, acc_w, acc_f, acc_r, np, nwp, pair_c,
                     pair_sf);
}

hipError_t sort_window(uint32_t nwr, int bits, uint32_t* keys_in, uint32_t* keys_out, uint32_t* vals_in,
                       uint32_t* vals_out, void* temp, size_t temp_bytes, hipStream_t s) {
  if (nwr == 0) return hipSuccess;
  size_t tb = temp_bytes;
  return hipcub::DeviceRadixSort::SortPairs(temp, tb, keys_in, keys_out, vals_in, vals_out, (int)nwr, 0, bits, s);
}

void launch_window_heads(const uint32_t* keys, uint32_t nwin, uint32_t* whead, hipStream_t s) {
  if (nwin == 0) return;
  hipLaunchKernelGGL(k_window_heads, dim3(blocks(nwin, 256)), dim3(256), 0, s, keys, nwin, whead);
}

void launch_window_offsets(const uint32_t* keys, const uint32_t* whead, uint32_t nwin, uint32_t nwp,
                           uint32_t* wseg_off, hipStream_t s) {
  hipLaunchKernelGGL(k_window_offsets, dim3(blocks(nwin ? nwin : 1, 256)), dim3(256), 0, s, keys, whead, nwin, nwp,
                     wseg_off);
}

void launch_pack_flags(const uint32_t* flags, const uint32_t* fscan, uint32_t n, uint64_t* bits, uint32_t* tile_h0,
                       hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_pack_flags, dim3(blocks(n, 256)), dim3(256), 0, s, flags, n, bits);
  const uint32_t ntiles = (n + kTile - 1) / kTile;
  hipLaunchKernelGGL(k_tile_h0, dim3(blocks(ntiles, 256)), dim3(256), 0, s, fscan, n, tile_h0);
}

void launch_export_vox(const uint2* rec, uint32_t n, const uint32_t* pair_vox, uint2* out, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_export_vox, dim3(blocks(n, 256)), dim3(256), 0, s, rec, n, pair_vox, out);
}

}  // namespace mrf
