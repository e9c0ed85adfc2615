// Loopy-BP message kernels of the MRFMap inference path.
//
//  * k_pair_msg: voxel->ray message log2 n of every (camera, voxel) pair
//    (gvdbkernels.cu:764-773) into a compact per-pair array.
//  * k_ray_messages: L lanes per ray over sliced-ELL pixel tiles; reads log2 n
//    of each record's pair, writes p*l into the record's slot of the
//    camera's voxel-sorted segment array.  Replaces
//    the pre pass (rayStartComputingMessagesFromBrickPre,
//    src/cuda/gvdbkernels.cu:689-819), the infinity term
//    (gvdbkernels.cu:1196-1208) and the post pass
//    (rayStartComputingMessagesFromBrickPost, gvdbkernels.cu:824-1038) with a
//    single launch.
//  * k_seg_sum / k_seg_update: per pair, the fixed-order sum S of
//    its slots, m = logit(clamp(S/C)), B += m - M_c, M_c = m
//    (compute_gvdb_message_to_factors, gvdbkernels.cu:396-496, live 457-479).
//  * residual / marginal / readback kernels (gvdbkernels.cu:469-476, 360-373).
#include <hip/hip_fp16.h>

#include "mrf_kernels.h"

namespace mrf {

namespace {

// ---------------------------------------------------------------- storage
template <int MT>
struct Msg;
// kOut: factor applied to ray->voxel messages p*l before storage (FP16 keeps
// them out of its subnormal range: p*l >= 1e-4 * l).
template <>
struct Msg<kMsgF32> {
  typedef float T;
  static constexpr float kOut = 1.0f;
  __device__ static float load(const T* p, uint32_t i) { return p[i]; }
  __device__ static void store(T* p, uint32_t i, float v) { p[i] = v; }
};
template <>
struct Msg<kMsgF16> {
  typedef __half T;
  static constexpr float kOut = 4096.0f;
  __device__ static float load(const T* p, uint32_t i) { return __half2float(p[i]); }
  __device__ static void store(T* p, uint32_t i, float v) { p[i] = __float2half_rn(v); }
};
template <>
struct Msg<kMsgBF16> {
  typedef uint16_t T;
  static constexpr float kOut = 1.0f;
  __device__ static float load(const T* p, uint32_t i) { return __uint_as_float((uint32_t)p[i] << 16); }
  __device__ static void store(T* p, uint32_t i, float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    p[i] = (uint16_t)(u >> 16);
  }
};

// ---------------------------------------------------------------- wave ops
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xF, true));
}

// Inclusive prefix sum over the 64 lanes (all lanes must be active):
// row_shr 1/2/4/8 inside each row of 16, then row_bcast 15 and 31.
__device__ __forceinline__ float wave_scan(float x) {
  x += dpp<0x111, 0xF>(x);
  x += dpp<0x112, 0xF>(x);
  x += dpp<0x114, 0xF>(x);
  x += dpp<0x118, 0xF>(x);
  x += dpp<0x142, 0xA>(x);
  x += dpp<0x143, 0xC>(x);
  return x;
}

__device__ __forceinline__ float lane63(float x) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
}

// ---------------------------------------------------------------- sensor model
// sigma(z), bias(z): p0 + p1 z (+ p2 z^2) per 20x20-pixel lookup cell
// (get_depth_sigma / get_depth_bias, gvdbkernels.cu:46-55).
// A linear cell has c2 = 0, for which fmaf(z, fmaf(z, 0, c1), c0) is
// fmaf(z, c1, c0) exactly.

// 0.398942 e^{-u^2/2}, the exponential as a power of 2 with the constants
// folded: e^{-u^2/2} = 2^{u^2 (-log2(e) / 2)}
constexpr float kNegHalfLog2e = -0.72134752044448170f;
__device__ __forceinline__ float gauss_a(float u) { return 0.398942f * __builtin_amdgcn_exp2f((u * u) * kNegHalfLog2e); }

// psi(s; d) = N(d; s + bias(s), (3 sigma(s))^2) (depth_potential,
// gvdbkernels.cu:57-72), with the likelihood width 3 sigma(s) evaluated
// from sigma's coefficients times 3 (sig3)
__device__ __forceinline__ float depth_psi(float s, float d, const Poly& sig3, const Poly& bia) {
  const float diff = fabsf(d - (s + bia(s)));
  const float inv = __builtin_amdgcn_rcpf(sig3(s));
  return gauss_a(diff * inv) * inv;
}

// ---------------------------------------------------------------- segments
// A camera's incidences are stored pair-major: pair q owns positions
// [pair_off[q], pair_off[q+1]).  Segment heads are marked in a bitmask and
// each tile of kTile positions knows how many heads precede it, so the pair
// of every position follows from one 64-bit word and a popcount.  All
// segment kernels below work on position tiles (one wavefront per tile), so
// a pair with hundreds of thousands of incidences (the voxels at a camera's
// centre, which every ray of that camera crosses) is spread over many
// wavefronts instead of serialising one.

// Per-lane pair index of positions base + lane (word = head bits of the
// 64 positions starting at base, before = heads before base).
__device__ __forceinline__ uint32_t pair_of(uint64_t word, uint32_t before, int lane) {
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  return before + (uint32_t)__popcll(word & upto) - 1u;
}

// Voxel -> ray message of every pair: n = clamp(1/(1+exp(B - M_c))), camera
// c's own previous message excluded from the belief (gvdbkernels.cu:764-773),
// stored as log2 n (the reference's ln n / ln 2): the ray kernels sum
// l log2 n and take exp2 of the sums (one multiply fewer per exponential than
// e^x = 2^(x log2 e)).
template <int MT>
__device__ __forceinline__ void pair_msg_qv(uint32_t q, uint32_t v, const float* __restrict__ B,
                                            const float* __restrict__ M, typename Msg<MT>::T* __restrict__ apair) {
  const float qv = B[v] - M[q];
  float nn = 1.0f / (1.0f + __expf(qv));
  nn = fminf(fmaxf(nn, kFEps), 1.0f - kFEps);
  Msg<MT>::store(apair, q, __builtin_amdgcn_logf(nn));
}

template <int MT>
__device__ __forceinline__ void pair_msg_one(uint32_t q, const uint32_t* __restrict__ pair_vox,
                                             const float* __restrict__ B, const float* __restrict__ M,
                                             typename Msg<MT>::T* __restrict__ apair) {
  pair_msg_qv<MT>(q, pair_vox[q], B, M, apair);
}

// The window pairs [0, nwp) change every iteration; of the front-only pairs,
// all while *msg_all (the first iterations of a call), afterwards only the
// list dyn (front-only pairs at voxels that some camera reaches with window
// records, each with its voxel, so that its message and belief are loaded
// together).  One launch: the window pairs one per thread, the others
// strided over the grid.
template <int MT>
__global__ __launch_bounds__(256) void k_pair_msg(const uint32_t* __restrict__ pair_vox, uint32_t nwp,
                                                  uint32_t npairs, const uint2* __restrict__ dyn, uint32_t ndyn,
                                                  const float* __restrict__ B, const float* __restrict__ M,
                                                  typename Msg<MT>::T* __restrict__ apair,
                                                  const uint32_t* __restrict__ msg_all) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  for (uint32_t q = tid; q < nwp; q += stride) pair_msg_one<MT>(q, pair_vox, B, M, apair);
  if (*msg_all) {
    for (uint32_t q = nwp + tid; q < npairs; q += stride) pair_msg_one<MT>(q, pair_vox, B, M, apair);
  } else {
    for (uint32_t k = tid; k < ndyn; k += stride) {
      const uint2 e = dyn[k];  // {pair, voxel}
      pair_msg_qv<MT>(e.x, e.y, B, M, apair);
    }
  }
}

// ---------------------------------------------------------------- list layout
// Copy every ray's records into the front and window regions of the final
// record buffer (sliced: lane l of slice t at region offset + k * 64 + l),
// padding each lane to its slice's length with gap records.  One wavefront
// is one slice, so record k of the 64 rays moves together: per row the
// voxel flags any[v] = 1 (every live record) and wany[v] = 1 (live window
// records) are set, skipping a lane whose voxel equals that of the ray to
// its left or above in the pixel tile (the flags say only whether), and
// counts[0] / counts[1] gain the row's live / live window records; per
// wavefront counts[2] gains its rays' records (gap records included) and
// counts[3] its rays with records.
__global__ __launch_bounds__(256) void k_reslice(FrontRec* __restrict__ front, RayRec* __restrict__ lanes,
                                                 uint32_t npix, const uint2* __restrict__ sell,
                                                 const uint32_t* __restrict__ off_f, const uint32_t* __restrict__ len_f,
                                                 const uint32_t* __restrict__ off_w, const uint32_t* __restrict__ len_w,
                                                 uint32_t base_w, uint2* __restrict__ inc, uint8_t* __restrict__ any,
                                                 uint8_t* __restrict__ wany, unsigned long long* __restrict__ counts) {
  const uint32_t ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= npix) return;  // uniform per wavefront: npix is a multiple of 64
  const uint32_t t = ray >> 6, l = ray & 63u;
  const FrontRec f = front[ray];
  const uint32_t rnw = lanes[ray].nw;
  const uint2* __restrict__ src = sell + f.off;
  const uint32_t fo = off_f[t] + l, wo = base_w + off_w[t] + l;
  const uint2 gap = make_uint2(kGapVox, 0u);
  // a voxel this lane need not flag: the left or the upper ray has it too
  auto dup = [&](uint32_t v) {
    const uint32_t left = (uint32_t)__shfl_up((int)v, 1, 64), up = (uint32_t)__shfl_up((int)v, 8, 64);
    return ((l & 7u) != 0u && left == v) || (l >= 8u && up == v);
  };
  // rows in batches of kB: the batch's loads are in flight together (the
  // bounds lf, lw are uniform per wavefront, so every lane takes part in the
  // shuffles)
  constexpr uint32_t kB = 8;
  uint64_t nl = 0, nw = 0;
  const uint32_t lf = len_f[t], lw = len_w[t];
  for (uint32_t k0 = 0; k0 < lf; k0 += kB) {
    uint2 e[kB];
#pragma unroll
    for (uint32_t u = 0; u < kB; ++u) {
      e[u] = gap;
      if (k0 + u < f.nf) e[u] = src[(size_t)(k0 + u) * kSlice];
    }
#pragma unroll
    for (uint32_t u = 0; u < kB; ++u) {
      if (k0 + u < lf) {
        inc[fo + (size_t)(k0 + u) * kSlice] = e[u];
        const bool live = e[u].x != kGapVox;
        nl += (uint64_t)__popcll(__ballot(live));
        const bool d = dup(e[u].x);
        if (live && !d) any[e[u].x] = 1;
      }
    }
  }
  const uint2* __restrict__ srcw = src + (size_t)f.nf * kSlice;
  for (uint32_t k0 = 0; k0 < lw; k0 += kB) {
    uint2 e[kB];
#pragma unroll
    for (uint32_t u = 0; u < kB; ++u) {
      e[u] = gap;
      if (k0 + u < rnw) e[u] = srcw[(size_t)(k0 + u) * kSlice];
    }
#pragma unroll
    for (uint32_t u = 0; u < kB; ++u) {
      if (k0 + u < lw) {
        inc[wo + (size_t)(k0 + u) * kSlice] = e[u];
        const bool live = e[u].x != kGapVox;
        const uint64_t n = (uint64_t)__popcll(__ballot(live));
        nl += n;
        nw += n;
        const bool d = dup(e[u].x);
        if (live && !d) {
          any[e[u].x] = 1;
          wany[e[u].x] = 1;
        }
      }
    }
  }
  // the ray's records (gap records included) and whether it has any
  uint32_t nr = f.nf + rnw;
  const uint64_t nray = (uint64_t)__popcll(__ballot(nr != 0u));
  for (int o = 32; o > 0; o >>= 1) nr += (uint32_t)__shfl_xor((int)nr, o, 64);
  if (l == 0) {
    if (nl) atomicAdd(&counts[0], (unsigned long long)nl);
    if (nw) atomicAdd(&counts[1], (unsigned long long)nw);
    if (nr) atomicAdd(&counts[2], (unsigned long long)nr);
    if (nray) atomicAdd(&counts[3], (unsigned long long)nray);
  }
  front[ray].off = fo;
  lanes[ray].off_w = wo;
}

// Brick entries of the window records, one wavefront per slice, while the
// records still hold voxel ids (slot * 512 + voxel in the brick): record k
// of lane l enters a brick when it is a voxel record and the previous record
// of its ray is a gap record, lies in another slot, or does not exist.  The
// 64 lanes' flags of window row k form one 64-bit word.  tail[lane]: when
// the first window record continues the brick of the last front record, the
// number of front records since that brick's entry (at most 22, the longest
// crossing of a brick), else 0.
__global__ __launch_bounds__(256) void k_brick_starts(const FrontRec* __restrict__ front,
                                                      const RayRec* __restrict__ lanes,
                                                      const uint32_t* __restrict__ len_w, uint32_t nslices,
                                                      const uint2* __restrict__ inc, uint32_t base_w,
                                                      uint64_t* __restrict__ bstart, uint8_t* __restrict__ tail) {
  const uint32_t t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (t >= nslices) return;  // uniform per wavefront
  const uint32_t l = threadIdx.x & 63u, ray = t * kSlice + l;
  const FrontRec f = front[ray];
  const RayRec r = lanes[ray];
  const uint32_t sw = __builtin_amdgcn_readfirstlane(r.off_w - l);  // the slice's window region
  const uint32_t row0 = (sw - base_w) / kSlice;
  const uint32_t lw = len_w[t];
  const uint32_t last = f.nf ? inc[f.off + (size_t)(f.nf - 1) * kSlice].x : kGapVox;
  uint32_t prev = last;
  for (uint32_t k = 0; k < lw; ++k) {  // padding records are gap records
    const uint32_t cur = inc[sw + (size_t)k * kSlice + l].x;
    const bool bs = cur != kGapVox && (prev == kGapVox || (prev >> kBrickLog * 3) != (cur >> kBrickLog * 3));
    const unsigned long long m = __ballot(bs);
    if (l == 0) bstart[row0 + k] = m;
    prev = cur;
  }
  uint32_t tl = 0;
  if (f.nf && r.nw) {
    const uint32_t w0 = inc[sw + l].x;
    if (w0 != kGapVox && last != kGapVox && (w0 >> kBrickLog * 3) == (last >> kBrickLog * 3)) {
      uint32_t k = f.nf - 1;
      while (k > 0) {
        const uint32_t pv = inc[f.off + (size_t)(k - 1) * kSlice].x;
        if (pv == kGapVox || (pv >> kBrickLog * 3) != (w0 >> kBrickLog * 3)) break;
        --k;
      }
      tl = f.nf - k;
    }
  }
  tail[ray] = (uint8_t)tl;
}

// Front chunks: a slice with window records needs ceil(len_f / kChunk) of
// them, a slice without none.
// (cnt[nslices] = 0, so that an exclusive scan of nslices + 1 counts ends
// with the total)
__global__ void k_chunk_counts(const uint32_t* __restrict__ len_f, const uint32_t* __restrict__ len_w,
                               uint32_t nslices, uint32_t* __restrict__ cnt) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t > nslices) return;
  cnt[t] = (t < nslices && len_w[t]) ? (len_f[t] + kChunk - 1) / kChunk : 0u;
}

// Chunk descriptors {chunk, slice, first row, end row} and, per lane of every
// slice, its front span {first record, records to sum: nf, 0 without window}.
__global__ void k_chunk_desc(const uint32_t* __restrict__ chunk_off, const uint32_t* __restrict__ len_f,
                             uint32_t nslices, uint4* __restrict__ desc) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nslices) return;
  for (uint32_t w = chunk_off[t]; w < chunk_off[t + 1]; ++w) {
    const uint32_t k0 = (w - chunk_off[t]) * kChunk;
    desc[w] = make_uint4(w, t, k0, min(k0 + kChunk, len_f[t]));
  }
}
__global__ void k_front_span(const FrontRec* __restrict__ front, const RayRec* __restrict__ lanes, uint32_t n,
                             uint2* __restrict__ span) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < n) span[r] = make_uint2(front[r].off, lanes[r].nw ? front[r].nf : 0u);
}
__global__ void k_gather_desc(const uint4* __restrict__ desc, const uint32_t* __restrict__ idx, uint32_t n,
                              uint4* __restrict__ out) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = desc[idx[j]];
}

// ---------------------------------------------------------------- ray kernels
// Front sums.  Front records have psi(s) = 0, so all they contribute to a ray
// is the visibility sum Lambda = sum l log2 n (in units of ln 2; the reference's visibility
// includes the voxel itself, gvdbkernels.cu:781-783); their messages are
// the constant F_EPS * l, summed per pair once per traversal, and their
// length sum is geometric (RayRec::ks/kc).  Chunk c of a slice is records
// [c * kChunk, (c + 1) * kChunk) of its 64 rays, one wavefront at a time,
// so the work splits into near-equal pieces whatever the ray lengths; each
// lane writes its Kahan-compensated partial sum.  Rays without window
// records are skipped.
//
// While the log2 n of all pairs are recomputed (the first two iterations of
// a call) every chunk is summed in record order, and the records whose log2 n
// stops changing (bit clear in the chunk-lane mask fdmask) are also summed
// on their own into fstatic.  Afterwards only the chunks with a changing
// record are visited (dyn_desc), and each lane adds its changing records,
// copied in order into a compact block of the chunk (dfront, rows of 64
// lanes, {first row, rows} in dcr), to its kept static sum: static records
// are not read again.  The two orders of summation differ only by FP32
// rounding.
template <int MI, int U>
__global__ __launch_bounds__(256) void k_front_sums(const uint2* __restrict__ fspan, const uint4* __restrict__ desc,
                                                    uint32_t nchunks, const uint4* __restrict__ dyn_desc,
                                                    uint32_t ndyn_chunks, const uint2* __restrict__ inc,
                                                    const typename Msg<MI>::T* __restrict__ Xi,
                                                    float* __restrict__ partial,
                                                    const uint32_t* __restrict__ msg_all,
                                                    const uint32_t* __restrict__ fdmask, float* __restrict__ fstatic,
                                                    const uint2* __restrict__ dfront, const uint2* __restrict__ dcr) {
  // a wave per entry of the descriptor list (one wave-uniform load per
  // chunk), the waves past the list's end exit at once.  The grid covers
  // every chunk; the loop strides for a smaller one.
  const bool all = *msg_all != 0u;
  const uint32_t n = all ? nchunks : ndyn_chunks;
  const uint32_t nwaves = gridDim.x * (blockDim.x >> 6);
  const uint32_t l = threadIdx.x & 63u;
  for (uint32_t i = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)); i < n;
       i += nwaves) {
    if (all) {
      const uint4 d = desc[i];  // {chunk, slice, first row, end row}
      const uint2 fs = fspan[d.y * kSlice + l];
      const uint32_t k1 = min(d.z + kChunk, fs.y);
      const uint2* __restrict__ rf = inc + fs.x;
      const uint32_t m = fdmask[(size_t)d.x * kSlice + l];
      Kahan acc = {0.f, 0.f}, accs = {0.f, 0.f};
      for (uint32_t kb = d.z; kb < d.w; kb += U) {
        uint2 e[U];
        float x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t k = kb + u;
          e[u] = k < k1 ? rf[(size_t)k * kSlice] : make_uint2(kGapVox, 0u);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = e[u].x != kGapVox ? Msg<MI>::load(Xi, e[u].x) : 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t k = kb + u;
          if (k < k1) {
            const float v = __uint_as_float(e[u].y) * x[u];
            acc.add(v);
            if (((m >> (k - d.z)) & 1u) == 0u) accs.add(v);
          }
        }
      }
      partial[(size_t)d.x * kSlice + l] = acc.at(0.0f);
      fstatic[(size_t)d.x * kSlice + l] = accs.at(0.0f);
    } else {
      const uint32_t w = dyn_desc[i].x;  // the chunk
      const uint2 cr = dcr[i];           // {first row, rows} of its compact block
      const uint32_t cnt = (uint32_t)__popc(fdmask[(size_t)w * kSlice + l]);
      const uint2* __restrict__ rf = dfront + (size_t)cr.x * kSlice + l;
      Kahan acc = {fstatic[(size_t)w * kSlice + l], 0.f};
      for (uint32_t rb = 0; rb < cr.y; rb += U) {
        uint2 e[U];
        float x[U];
#pragma unroll
        for (int u = 0; u < U; ++u) e[u] = rb + u < cnt ? rf[(size_t)(rb + u) * kSlice] : make_uint2(kGapVox, 0u);
#pragma unroll
        for (int u = 0; u < U; ++u) x[u] = e[u].x != kGapVox ? Msg<MI>::load(Xi, e[u].x) : 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u)
          if (rb + u < cnt) acc.add(__uint_as_float(e[u].y) * x[u]);
      }
      partial[(size_t)w * kSlice + l] = acc.at(0.0f);
    }
  }
}

// Per front chunk (the layout of k_front_sums): a record changes when it
// belongs to a window pair or to a front-only pair at a voxel that some
// camera reaches with window records (pflag[q - nwp] != 0); the others'
// log2 n stop changing after the first two iterations of a call.  Per lane,
// dmask has bit k - first row set for each changing record k; per chunk,
// rows = the most changing records of a lane (the rows of its compact
// block), 0 when no record changes (the chunk's partial sums can be kept).
// The live records that do not change, which the front sums stop reading
// after the first two iterations of a call, are counted into nstatic.
__global__ __launch_bounds__(1024) void k_chunk_static(const uint2* __restrict__ fspan,
                                                      const uint4* __restrict__ desc, uint32_t nchunks,
                                                      const uint2* __restrict__ inc,
                                                      uint32_t nwp, const uint32_t* __restrict__ pflag,
                                                      uint32_t* __restrict__ dyn_flag, uint32_t* __restrict__ dmask,
                                                      uint32_t* __restrict__ rows,
                                                      unsigned long long* __restrict__ nstatic) {
  __shared__ uint32_t block_kept[16];
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t l = threadIdx.x & 63u;
  if (w >= nchunks) {  // uniform per wavefront; still takes part in the block's sum
    if (l == 0) block_kept[threadIdx.x >> 6] = 0u;
    __syncthreads();
    return;
  }
  const uint4 d = desc[w];
  const uint2 fs = fspan[d.y * kSlice + l];
  const uint32_t k0 = d.z, k1 = min(k0 + kChunk, fs.y);
  const uint2* __restrict__ rf = inc + fs.x;
  // all loads of the chunk first, then the flags of its front-only pairs
  uint32_t q[kChunk], pf[kChunk];
#pragma unroll
  for (uint32_t b = 0; b < kChunk; ++b) q[b] = k0 + b < k1 ? rf[(size_t)(k0 + b) * kSlice].x : kGapVox;
#pragma unroll
  for (uint32_t b = 0; b < kChunk; ++b) pf[b] = q[b] != kGapVox && q[b] >= nwp ? pflag[q[b] - nwp] : 0u;
  uint32_t m = 0, live = 0;
#pragma unroll
  for (uint32_t b = 0; b < kChunk; ++b) {
    if (q[b] == kGapVox) continue;
    ++live;
    if (q[b] < nwp || pf[b] != 0u) m |= 1u << b;
  }
  dmask[(size_t)w * kSlice + l] = m;
  uint32_t nr = (uint32_t)__popc(m);
  uint32_t kept = live - nr;  // this lane's records whose sums are kept
  for (int o = 32; o > 0; o >>= 1) {
    nr = max(nr, (uint32_t)__shfl_xor((int)nr, o, 64));
    kept += (uint32_t)__shfl_xor((int)kept, o, 64);
  }
  if (l == 0) {
    dyn_flag[w] = nr ? 1u : 0u;
    rows[w] = nr;
    block_kept[threadIdx.x >> 6] = kept;
  }
  // one atomic per block, not per chunk: they all go to one address
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t sum = 0;
    for (uint32_t i = 0; i < (blockDim.x >> 6); ++i) sum += block_kept[i];
    if (sum) atomicAdd(nstatic, (unsigned long long)sum);
  }
}

// rows of every changing chunk's compact block, in list order
__global__ void k_dyn_rows(const uint4* __restrict__ dyn_desc, uint32_t n, const uint32_t* __restrict__ rows,
                           uint32_t* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = rows[dyn_desc[i].x];
}

// The compact blocks: per entry of the changing-chunk list (one wave), each
// lane copies its changing records of the chunk in order to rows
// first row + 0, 1, ... of the block; dcr[i] = {first row, rows}.
__global__ __launch_bounds__(256) void k_fill_dfront(const uint2* __restrict__ fspan,
                                                     const uint4* __restrict__ dyn_desc, uint32_t n,
                                                     const uint2* __restrict__ inc,
                                                     const uint32_t* __restrict__ dmask,
                                                     const uint32_t* __restrict__ rows,
                                                     const uint32_t* __restrict__ row_off, uint2* __restrict__ dfront,
                                                     uint2* __restrict__ dcr) {
  const uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (i >= n) return;  // uniform per wavefront
  const uint32_t l = threadIdx.x & 63u;
  const uint4 d = dyn_desc[i];
  const uint2 fs = fspan[d.y * kSlice + l];
  const uint2* __restrict__ rf = inc + fs.x;
  const uint32_t m = dmask[(size_t)d.x * kSlice + l];
  const uint32_t r0 = row_off[i];
  uint2 e[kChunk];
#pragma unroll
  for (uint32_t b = 0; b < kChunk; ++b) e[b] = (m >> b) & 1u ? rf[(size_t)(d.z + b) * kSlice] : make_uint2(0u, 0u);
  uint32_t r = 0;
#pragma unroll
  for (uint32_t b = 0; b < kChunk; ++b)
    if ((m >> b) & 1u) dfront[(size_t)(r0 + r++) * kSlice + l] = e[b];
  if (l == 0) dcr[i] = make_uint2(r0, rows[d.x]);
}

// ---------------------------------------------------------------- buffer access
// Raw buffer loads and stores through a wave-uniform resource (base address,
// size in bytes): the hardware's range check makes an access at or past the
// size read 0 or be dropped, and constant offsets fold into the instruction.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format, no swizzle
__device__ __forceinline__ Rsrc make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kRsrcWord3);
}

template <int MT>
struct MsgB;
template <>
struct MsgB<kMsgF32> {
  static constexpr uint32_t kShift = 2;
  __device__ static float load(Rsrc r, uint32_t off) {
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
  }
  __device__ static void store(Rsrc r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)off, 0, 0);
  }
};
template <>
struct MsgB<kMsgF16> {
  static constexpr uint32_t kShift = 1;
  __device__ static float load(Rsrc r, uint32_t off) {
    return __half2float(__ushort_as_half((unsigned short)__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0)));
  }
  __device__ static void store(Rsrc r, uint32_t off, float v) {
    __builtin_amdgcn_raw_buffer_store_b16(__half_as_ushort(__float2half_rn(v)), r, (int)off, 0, 0);
  }
};
template <>
struct MsgB<kMsgBF16> {
  static constexpr uint32_t kShift = 1;
  __device__ static float load(Rsrc r, uint32_t off) {
    return __uint_as_float((uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0) << 16);
  }
  __device__ static void store(Rsrc r, uint32_t off, float v) {
    uint32_t u = __float_as_uint(v);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)(u >> 16), r, (int)off, 0, 0);
  }
};

// Window of every ray: one thread per ray, the 64 rays of a slice in one
// wavefront; record k of all lanes of a slice is contiguous.  Lambda starts
// from the ray's front partial sums, the travelled length from the front's
// Kahan state.  Pass 1: vpsi_k = exp(Lambda_k) psi(s_k), tau_k = vpsi_k
// (1 - n_k), T = sum tau (+ V_n when the walk ran out and d lies beyond,
// gvdbkernels.cu:1196-1208).  Pass 2 restarts at the window with the saved
// sums and writes p = mu1 / (mu1 + mu0 + 1e-7), mu1 = P + vpsi,
// mu0 = P + R / n (gvdbkernels.cu:936-951) times l into the record's own
// slot of the window region (coalesced; the segment sums gather them in
// voxel order).  The file is built without FP contraction, so pass 2
// recomputes pass 1's terms bit for bit.
//
// Records are processed in batches of UW, software-pipelined: while a batch
// is computed, the message gathers of the next batch and the record loads of
// the one after are in flight (two batches of lengths and messages and one
// of records in registers, used in turn, so nothing is copied).  Every
// access goes through a buffer resource: the records and the message slots
// of the slice's window region, and the pair messages.  Loads past the
// region read {0, 0} (pair 0, length 0); a gap record's pair index is past
// the message array, so its message reads 0, and since every pair's
// log2 n <= log2(1 - F_EPS) < 0, a record is live exactly when its message is
// nonzero.  All lanes run to the slice's longest ray (uniform loop control);
// slots past a lane's last record are computed too: they have length 0, so
// the visibility and length sums add exact zeros, their taus are left out
// of T and P by a select, their stores land in the slice's padding or past
// the region (dropped), and no kernel reads those slots.
// UW records of a lane from row k0 on (record k at byte lane_off + k * 64 * 8)
template <int UW>
__device__ __forceinline__ void win_load(Rsrc rr, uint32_t lane_off, uint32_t k0, uint2 (&e)[UW]) {
#pragma unroll
  for (int u = 0; u < UW; ++u) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(rr, (int)(lane_off + (k0 + (uint32_t)u) * (kSlice * 8u)), 0, 0);
    e[u] = make_uint2((uint32_t)v[0], (uint32_t)v[1]);
  }
}
// gathers of the batch's messages; the lengths move to their own registers
template <int MI, int UW>
__device__ __forceinline__ void win_issue(Rsrc rx, const uint2 (&e)[UW], float (&ln)[UW], float (&x)[UW]) {
#pragma unroll
  for (int u = 0; u < UW; ++u) {
    x[u] = MsgB<MI>::load(rx, e[u].x << MsgB<MI>::kShift);
    ln[u] = __uint_as_float(e[u].y);
  }
}

struct WinRay {
  float z0, inv, d;
  Poly sig3, bia;  // 3 sigma(z) (the likelihood width) and bias(z)
};

// depth at the exit of a record: z0 + (travelled chord length) * inv
__device__ __forceinline__ float win_depth(const WinRay& r, const Kahan& kS, float ln) {
  return fmaf(kS.at(ln), r.inv, r.z0);
}

// terms of one record shared by both passes: a = l log2 n, vpsi, n
__device__ __forceinline__ void win_terms(const WinRay& r, const Kahan& kL, const Kahan& kS, float ln, float lnn,
                                          float& a, float& vpsi, float& nn) {
  a = ln * lnn;
  const float s = win_depth(r, kS, ln);
  const float ps = depth_psi(s, r.d, r.sig3, r.bia);
  // a gap or padding record (log2 n = 0) gets a vpsi but n = 1 exactly, so
  // its tau vpsi (1 - n) is exactly 0; its own message is never read
  vpsi = __builtin_amdgcn_exp2f(kL.at(a)) * ps;
  nn = __builtin_amdgcn_exp2f(lnn);
}

template <int UW>
__device__ __forceinline__ void win_pass1(const WinRay& r, uint32_t k0, uint32_t nw, const float (&ln)[UW],
                                          const float (&x)[UW], Kahan& kL, Kahan& kS, float& T) {
#pragma unroll
  for (int u = 0; u < UW; ++u) {
    const bool valid = k0 + u < nw;
    float a, vpsi, nn;
    win_terms(r, kL, kS, ln[u], x[u], a, vpsi, nn);
    T += valid ? vpsi * (1.0f - nn) : 0.0f;
    kL.add(a);  // a = 0 and l = 0 past the lane's last record: no select needed
    kS.add(ln[u]);
  }
}

// Pass 1 of the reference's pre pass with its flush-to-zero arithmetic
// (MRF_VIS_BRICK_LOCAL): a tau is non-zero there only when a = 0.398942
// e^{-u^2/2}, psi = a / (3 sigma), b = V (1 - n) and b psi are all at least
// the smallest normal float (gvdbkernels.cu:57-72, 792-795).  Until the
// first such tau the visibility sum restarts at every brick entry (bit l of
// the row's word in bsw; the restore at gvdbkernels.cu:708-715 tests the
// accumulated tau), and only such taus enter T.  The terms are those of
// win_terms, so a ray whose restarts change nothing gets the T of pass 1.
constexpr float kFtzMin = 0x1p-126f;

template <int UW>
__device__ __forceinline__ void win_pass1_bl(const WinRay& r, uint32_t k0, uint32_t nw, const float (&ln)[UW],
                                             const float (&x)[UW], const uint64_t* __restrict__ bsw, uint32_t l,
                                             Kahan& kL, Kahan& kS, float& T, bool& nz) {
#pragma unroll
  for (int u = 0; u < UW; ++u) {
    const bool valid = k0 + u < nw;
    const bool restart = valid && !nz && ((bsw[k0 + u] >> l) & 1ull) != 0ull;
    kL.s = restart ? 0.0f : kL.s;
    kL.c = restart ? 0.0f : kL.c;
    const float lnn = x[u];
    const float a = ln[u] * lnn;
    const float s = win_depth(r, kS, ln[u]);
    const float diff = fabsf(r.d - (s + r.bia(s)));
    const float inv = __builtin_amdgcn_rcpf(r.sig3(s));
    const float uu = diff * inv;
    const float ap = gauss_a(uu);
    const float ps = ap * inv;  // depth_psi's product order
    const float V = __builtin_amdgcn_exp2f(kL.at(a));
    const float nn = __builtin_amdgcn_exp2f(lnn);
    const float b = V * (1.0f - nn);
    // every factor at least 2^-126 (log2 n = 0, a gap or padding record, gives b = 0)
    const bool counted = valid && fminf(fminf(ap, ps), fminf(b, b * ps)) >= kFtzMin;
    T += counted ? (V * ps) * (1.0f - nn) : 0.0f;
    nz = nz || counted;
    kL.add(a);  // a = 0 and l = 0 past the lane's last record: no select needed
    kS.add(ln[u]);
  }
}

template <int UW, bool BL>
__device__ __forceinline__ void win_pass1_any(const WinRay& r, uint32_t k0, uint32_t nw, const float (&ln)[UW],
                                              const float (&x)[UW], const uint64_t* __restrict__ bsw, uint32_t l,
                                              Kahan& kL, Kahan& kS, float& T, bool& nz) {
  if constexpr (BL)
    win_pass1_bl<UW>(r, k0, nw, ln, x, bsw, l, kL, kS, T, nz);
  else
    win_pass1<UW>(r, k0, nw, ln, x, kL, kS, T);
}

template <int MO, int UW>
__device__ __forceinline__ void win_pass2(const WinRay& r, uint32_t k0, uint32_t nw, const float (&ln)[UW],
                                          const float (&x)[UW], float T, Rsrc ro, uint32_t lane_off, Kahan& kL,
                                          Kahan& kS, float& P) {
  typedef typename Msg<MO>::T TO;
#pragma unroll
  for (int u = 0; u < UW; ++u) {
    const bool valid = k0 + u < nw;
    float a, vpsi, nn;
    win_terms(r, kL, kS, ln[u], x[u], a, vpsi, nn);
    const float tau = vpsi * (1.0f - nn);
    const float R = (T - P) - tau;
    const float mu1 = P + vpsi;
    const float mu0 = P + R * __builtin_amdgcn_rcpf(nn);
    float p = mu1 * __builtin_amdgcn_rcpf(mu1 + mu0 + kMsgNormEps);
    p = fminf(fmaxf(p, kFEps), 1.0f - kFEps);
    MsgB<MO>::store(ro, lane_off + (k0 + (uint32_t)u) * (kSlice * (uint32_t)sizeof(TO)), p * ln[u] * Msg<MO>::kOut);
    kL.add(a);  // a = 0 and l = 0 past the lane's last record: no select needed
    kS.add(ln[u]);
    P += valid ? tau : 0.0f;
  }
}

template <int MI, int MO, int UW, bool BL>
__global__ __launch_bounds__(64) void k_ray_window(const RayRec* __restrict__ lanes, const uint32_t* __restrict__ len_w,
                                                    uint32_t nslices, const uint32_t* __restrict__ chunk_off,
                                                    const float* __restrict__ partial, const uint2* __restrict__ inc,
                                                    uint32_t base_w,
                                                    const typename Msg<MI>::T* __restrict__ Xi, uint32_t npairs,
                                                    typename Msg<MO>::T* __restrict__ Xo,
                                                    const float* __restrict__ sigt, const float* __restrict__ biat,
                                                    int lookup_n, const FrontRec* __restrict__ front,
                                                    const uint64_t* __restrict__ bstart,
                                                    const uint8_t* __restrict__ tail) {
  typedef typename Msg<MI>::T TI;
  typedef typename Msg<MO>::T TO;
  // one wave (one slice) per block, so that a finished wave frees its slot at
  // once; XCD-aware order: consecutive blocks (neighbouring pixel tiles) share
  // an XCD's L2
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t ray = lb * 64u + threadIdx.x;  // slice * 64 + lane
  const uint32_t t = ray >> 6, l = ray & 63u;
  if (t >= nslices) return;
  const uint32_t lw = len_w[t];
  if (lw == 0) return;
  const RayRec rec = lanes[ray];
  const uint32_t nw = rec.nw;
  if (nw == 0) return;  // no lane-crossing operations below

  // the front's visibility sum: the partial sums of the slice's front chunks, loaded 8 at a time
  Kahan kL = {0.f, 0.f};
  const uint32_t c1 = chunk_off[t + 1];
  for (uint32_t c0 = chunk_off[t]; c0 < c1; c0 += 8) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = c0 + i < c1 ? partial[(size_t)(c0 + i) * kSlice + l] : 0.0f;
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (c0 + i < c1) kL.add(v[i]);
  }
  Kahan kS = {rec.ks, rec.kc};
  WinRay r;
  load_cell(sigt, biat, lookup_n, rec.cell_flags & kCellMask, r.sig3, r.bia);
  r.sig3.c0 *= 3.0f;
  r.sig3.c1 *= 3.0f;
  r.sig3.c2 *= 3.0f;
  r.z0 = rec.z0;
  r.inv = rec.inv_norm;
  r.d = rec.depth;
  const Kahan L0 = kL, S0 = kS;
  // the slice's window region (record k of lane l at sw + k * 64 + l), wave-uniform
  const uint32_t sw = __builtin_amdgcn_readfirstlane(rec.off_w - l);
  const uint32_t lws = __builtin_amdgcn_readfirstlane(lw);
  const Rsrc rr = make_rsrc(inc + sw, lws * kSlice * 8u);
  const Rsrc rx = make_rsrc(Xi, npairs * (uint32_t)sizeof(TI));
  const Rsrc ro = make_rsrc(Xo + (sw - base_w), lws * kSlice * (uint32_t)sizeof(TO));
  const uint32_t lr = l * 8u, lo = l * (uint32_t)sizeof(TO);

  float lA[UW], xA[UW], lB[UW], xB[UW];
  uint2 e[UW];
  // pass 1
  float T = 0.0f;  // sum of the taus: positive terms, a plain sum
  bool nz = false;
  const uint64_t* bsw = nullptr;
  if constexpr (BL) {
    // the pre pass enters the window with the visibility sum of the front
    // records in the first window record's brick (every front tau is 0);
    // bit l of word k of bsw: window record k of this lane enters a brick
    // (the slice's rows of the word array, padded past the last slice)
    // (at most 22 records, the longest crossing of a brick; 8 record loads,
    // then their 8 gathers, in flight at a time, summed in record order)
    const FrontRec fr = front[ray];
    const uint32_t tl = tail[ray];
    const uint2* __restrict__ tr = inc + fr.off + (size_t)(fr.nf - tl) * kSlice;
    Kahan kp = {0.f, 0.f};
    for (uint32_t b = 0; b < tl; b += 8) {
      uint2 et[8];
      float tx[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) et[i] = b + i < tl ? tr[(size_t)(b + i) * kSlice] : make_uint2(0u, 0u);
#pragma unroll
      for (int i = 0; i < 8; ++i) tx[i] = b + i < tl ? Msg<MI>::load(Xi, et[i].x) : 0.0f;
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (b + i < tl) kp.add(__uint_as_float(et[i].y) * tx[i]);
    }
    kL = kp;
    bsw = bstart + (sw - base_w) / kSlice;
  }
  win_load<UW>(rr, lr, 0, e);
  win_issue<MI, UW>(rx, e, lA, xA);
  win_load<UW>(rr, lr, UW, e);
  for (uint32_t k0 = 0;; k0 += 2 * UW) {  // bounded by the slice's longest ray: uniform control flow
    win_issue<MI, UW>(rx, e, lB, xB);
    win_load<UW>(rr, lr, k0 + 2 * UW, e);
    win_pass1_any<UW, BL>(r, k0, nw, lA, xA, bsw, l, kL, kS, T, nz);
    if (k0 + UW >= lws) break;
    win_issue<MI, UW>(rx, e, lA, xA);
    win_load<UW>(rr, lr, k0 + 3 * UW, e);
    win_pass1_any<UW, BL>(r, k0 + UW, nw, lB, xB, bsw, l, kL, kS, T, nz);
    if (k0 + 2 * UW >= lws) break;
  }
  if (rec.cell_flags & kInfTerm) T += __builtin_amdgcn_exp2f(kL.at(0.0f));  // decided at traversal

  // pass 2
  kL = L0;
  kS = S0;
  float P = 0.0f;  // prefix sum of the taus
  win_load<UW>(rr, lr, 0, e);
  win_issue<MI, UW>(rx, e, lA, xA);
  win_load<UW>(rr, lr, UW, e);
  for (uint32_t k0 = 0;; k0 += 2 * UW) {
    win_issue<MI, UW>(rx, e, lB, xB);
    win_load<UW>(rr, lr, k0 + 2 * UW, e);
    win_pass2<MO, UW>(r, k0, nw, lA, xA, T, ro, lo, kL, kS, P);
    if (k0 + UW >= lws) break;
    win_issue<MI, UW>(rx, e, lA, xA);
    win_load<UW>(rr, lr, k0 + 3 * UW, e);
    win_pass2<MO, UW>(r, k0 + UW, nw, lB, xB, T, ro, lo, kL, kS, P);
    if (k0 + 2 * UW >= lws) break;
  }
}

// Segment sums S_q = sum of p*l over pair q's positions, in a fixed order.
// One wavefront per tile of kTile = 512 positions; lane l owns the 8
// consecutive positions j0 + 8l .. j0 + 8l + 7 (two 128-bit loads for FP32).
// Each lane sums its runs serially and writes the segments that start and
// end inside it; the runs that cross lanes are joined by one segmented
// inclusive scan over the 64 lanes (DPP).  Of the segments that cross tiles,
// the part that started in an earlier tile goes to pf, the part that
// continues into the next tile to pl, and a tile lying wholly inside one
// segment to mid; k_seg_update joins them.

template <int CTRL, int ROWS>
__device__ __forceinline__ int dppi(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, ROWS, 0xF, true);
}

// (x, f) <- (x_left, f_left) (+) (x, f): add the left value unless a segment
// starts at or before this lane.
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_step(float& x, int& f) {
  const float y = dpp<CTRL, ROWS>(x);
  const int g = dppi<CTRL, ROWS>(f);
  x = f ? x : x + y;
  f |= g;
}

__device__ __forceinline__ void wave_seg_scan(float& x, int& f) {
  seg_step<0x111, 0xF>(x, f);
  seg_step<0x112, 0xF>(x, f);
  seg_step<0x114, 0xF>(x, f);
  seg_step<0x118, 0xF>(x, f);
  seg_step<0x142, 0xA>(x, f);
  seg_step<0x143, 0xC>(x, f);
}

__device__ __forceinline__ int wave_scan_i(int x) {
  x += dppi<0x111, 0xF>(x);
  x += dppi<0x112, 0xF>(x);
  x += dppi<0x114, 0xF>(x);
  x += dppi<0x118, 0xF>(x);
  x += dppi<0x142, 0xA>(x);
  x += dppi<0x143, 0xC>(x);
  return x;
}

// 8 consecutive messages from p0 (a multiple of 8), zero past j1; with an
// index array, message j is X[idx[j]]
template <int MT>
__device__ __forceinline__ void load8(const typename Msg<MT>::T* __restrict__ X, const uint32_t* __restrict__ idx,
                                      uint32_t p0, uint32_t j1, float (&x)[8]) {
  typedef typename Msg<MT>::T T;
  if (idx) {
    if (p0 + 8 <= j1) {
      const uint4 a = *reinterpret_cast<const uint4*>(idx + p0);
      const uint4 b = *reinterpret_cast<const uint4*>(idx + p0 + 4);
      const uint32_t ix[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
      for (int i = 0; i < 8; ++i) x[i] = Msg<MT>::load(X, ix[i]);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) x[i] = p0 + i < j1 ? Msg<MT>::load(X, idx[p0 + i]) : 0.0f;
    }
    return;
  }
  if (p0 + 8 <= j1) {
    constexpr int kVec = (int)(8 * sizeof(T) / sizeof(uint4));
    union {
      uint4 v[kVec];
      T t[8];
    } u;
    const uint4* src = reinterpret_cast<const uint4*>(X + p0);
#pragma unroll
    for (int k = 0; k < kVec; ++k) u.v[k] = src[k];
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = Msg<MT>::load(u.t, (uint32_t)i);
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) x[i] = p0 + i < j1 ? Msg<MT>::load(X, p0 + i) : 0.0f;
  }
}

template <int MT>
__global__ __launch_bounds__(256) void k_seg_sum(const uint64_t* __restrict__ heads, const uint32_t* __restrict__ tile_h0,
                                                 uint32_t nlive, const typename Msg<MT>::T* __restrict__ X,
                                                 const uint32_t* __restrict__ idx,
                                                 float* __restrict__ S,
                                                 float* __restrict__ pf_v,
                                                 float* __restrict__ pl_v, float* __restrict__ mid) {
  const int lane = threadIdx.x & 63;
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t j0 = w * kTile;
  if (j0 >= nlive) return;  // uniform per wavefront
  const uint32_t j1 = min(j0 + kTile, nlive);
  const uint32_t p0 = j0 + 8u * (uint32_t)lane;
  // head bits of this lane's 8 positions (zero past nlive)
  const uint32_t hb = p0 < j1 ? (uint32_t)(heads[p0 >> 6] >> (p0 & 63)) & 0xFFu : 0u;
  float x[8];
  load8<MT>(X, idx, p0, j1, x);
  const int nh = __popc(hb);
  const int incl = wave_scan_i(nh);
  const uint32_t before = tile_h0[w] + (uint32_t)(incl - nh);  // heads before p0
  // serial runs; segments that start and end in the lane are complete
  float run = 0.0f, headpart = 0.0f;
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    if ((hb >> i) & 1u) {
      if (cnt)
        S[before + (uint32_t)cnt - 1u] = run;
      else
        headpart = run;
      ++cnt;
      run = 0.0f;
    }
    run += x[i];
  }
  // join the runs across lanes: U = open run at the end of this lane
  float U = run;
  int F = cnt > 0;
  wave_seg_scan(U, F);
  float in = __shfl_up(U, 1);
  int fin = __shfl_up(F, 1);
  if (lane == 0) {
    in = 0.0f;
    fin = 0;
  }
  const bool open_first = ((heads[j0 >> 6] >> (j0 & 63)) & 1ull) == 0ull;  // tile starts inside a segment
  // the segment that ends just before this lane's first head
  bool pf_here = false;
  float pfv = 0.0f;
  if (cnt) {
    const float v = in + headpart;
    if (fin)
      S[before - 1u] = v;
    else if (open_first) {
      pf_here = true;  // first head of the tile: the segment began in an earlier tile
      pfv = v;
    }
  }
  // the run open at the end of the tile
  const float Ut = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(U), 63));
  const int Ft = __builtin_amdgcn_readlane(F, 63);
  const int ntile = __builtin_amdgcn_readlane(incl, 63);
  const uint32_t qt = tile_h0[w] + (uint32_t)ntile - 1u;  // pair of the tile's last head
  const bool ends_here = j1 >= nlive || (heads[j1 >> 6] & 1ull) != 0ull;
  const unsigned long long pfmask = __ballot(pf_here);
  const float pfv0 =
      pfmask ? __int_as_float(__builtin_amdgcn_readlane(__float_as_int(pfv), __builtin_ctzll(pfmask))) : 0.0f;
  if (lane == 0) {
    float pv = pfmask ? pfv0 : 0.0f, lv = 0.0f, mv = 0.0f;
    if (ends_here) {
      if (Ft)
        S[qt] = Ut;
      else
        pv = Ut;  // no head in the tile: the whole tile ends an earlier segment
    } else if (Ft) {
      lv = Ut;  // the tile's last segment continues into the next tile
    } else {
      mv = Ut;  // the tile lies inside one segment
    }
    pf_v[w] = pv;
    pl_v[w] = lv;
    mid[w] = mv;
  }
}


// Camera message of each (camera, voxel) pair from its message sum:
// a = clamp(S/C), m = ln a - ln(1-a), B += m - M_c, M_c = m, for C > 0.1f
// (compute_gvdb_message_to_factors, gvdbkernels.cu:408, 457-479).
// S = (sum of the pair's constant front messages) + (its window segment
// sum); window pairs are numbered first.  A window segment inside one tile
// of the segment kernel has its sum in S; one that crosses tiles is joined
// from its parts: the head tile's pl, the tiles inside it (mid), the tail
// tile's pf.

// The update of pair q from its loaded operands: C, the front sum sf, the
// window sum Sw (0 for a front-only pair), the voxel v, its belief Bv and
// the pair's previous message mo.  Every operand is loaded before this test
// (the belief right after the voxel), so that a pair costs two memory round
// trips, not four.
__device__ __forceinline__ void apply_pair(uint32_t q, float C, float sf, float Sw, uint32_t v, float Bv, float mo,
                                           float* __restrict__ B, float* __restrict__ M,
                                           uint8_t* __restrict__ upd, bool mark) {
  if (!(C > kCumLenMin)) return;
  float a = (sf + Sw) / C;
  a = fminf(fmaxf(a, kFEps), 1.0f - kFEps);
  const float m = __logf(a) - __logf(1.0f - a);
  M[q] = m;
  B[v] = (Bv + m) - mo;
  if (mark) upd[v] = 1;
}

// The window pairs [0, nwp), one per thread, and while *all_pairs (the
// first iteration of a call) the front-only pairs [nwp, npairs), strided
// over the grid: a front-only pair's m is the same every iteration, and once
// it is in M the update B = (B + m) - M would add nothing but FP32 rounding.
// The update flags (a voxel's belief has been set by a message) are written
// in that first iteration only: every later iteration updates the same
// voxels.  A camera has one pair per voxel, so no two threads share a B[v].
__global__ __launch_bounds__(256) void k_seg_update(const uint32_t* __restrict__ pair_vox,
                                                    const float* __restrict__ pair_c,
                                                    const float* __restrict__ pair_sf, uint32_t nwp, uint32_t npairs,
                                                    const uint32_t* __restrict__ wseg_off, const float* __restrict__ S,
                                                    const float* __restrict__ pl_v, const float* __restrict__ pf_v,
                                                    const float* __restrict__ mid, float out_scale,
                                                    float* __restrict__ B, float* __restrict__ M,
                                                    uint8_t* __restrict__ upd, const uint32_t* __restrict__ all_pairs) {
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  const bool first = *all_pairs != 0u;
  for (uint32_t q = tid; q < nwp; q += stride) {
    const uint32_t o0 = wseg_off[q], o1 = wseg_off[q + 1];
    const float Sq = S[q];  // stale for a segment that crosses tiles: not used then
    const float C = pair_c[q], sf = pair_sf[q], mo = M[q];
    const uint32_t v = pair_vox[q];
    const float Bv = B[v];
    const uint32_t w = o0 / kTile, wt = (o1 - 1u) / kTile;
    float Sw = Sq;
    if (w != wt) {
      float acc = 0.0f;
      for (uint32_t t = w + 1; t < wt; ++t) acc += mid[t];
      Sw = (pl_v[w] + acc) + pf_v[wt];
    }
    apply_pair(q, C, sf, Sw / out_scale, v, Bv, mo, B, M, upd, first);
  }
  if (first)
    for (uint32_t q = nwp + tid; q < npairs; q += stride) {
      const float C = pair_c[q], sf = pair_sf[q], mo = M[q];
      const uint32_t v = pair_vox[q];
      apply_pair(q, C, sf, 0.0f, v, B[v], mo, B, M, upd, true);
    }
}

// Beliefs of an external map (distinct voxels): B = the given log-odds,
// marked as updated so that the marginal reads back.
__global__ void k_set_voxels(const uint32_t* __restrict__ vox, const float* __restrict__ logodds, uint32_t n,
                             float* __restrict__ B, uint8_t* __restrict__ upd) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  B[vox[i]] = logodds[i];
  upd[vox[i]] = 1;
}

// Summed chord length per voxel over one camera's pairs (a camera's pairs
// are distinct voxels, so cameras launched one after another need no atomics).
__global__ void k_cum_add(const uint32_t* __restrict__ pair_vox, const float* __restrict__ pair_c, uint32_t npairs,
                          float* __restrict__ cum) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < npairs) cum[pair_vox[q]] += pair_c[q];
}

// Voxels that no remaining ray reaches go back to the prior and read 0
// (reset_empty_voxels, gvdbkernels.cu:498-517) ...
__global__ void k_reset_empty(const float* __restrict__ cum, size_t n, float prior_logodds, float* __restrict__ B,
                              uint8_t* __restrict__ upd) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n || !(cum[v] < 0.05f)) return;
  B[v] = prior_logodds;
  upd[v] = 0;
}

// ... and lose the moved camera's message.
__global__ void k_reset_pair_msgs(const uint32_t* __restrict__ pair_vox, uint32_t npairs, const float* __restrict__ cum,
                                  float* __restrict__ M) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < npairs && cum[pair_vox[q]] < 0.05f) M[q] = 0.0f;
}

// Camera messages across a re-traversal: old pairs -> voxels -> new pairs.
__global__ void k_scatter_pairs(const uint32_t* __restrict__ pair_vox, const float* __restrict__ M, uint32_t npairs,
                                float* __restrict__ dense) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < npairs) dense[pair_vox[q]] = M[q];
}

__global__ void k_gather_pairs(const uint32_t* __restrict__ pair_vox, const float* __restrict__ dense, uint32_t npairs,
                               float* __restrict__ M) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < npairs) M[q] = dense[pair_vox[q]];
}

// marginal occupancy 1 - clamp(1/(1+e^B)) (gvdbkernels.cu:469-476)
__device__ __forceinline__ float marginal(float b) {
  float neg = 1.0f / (1.0f + expf(b));
  neg = fminf(fmaxf(neg, kFEps), 1.0f - kFEps);
  return 1.0f - neg;
}

__global__ void k_marginals(const float* __restrict__ B, float* __restrict__ A, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) A[i] = marginal(B[i]);
}

// max |marginal(B) - Aprev| as an IEEE double reduced through the ordered bit
// pattern of non-negative doubles; Aprev is refreshed in the same sweep.
__global__ __launch_bounds__(256) void k_residual(const float* __restrict__ B, float* __restrict__ A, size_t n,
                                                   unsigned long long* __restrict__ d_max) {
  __shared__ double sh[256];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float a = marginal(B[i]);
    const double dlt = fabs((double)a - (double)A[i]);
    A[i] = a;
    m = dlt > m ? dlt : m;
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(d_max, (unsigned long long)__double_as_longlong(sh[0]));
}

// max |alpha(v) - ref(v)| with alpha the readback marginal (A, or 0 where
// never updated); A holds the marginals of the current beliefs.
__global__ __launch_bounds__(256) void k_ref_error(const float* __restrict__ A, const uint8_t* __restrict__ upd,
                                                   const float* __restrict__ ref, size_t n,
                                                   unsigned long long* __restrict__ d_max) {
  __shared__ double sh[256];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float a = upd[i] ? A[i] : 0.0f;
    const double dlt = fabs((double)a - (double)ref[i]);
    m = dlt > m ? dlt : m;
  }
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(d_max, (unsigned long long)__double_as_longlong(sh[0]));
}

__global__ void k_alphas(const float* __restrict__ B, const uint8_t* __restrict__ upd, float* __restrict__ out,
                         size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = upd[i] ? marginal(B[i]) : 0.0f;
}

__global__ void k_fill_f32(float* p, float v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_countdown(uint32_t* p) {
  if (threadIdx.x == 0 && *p) *p -= 1u;
}

// window-touched voxels: W[v] = 1 for the voxel of every window pair
__global__ void k_mark_window_voxels(const uint32_t* __restrict__ pair_vox, uint32_t nwp, uint32_t* __restrict__ W) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nwp) W[pair_vox[q]] = 1u;
}

__global__ void k_dyn_flags(const uint32_t* __restrict__ pair_vox, uint32_t n, const uint32_t* __restrict__ W,
                            uint32_t* __restrict__ flags) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) flags[k] = W[pair_vox[k]];
}

__global__ void k_dyn_compact(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, uint32_t n,
                              uint32_t q0, uint32_t* __restrict__ dyn) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && flags[k]) dyn[pos[k]] = q0 + k;
}

__global__ void k_dyn_pairs(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, uint32_t n,
                            uint32_t q0, const uint32_t* __restrict__ pair_vox, uint2* __restrict__ dyn) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n && flags[k]) dyn[pos[k]] = make_uint2(q0 + k, pair_vox[q0 + k]);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_u8(uint8_t* p, uint8_t v, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_zero_u32(uint32_t* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

inline unsigned grid_for(size_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }
// grid of the per-pair kernels: a thread per window pair, at most 4096
// blocks strided over the front-only pairs
inline unsigned pair_grid(const CamPass& a) {
  return std::max(1u, std::max(grid_for(a.nwp, 256), std::min<unsigned>(grid_for(a.npairs - a.nwp, 256), 4096u)));
}

}  // namespace

template <int MT>
void launch_pair_msgs_t(const CamPass& a, hipStream_t s) {
  typedef typename Msg<MT>::T T;
  hipLaunchKernelGGL(k_pair_msg<MT>, dim3(pair_grid(a)), dim3(256), 0, s, a.pair_vox, a.nwp, a.npairs, a.dyn, a.ndyn,
                     a.B, a.M, (T*)a.apair, a.msg_all);
}

void launch_pair_msgs(const CamPass& a, hipStream_t s) {
  if (a.npairs == 0) return;
  switch (msg_in_type(a.msg_type)) {
    case kMsgF16:
      launch_pair_msgs_t<kMsgF16>(a, s);
      break;
    case kMsgBF16:
      launch_pair_msgs_t<kMsgBF16>(a, s);
      break;
    default:
      launch_pair_msgs_t<kMsgF32>(a, s);
  }
}

template <int MI, int MO, int UF, int UW>
void launch_ray_t(const CamPass& a, hipStream_t s, int part) {
  const typename Msg<MI>::T* Xi = (const typename Msg<MI>::T*)a.apair;
  if (a.nchunks && part != 2)
    hipLaunchKernelGGL((k_front_sums<MI, UF>), dim3(grid_for(a.nchunks, 4)), dim3(256), 0, s,
                       a.fspan, a.chunk_desc, a.nchunks, a.dyn_desc, a.ndyn_chunks, a.inc, Xi, a.partial, a.msg_all,
                       a.fdmask, a.fstatic, a.dfront, a.dcr);
  if (part == 1) return;
  const uint32_t blocks = (a.nslices * kSlice + 63) / 64;
  const unsigned nb = (blocks + 7) / 8 * 8;  // multiple of the 8 XCDs for the block remapping
  if (a.brick_local)
    hipLaunchKernelGGL((k_ray_window<MI, MO, UW, true>), dim3(nb), dim3(64), 0, s, a.lanes, a.len_w, a.nslices,
                       a.chunk_off, a.partial, a.inc, a.base_w, Xi, a.npairs, (typename Msg<MO>::T*)a.Xo, a.sig, a.bia,
                       a.lookup_n, a.front, a.bstart, a.tail);
  else
    hipLaunchKernelGGL((k_ray_window<MI, MO, UW, false>), dim3(nb), dim3(64), 0, s, a.lanes, a.len_w, a.nslices,
                       a.chunk_off, a.partial, a.inc, a.base_w, Xi, a.npairs, (typename Msg<MO>::T*)a.Xo, a.sig, a.bia,
                       a.lookup_n, a.front, a.bstart, a.tail);
}

template <int MI, int MO>
void launch_ray_sched(const CamPass& a, hipStream_t s, int part) {
  switch (a.ray_sched) {
    case kSchedF8W8: launch_ray_t<MI, MO, 8, 8>(a, s, part); break;
    case kSchedF16W4: launch_ray_t<MI, MO, 16, 4>(a, s, part); break;
    case kSchedF16W8: launch_ray_t<MI, MO, 16, 8>(a, s, part); break;
    case kSchedF4W4: launch_ray_t<MI, MO, 4, 4>(a, s, part); break;
    case kSchedF16W16: launch_ray_t<MI, MO, 16, 16>(a, s, part); break;
    default: launch_ray_t<MI, MO, 8, 4>(a, s, part);
  }
}

void launch_ray_messages(const CamPass& a, hipStream_t s, int part) {
  if (a.nslices == 0 || a.nwin == 0) return;
  switch (a.msg_type) {
    case kMsgF16: launch_ray_sched<kMsgF16, kMsgF16>(a, s, part); break;
    case kMsgBF16: launch_ray_sched<kMsgBF16, kMsgBF16>(a, s, part); break;
    case kMsgMixedF16: launch_ray_sched<kMsgF32, kMsgF16>(a, s, part); break;
    case kMsgMixedBF16: launch_ray_sched<kMsgF32, kMsgBF16>(a, s, part); break;
    default: launch_ray_sched<kMsgF32, kMsgF32>(a, s, part);
  }
}


void launch_epilogue(const CamPass& a, hipStream_t s, int part) {
  if (a.npairs == 0) return;
  const int mo = msg_out_type(a.msg_type);
  const float out_scale = mo == kMsgF16 ? Msg<kMsgF16>::kOut : (mo == kMsgBF16 ? Msg<kMsgBF16>::kOut : 1.0f);
  if (a.nwin && part != 2) {
    const uint32_t ntiles = (a.nwin + kTile - 1) / kTile;
    dim3 grid(grid_for(ntiles, 4)), block(256);
    switch (mo) {
      case kMsgF16:
        hipLaunchKernelGGL(k_seg_sum<kMsgF16>, grid, block, 0, s, a.heads, a.tile_h0, a.nwin, (const __half*)a.Xo,
                           a.win_idx, a.S, a.pf_v, a.pl_v, a.mid);
        break;
      case kMsgBF16:
        hipLaunchKernelGGL(k_seg_sum<kMsgBF16>, grid, block, 0, s, a.heads, a.tile_h0, a.nwin, (const uint16_t*)a.Xo,
                           a.win_idx, a.S, a.pf_v, a.pl_v, a.mid);
        break;
      default:
        hipLaunchKernelGGL(k_seg_sum<kMsgF32>, grid, block, 0, s, a.heads, a.tile_h0, a.nwin, (const float*)a.Xo,
                           a.win_idx, a.S, a.pf_v, a.pl_v, a.mid);
    }
  }
  if (part == 1) return;
  // the segments that cross tiles are joined by k_seg_update
  hipLaunchKernelGGL(k_seg_update, dim3(pair_grid(a)), dim3(256), 0, s, a.pair_vox, a.pair_c, a.pair_sf, a.nwp,
                     a.npairs, a.wseg_off, a.S, a.pl_v, a.pf_v, a.mid, out_scale, a.B, a.M, a.upd, a.all_pairs);
}

void launch_reslice(FrontRec* front, RayRec* lanes, uint32_t npix, const uint2* sell, const uint32_t* off_f,
                    const uint32_t* len_f, const uint32_t* off_w, const uint32_t* len_w, uint32_t base_w, uint2* inc,
                    uint8_t* any, uint8_t* wany, unsigned long long* counts, hipStream_t s) {
  if (npix == 0) return;
  hipLaunchKernelGGL(k_reslice, dim3(grid_for(npix, 256)), dim3(256), 0, s, front, lanes, npix, sell, off_f, len_f,
                     off_w, len_w, base_w, inc, any, wany, counts);
}

void launch_brick_starts(const FrontRec* front, const RayRec* lanes, const uint32_t* len_w, uint32_t nslices,
                         const uint2* inc, uint32_t base_w, uint64_t* bstart, uint8_t* tail, hipStream_t s) {
  if (nslices == 0) return;
  hipLaunchKernelGGL(k_brick_starts, dim3(grid_for(nslices, 4)), dim3(256), 0, s, front, lanes, len_w, nslices, inc,
                     base_w, bstart, tail);
}

void launch_chunk_counts(const uint32_t* len_f, const uint32_t* len_w, uint32_t nslices, uint32_t* cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_chunk_counts, dim3(grid_for(nslices + 1, 256)), dim3(256), 0, s, len_f, len_w, nslices, cnt);
}

void launch_chunk_desc(const uint32_t* chunk_off, const uint32_t* len_f, uint32_t nslices, uint4* desc, hipStream_t s) {
  if (nslices == 0) return;
  hipLaunchKernelGGL(k_chunk_desc, dim3(grid_for(nslices, 256)), dim3(256), 0, s, chunk_off, len_f, nslices, desc);
}
void launch_front_span(const FrontRec* front, const RayRec* lanes, uint32_t n, uint2* span, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_front_span, dim3(grid_for(n, 256)), dim3(256), 0, s, front, lanes, n, span);
}
void launch_gather_desc(const uint4* desc, const uint32_t* idx, uint32_t n, uint4* out, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_gather_desc, dim3(grid_for(n, 256)), dim3(256), 0, s, desc, idx, n, out);
}

void launch_chunk_static(const CamPass& a, const uint32_t* pflag, uint32_t* dyn_flag, uint32_t* dmask, uint32_t* rows,
                         unsigned long long* nstatic, hipStream_t s) {
  if (a.nchunks == 0) return;
  hipLaunchKernelGGL(k_chunk_static, dim3(grid_for(a.nchunks, 16)), dim3(1024), 0, s, a.fspan, a.chunk_desc, a.nchunks,
                     a.inc, a.nwp, pflag, dyn_flag, dmask, rows, nstatic);
}

void launch_dyn_rows(const uint4* dyn_desc, uint32_t n, const uint32_t* rows, uint32_t* out, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_dyn_rows, dim3(grid_for(n, 256)), dim3(256), 0, s, dyn_desc, n, rows, out);
}

void launch_fill_dfront(const CamPass& a, const uint32_t* rows, const uint32_t* row_off, uint2* dfront, uint2* dcr,
                        hipStream_t s) {
  if (a.ndyn_chunks == 0) return;
  hipLaunchKernelGGL(k_fill_dfront, dim3(grid_for(a.ndyn_chunks, 4)), dim3(256), 0, s, a.fspan, a.dyn_desc,
                     a.ndyn_chunks, a.inc, a.fdmask, rows, row_off, dfront, dcr);
}

void launch_marginals(const float* B, float* A, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_marginals, dim3(grid_for(n, 256)), dim3(256), 0, s, B, A, n);
}

void launch_residual(const float* B, float* A, size_t n, unsigned long long* d_max, hipStream_t s) {
  if (n == 0) return;
  unsigned g = grid_for(n, 256);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_residual, dim3(g), dim3(256), 0, s, B, A, n, d_max);
}

void launch_ref_error(const float* A, const uint8_t* upd, const float* ref, size_t n, unsigned long long* d_max,
                      hipStream_t s) {
  launch_zero_words(d_max, 8, s);
  if (n == 0) return;
  size_t g = grid_for(n, 256);
  if (g > 2048) g = 2048;
  hipLaunchKernelGGL(k_ref_error, dim3(g), dim3(256), 0, s, A, upd, ref, n, d_max);
}

void launch_alphas(const float* B, const uint8_t* upd, float* out, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_alphas, dim3(grid_for(n, 256)), dim3(256), 0, s, B, upd, out, n);
}

void launch_set_voxels(const uint32_t* vox, const float* logodds, uint32_t n, float* B, uint8_t* upd, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_set_voxels, dim3(grid_for(n, 256)), dim3(256), 0, s, vox, logodds, n, B, upd);
}

void launch_cum_add(const uint32_t* pair_vox, const float* pair_c, uint32_t npairs, float* cum, hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_cum_add, dim3(grid_for(npairs, 256)), dim3(256), 0, s, pair_vox, pair_c, npairs, cum);
}

void launch_reset_empty(const float* cum, size_t n, float prior_logodds, float* B, uint8_t* upd, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_reset_empty, dim3(grid_for(n, 256)), dim3(256), 0, s, cum, n, prior_logodds, B, upd);
}

void launch_reset_pair_msgs(const uint32_t* pair_vox, uint32_t npairs, const float* cum, float* M, hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_reset_pair_msgs, dim3(grid_for(npairs, 256)), dim3(256), 0, s, pair_vox, npairs, cum, M);
}

void launch_scatter_pairs(const uint32_t* pair_vox, const float* M, uint32_t npairs, float* dense, hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_scatter_pairs, dim3(grid_for(npairs, 256)), dim3(256), 0, s, pair_vox, M, npairs, dense);
}

void launch_gather_pairs(const uint32_t* pair_vox, const float* dense, uint32_t npairs, float* M, hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_gather_pairs, dim3(grid_for(npairs, 256)), dim3(256), 0, s, pair_vox, dense, npairs, M);
}

void launch_fill_f32(float* p, float v, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_countdown(uint32_t* p, hipStream_t s) { hipLaunchKernelGGL(k_countdown, dim3(1), dim3(64), 0, s, p); }

void launch_mark_window_voxels(const uint32_t* pair_vox, uint32_t nwp, uint32_t* W, hipStream_t s) {
  if (nwp) hipLaunchKernelGGL(k_mark_window_voxels, dim3(grid_for(nwp, 256)), dim3(256), 0, s, pair_vox, nwp, W);
}

void launch_dyn_flags(const uint32_t* pair_vox, uint32_t n, const uint32_t* W, uint32_t* flags, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_dyn_flags, dim3(grid_for(n, 256)), dim3(256), 0, s, pair_vox, n, W, flags);
}

void launch_dyn_compact(const uint32_t* flags, const uint32_t* pos, uint32_t n, uint32_t q0, uint32_t* dyn,
                        hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_dyn_compact, dim3(grid_for(n, 256)), dim3(256), 0, s, flags, pos, n, q0, dyn);
}

void launch_dyn_pairs(const uint32_t* flags, const uint32_t* pos, uint32_t n, uint32_t q0, const uint32_t* pair_vox,
                      uint2* dyn, hipStream_t s) {
  if (n) hipLaunchKernelGGL(k_dyn_pairs, dim3(grid_for(n, 256)), dim3(256), 0, s, flags, pos, n, q0, pair_vox, dyn);
}

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
}

void launch_zero_words(void* p, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 4;
  if (n == 0) return;
  hipLaunchKernelGGL(k_zero_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, (uint32_t*)p, n);
}

}  // namespace mrf
