// Loopy-BP message kernels of the MRFMap inference path.
//
//  * k_ray_messages: one wavefront per ray.  Replaces the pre pass
//    (rayStartComputingMessagesFromBrickPre, src/cuda/gvdbkernels.cu:689-819),
//    the infinity term (gvdbkernels.cu:1196-1208) and the post pass
//    (rayStartComputingMessagesFromBrickPost, gvdbkernels.cu:824-1038) with a
//    single launch: 64 incidences per step, DPP prefix sums for
//    Lambda = sum l ln n, for the travelled length and for the exclusive sum
//    of tau, a wave reduction for T, and a scatter of p*l into the
//    camera's ray->voxel accumulator.
//  * k_cum_len / pair build: the per-camera summed chord length C(v)
//    (rayComputeCumLens, gvdbkernels.cu:519-571) is geometry only, so it is
//    computed once per traversal with exact 2^-32 fixed-point atomics and
//    compacted into an ascending (voxel, C) pair list.
//  * k_epilogue: per (camera, touched voxel) pair: camera message
//    m = logit(clamp(S/C)), B += m - M_c, M_c = m, S = 0
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
struct Poly {
  float c0, c1, c2;
  int n;
  __device__ __forceinline__ float operator()(float z) const {
    return n == 3 ? (z * z * c2 + z * c1 + c0) : (z * c1 + c0);
  }
};

// psi(s; d) = N(d; s + bias(s), (3 sigma(s))^2) (depth_potential, gvdbkernels.cu:57-72).
__device__ __forceinline__ float depth_psi(float s, float d, const Poly& sig, const Poly& bia) {
  const float diff = fabsf(d - (s + bia(s)));
  const float inv = __builtin_amdgcn_rcpf(3.0f * sig(s));
  const float u = diff * inv;
  return 0.398942f * __expf(-0.5f * u * u) * inv;
}

// Kahan-compensated running sum of the per-chunk totals along a ray.
struct Kahan {
  float s, c;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  // carry + in-chunk prefix, corrected by the compensation
  __device__ __forceinline__ float at(float prefix) const { return s + (prefix - c); }
};

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

// Voxel -> ray message of every pair: ln n = ln clamp(1/(1+exp(B - M_c))),
// camera c's own previous message excluded from the belief (gvdbkernels.cu:764-773).
__global__ __launch_bounds__(256) void k_pair_msg(const uint32_t* __restrict__ pair_vox, uint32_t npairs,
                                                  const float* __restrict__ B, const float* __restrict__ M,
                                                  float* __restrict__ apair) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  const uint32_t v = pair_vox[q];
  const float qv = B[v] - M[v];
  float nn = 1.0f / (1.0f + __expf(qv));
  nn = fminf(fmaxf(nn, kFEps), 1.0f - kFEps);
  apair[q] = __logf(nn);
}

// Copy each pair's message into every slot of its segment (one wavefront
// per tile of kTile positions, coalesced writes).
template <int MT>
__global__ __launch_bounds__(256) void k_expand(const uint64_t* __restrict__ heads, const uint32_t* __restrict__ tile_h0,
                                                uint32_t nlive, const float* __restrict__ apair,
                                                typename Msg<MT>::T* __restrict__ X) {
  constexpr int C = kTile / 64;
  const int lane = threadIdx.x & 63;
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t j0 = w * kTile;
  if (j0 >= nlive) return;
  const uint32_t j1 = min(j0 + kTile, nlive);
  uint32_t before = tile_h0[w];
  float v[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const uint32_t base = j0 + 64u * c;
    const uint64_t word = base < j1 ? heads[base >> 6] : 0ull;
    const uint32_t j = base + lane;
    v[c] = j < j1 ? apair[pair_of(word, before, lane)] : 0.0f;
    before += (uint32_t)__popcll(word);
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const uint32_t j = j0 + 64u * c + lane;
    if (j < j1) Msg<MT>::store(X, j, v[c]);
  }
}

// One lane per ray, one wavefront per slice (the 64 rays of an 8x8 pixel
// tile, records interleaved so that record k of all 64 rays is one coalesced
// load and the 64 lanes touch neighbouring voxels together).
//
// Pass 1: Lambda_k = sum_{j<=k} l_j ln n_j (the reference's visibility
// includes voxel k itself, gvdbkernels.cu:781-783), vpsi_k = exp(Lambda_k)
// psi(s_k), tau_k = vpsi_k (1 - n_k), T = sum tau (+ V_n when the walk ran
// out and d lies beyond, gvdbkernels.cu:1196-1208).  Until the first record
// with vpsi > 0 every earlier tau is exactly 0, so P = 0, mu1 = 0 and the
// message is exactly the clamp floor F_EPS: pass 1 writes those directly.
// Pass 2 restarts from that record with the saved sums and writes
// p = mu1 / (mu1 + mu0 + 1e-7), mu1 = P + vpsi, mu0 = P + R / n
// (gvdbkernels.cu:936-951).  Long sums along the ray are Kahan-compensated.
struct RayCtx {
  Poly sig, bia;
  float d, z0, inv;
};

template <int MI, int MO>
__global__ __launch_bounds__(256) void k_ray_messages(const RayRec* __restrict__ lanes,
                                                      const uint32_t* __restrict__ slice_len, uint32_t nslices,
                                                      const uint2* __restrict__ inc,
                                                      const typename Msg<MI>::T* __restrict__ Xi,
                                                      typename Msg<MO>::T* __restrict__ Xo,
                                                      const float* __restrict__ sigt, const float* __restrict__ biat,
                                                      int lookup_n) {
  const int lane = threadIdx.x & 63;
  // XCD-aware order: consecutive blocks (neighbouring pixel tiles) share an XCD's L2
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t t = lb * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (t >= nslices) return;
  const uint32_t len = slice_len[t];
  if (len == 0) return;
  const RayRec rec = lanes[t * kSlice + lane];
  const uint32_t cell = rec.cell_flags & 0x7FFFFFFFu;
  RayCtx rc;
  rc.sig.n = rc.bia.n = lookup_n;
  rc.sig.c0 = sigt[cell];
  rc.sig.c1 = sigt[cell + 1];
  rc.sig.c2 = lookup_n == 3 ? sigt[cell + 2] : 0.0f;
  rc.bia.c0 = biat[cell];
  rc.bia.c1 = biat[cell + 1];
  rc.bia.c2 = lookup_n == 3 ? biat[cell + 2] : 0.0f;
  rc.d = rec.depth;
  rc.z0 = rec.z0;
  rc.inv = rec.inv_norm;
  const uint2* __restrict__ rl = inc + rec.off;  // record k at rl[k * kSlice]

  Kahan kL = {0.f, 0.f}, kS = {0.f, 0.f}, kT = {0.f, 0.f};
  Kahan L0 = kL, S0 = kS;
  uint32_t i0 = len;  // first record with vpsi > 0
  for (uint32_t k = 0; k < len; ++k) {
    const uint2 e = rl[(size_t)k * kSlice];
    const bool live = e.x != kGapVox;
    const float lnn = live ? Msg<MI>::load(Xi, e.x) : 0.0f;
    const float ln = __uint_as_float(e.y);
    const Kahan Lb = kL, Sb = kS;
    kL.add(ln * lnn);
    kS.add(ln);
    const float s = rc.z0 + kS.at(0.0f) * rc.inv;
    const float vpsi = live ? __expf(kL.at(0.0f)) * depth_psi(s, rc.d, rc.sig, rc.bia) : 0.0f;
    if (i0 == len) {
      if (vpsi > 0.0f) {
        i0 = k;
        L0 = Lb;
        S0 = Sb;
      } else if (live) {
        Msg<MO>::store(Xo, e.x, kFEps * ln * Msg<MO>::kOut);
      }
    }
    if (i0 != len) kT.add(vpsi * (1.0f - __expf(lnn)));
  }
  float T = kT.at(0.0f);
  if (rec.cell_flags & kRanOut) {
    const float sn = rc.z0 + kS.at(0.0f) * rc.inv;
    if (rc.d > sn + rc.bia(sn)) T += __expf(kL.at(0.0f));
  }

  uint32_t kmin = i0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) kmin = min(kmin, (uint32_t)__shfl_xor((int)kmin, o));
  kL = L0;
  kS = S0;
  Kahan kP = {0.f, 0.f};
  for (uint32_t k = kmin; k < len; ++k) {
    if (k < i0) continue;
    const uint2 e = rl[(size_t)k * kSlice];
    const bool live = e.x != kGapVox;
    const float lnn = live ? Msg<MI>::load(Xi, e.x) : 0.0f;
    const float ln = __uint_as_float(e.y);
    kL.add(ln * lnn);
    kS.add(ln);
    const float s = rc.z0 + kS.at(0.0f) * rc.inv;
    const float vpsi = live ? __expf(kL.at(0.0f)) * depth_psi(s, rc.d, rc.sig, rc.bia) : 0.0f;
    const float nn = __expf(lnn);
    const float tau = vpsi * (1.0f - nn);
    const float P = kP.at(0.0f);
    const float R = (T - P) - tau;
    if (live) {
      const float mu1 = P + vpsi;
      const float mu0 = P + R * __builtin_amdgcn_rcpf(nn);
      float p = mu1 * __builtin_amdgcn_rcpf(mu1 + mu0 + kMsgNormEps);
      p = fminf(fmaxf(p, kFEps), 1.0f - kFEps);
      Msg<MO>::store(Xo, e.x, p * ln * Msg<MO>::kOut);
    }
    kP.add(tau);
  }
}

// Segment sums S_q = sum of p*l over pair q's positions, in a fixed order.
// One wavefront per tile: a segmented inclusive scan per 64-wide chunk, the
// open segment carried from chunk to chunk.  Segments that start and end in
// the tile are written to S directly; the part of a segment that started in
// an earlier tile goes to pf, the part of one that continues into the next
// tile to pl, and a tile lying wholly inside one segment to mid.
constexpr uint32_t kNoPair = 0xFFFFFFFFu;

template <int MT>
__global__ __launch_bounds__(256) void k_seg_sum(const uint64_t* __restrict__ heads, const uint32_t* __restrict__ tile_h0,
                                                 uint32_t nlive, const typename Msg<MT>::T* __restrict__ X,
                                                 float* __restrict__ S, uint32_t* __restrict__ pf_q,
                                                 float* __restrict__ pf_v, uint32_t* __restrict__ pl_q,
                                                 float* __restrict__ pl_v, float* __restrict__ mid) {
  const int lane = threadIdx.x & 63;
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint32_t j0 = w * kTile;
  if (j0 >= nlive) return;
  const uint32_t j1 = min(j0 + kTile, nlive);
  uint32_t before = tile_h0[w];
  const bool open_first = ((heads[j0 >> 6] >> (j0 & 63)) & 1ull) == 0ull;  // tile starts inside a segment
  float carry = 0.0f;                                                      // open segment's running sum
  bool carry_from_before = open_first;
  uint32_t pfq = kNoPair, plq = kNoPair;
  float pfv = 0.0f, plv = 0.0f, midv = 0.0f;
  bool is_mid = false;
  constexpr int C = kTile / 64;
  float xs[C];
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const uint32_t j = j0 + 64u * c + lane;
    xs[c] = j < j1 ? Msg<MT>::load(X, j) : 0.0f;
  }
#pragma unroll
  for (int c = 0; c < C; ++c) {
    const uint32_t base = j0 + 64u * c;
    if (base >= j1) break;
    const uint64_t word = heads[base >> 6];
    const uint64_t next_word = (base + 64 < nlive) ? heads[(base >> 6) + 1] : ~0ull;
    const uint32_t j = base + lane;
    const bool valid = j < j1;
    const uint32_t q = pair_of(word, before, lane);
    int head = (int)((word >> lane) & 1ull);
    float x = xs[c];
    if (lane == 0 && !head) x += carry;  // continue the open segment
    if (lane == 0) head = 1;
#pragma unroll
    for (int dlt = 1; dlt < 64; dlt <<= 1) {
      const float y = __shfl_up(x, dlt);
      const int h = __shfl_up(head, dlt);
      if (lane >= dlt && !head) {
        x += y;
        head |= h;
      }
    }
    // a segment ends at j when j + 1 starts another pair or is past the data
    const bool nxt_head = lane == 63 ? (next_word & 1ull) != 0ull : ((word >> (lane + 1)) & 1ull) != 0ull;
    const bool ends = valid && (j + 1 >= nlive || nxt_head);
    // the chunk's first segment is the continuation of the carried one
    const uint64_t headbits = word & (lane == 63 ? ~0ull : ((2ull << lane) - 1ull));
    const bool in_first = headbits == 0ull;  // no head at or before this lane in this chunk
    if (ends) {
      if (in_first && carry_from_before) {
        pfq = q;  // unique per tile: lane-uniform values written once below
        pfv = x;
      } else {
        S[q] = x;
      }
    }
    // the last valid lane carries its open segment into the next chunk
    const uint32_t last = min(63u, j1 - 1 - base);
    const float xl = __shfl(x, (int)last);
    const int el = __shfl((int)ends, (int)last);
    const int firstl = __shfl((int)in_first, (int)last);
    const uint32_t ql = (uint32_t)__shfl((int)q, (int)last);
    carry_from_before = el ? false : (firstl ? carry_from_before : false);
    carry = el ? 0.0f : xl;
    if (base + 64 >= j1 && !el) {  // tile ends inside a segment
      if (carry_from_before) {
        is_mid = true;
        midv = xl;
      } else {
        plq = ql;
        plv = xl;
      }
    }
    before += (uint32_t)__popcll(word);
  }
  // pf was set by the single lane holding the end of the tile's first segment
  const unsigned long long pfmask = __ballot(pfq != kNoPair);
  if (lane == 0) {
    pf_q[w] = kNoPair;
    pl_q[w] = plq;
    pl_v[w] = plv;
    mid[w] = is_mid ? midv : 0.0f;
  }
  if (pfmask) {
    const int src = __ffsll((long long)pfmask) - 1;
    const uint32_t q = (uint32_t)__shfl((int)pfq, src);
    const float v = __shfl(pfv, src);
    if (lane == 0) {
      pf_q[w] = q;
      pf_v[w] = v;
    }
  }
}

// Pairs whose segment crosses tiles: S = pl(head tile) + sum of mid over the
// tiles in between + pf(tail tile); one wavefront per head tile, the middle
// tiles reduced 64 at a time in a fixed order.
__global__ __launch_bounds__(256) void k_seg_span(const uint32_t* __restrict__ pair_off, const uint32_t* __restrict__ pl_q,
                                                  const float* __restrict__ pl_v, const float* __restrict__ pf_v,
                                                  const float* __restrict__ mid, uint32_t ntiles,
                                                  float* __restrict__ S) {
  const int lane = threadIdx.x & 63;
  const uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= ntiles) return;
  const uint32_t q = pl_q[w];
  if (q == kNoPair) return;
  const uint32_t wt = (pair_off[q + 1] - 1u) / kTile;  // tile holding the segment's last position
  float acc = 0.0f;
  for (uint32_t b = w + 1; b < wt; b += 64) {
    const uint32_t t = b + lane;
    acc += lane63(wave_scan(t < wt ? mid[t] : 0.0f));
  }
  if (lane == 0) S[q] = (pl_v[w] + acc) + pf_v[wt];
}

// Camera message of each (camera, voxel) pair from its segment sum:
// a = clamp(S/C), m = ln a - ln(1-a), B += m - M_c, M_c = m, for C > 0.1
// (compute_gvdb_message_to_factors, gvdbkernels.cu:408, 457-479).
__global__ __launch_bounds__(256) void k_seg_update(const uint32_t* __restrict__ pair_vox,
                                                    const float* __restrict__ pair_c, uint32_t npairs,
                                                    const float* __restrict__ S, float out_scale, float* __restrict__ B,
                                                    float* __restrict__ M, uint8_t* __restrict__ upd) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  const float C = pair_c[q];
  if (!(C > kCumLenMin)) return;
  const uint32_t v = pair_vox[q];
  float a = S[q] / (C * out_scale);
  a = fminf(fmaxf(a, kFEps), 1.0f - kFEps);
  const float m = __logf(a) - __logf(1.0f - a);
  const float mo = M[v];
  M[v] = m;
  B[v] = (B[v] + m) - mo;
  upd[v] = 1;
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

__global__ void k_alphas(const float* __restrict__ B, const uint8_t* __restrict__ upd, float* __restrict__ out,
                         size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = upd[i] ? marginal(B[i]) : 0.0f;
}

__global__ void k_fill_f32(float* p, float v, size_t n) {
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

}  // namespace

void launch_expand(const CamPass& a, hipStream_t s) {
  if (a.npairs == 0) return;
  hipLaunchKernelGGL(k_pair_msg, dim3(grid_for(a.npairs, 256)), dim3(256), 0, s, a.pair_vox, a.npairs, a.B, a.M,
                     a.apair);
  const uint32_t ntiles = (a.nlive + kTile - 1) / kTile;
  dim3 grid(grid_for(ntiles, 4)), block(256);
  switch (msg_in_type(a.msg_type)) {
    case kMsgF16:
      hipLaunchKernelGGL(k_expand<kMsgF16>, grid, block, 0, s, a.heads, a.tile_h0, a.nlive, a.apair, (__half*)a.X);
      break;
    case kMsgBF16:
      hipLaunchKernelGGL(k_expand<kMsgBF16>, grid, block, 0, s, a.heads, a.tile_h0, a.nlive, a.apair,
                         (uint16_t*)a.X);
      break;
    default:
      hipLaunchKernelGGL(k_expand<kMsgF32>, grid, block, 0, s, a.heads, a.tile_h0, a.nlive, a.apair, (float*)a.X);
  }
}

void launch_ray_messages(const CamPass& a, hipStream_t s) {
  if (a.nslices == 0) return;
  const int waves = 4;
  // grid padded to a multiple of the 8 XCDs for the block remapping
  const unsigned nb = ((a.nslices + waves - 1) / waves + 7) / 8 * 8;
  dim3 grid(nb), block(64 * waves);
#define MRF_RAY(MI, MO)                                                                                        \
  hipLaunchKernelGGL((k_ray_messages<MI, MO>), grid, block, 0, s, a.lanes, a.slice_len, a.nslices, a.inc,       \
                     (const typename Msg<MI>::T*)a.X, (typename Msg<MO>::T*)a.Xo, a.sig, a.bia, a.lookup_n)
  switch (a.msg_type) {
    case kMsgF16: MRF_RAY(kMsgF16, kMsgF16); break;
    case kMsgBF16: MRF_RAY(kMsgBF16, kMsgBF16); break;
    case kMsgMixedF16: MRF_RAY(kMsgF32, kMsgF16); break;
    case kMsgMixedBF16: MRF_RAY(kMsgF32, kMsgBF16); break;
    default: MRF_RAY(kMsgF32, kMsgF32);
  }
#undef MRF_RAY
}

void launch_segment_sum_f32(const uint64_t* heads, const uint32_t* tile_h0, uint32_t nlive, const float* X, float* S,
                            uint32_t* pf_q, float* pf_v, uint32_t* pl_q, float* pl_v, float* mid,
                            const uint32_t* pair_off, hipStream_t s) {
  if (nlive == 0) return;
  const uint32_t ntiles = (nlive + kTile - 1) / kTile;
  dim3 grid(grid_for(ntiles, 4)), block(256);
  hipLaunchKernelGGL(k_seg_sum<kMsgF32>, grid, block, 0, s, heads, tile_h0, nlive, X, S, pf_q, pf_v, pl_q, pl_v, mid);
  hipLaunchKernelGGL(k_seg_span, grid, block, 0, s, pair_off, pl_q, pl_v, pf_v, mid, ntiles, S);
}

void launch_epilogue(const CamPass& a, hipStream_t s) {
  if (a.npairs == 0) return;
  const uint32_t ntiles = (a.nlive + kTile - 1) / kTile;
  dim3 grid(grid_for(ntiles, 4)), block(256);
  float out_scale = 1.0f;
  switch (msg_out_type(a.msg_type)) {
    case kMsgF16:
      hipLaunchKernelGGL(k_seg_sum<kMsgF16>, grid, block, 0, s, a.heads, a.tile_h0, a.nlive, (const __half*)a.Xo,
                         a.S, a.pf_q, a.pf_v, a.pl_q, a.pl_v, a.mid);
      hipLaunchKernelGGL(k_seg_span, grid, block, 0, s, a.pair_off, a.pl_q, a.pl_v, a.pf_v, a.mid, ntiles, a.S);
      out_scale = Msg<kMsgF16>::kOut;
      break;
    case kMsgBF16:
      hipLaunchKernelGGL(k_seg_sum<kMsgBF16>, grid, block, 0, s, a.heads, a.tile_h0, a.nlive, (const uint16_t*)a.Xo,
                         a.S, a.pf_q, a.pf_v, a.pl_q, a.pl_v, a.mid);
      hipLaunchKernelGGL(k_seg_span, grid, block, 0, s, a.pair_off, a.pl_q, a.pl_v, a.pf_v, a.mid, ntiles, a.S);
      out_scale = Msg<kMsgBF16>::kOut;
      break;
    default:
      launch_segment_sum_f32(a.heads, a.tile_h0, a.nlive, (const float*)a.Xo, a.S, a.pf_q, a.pf_v, a.pl_q, a.pl_v,
                             a.mid, a.pair_off, s);
  }
  hipLaunchKernelGGL(k_seg_update, dim3(grid_for(a.npairs, 256)), dim3(256), 0, s, a.pair_vox, a.pair_c, a.npairs,
                     a.S, out_scale, a.B, a.M, a.upd);
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

void launch_alphas(const float* B, const uint8_t* upd, float* out, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_alphas, dim3(grid_for(n, 256)), dim3(256), 0, s, B, upd, out, n);
}

void launch_fill_f32(float* p, float v, size_t n, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(n, 256)), dim3(256), 0, s, p, v, n);
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
