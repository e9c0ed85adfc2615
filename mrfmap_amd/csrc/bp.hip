// Loopy-BP message kernels of the MRFMap inference path.
//
//  * k_pair_msg: voxel->ray message ln n of every (camera, voxel) pair
//    (gvdbkernels.cu:764-773) into a compact per-pair array.
//  * k_ray_messages: L lanes per ray over sliced-ELL pixel tiles; reads ln n
//    of each record's pair, writes p*l into the record's slot of the
//    camera's voxel-sorted segment array.  Replaces
//    the pre pass (rayStartComputingMessagesFromBrickPre,
//    src/cuda/gvdbkernels.cu:689-819), the infinity term
//    (gvdbkernels.cu:1196-1208) and the post pass
//    (rayStartComputingMessagesFromBrickPost, gvdbkernels.cu:824-1038) with a
//    single launch.
//  * k_seg_sum / k_seg_span / k_seg_update: per pair, the fixed-order sum S of
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
struct Poly {
  float c0, c1, c2;
  int n;
  __device__ __forceinline__ float operator()(float z) const {
    return n == 3 ? fmaf(z, fmaf(z, c2, c1), c0) : fmaf(z, c1, c0);
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
template <int MT>
__global__ __launch_bounds__(256) void k_pair_msg(const uint32_t* __restrict__ pair_vox, uint32_t npairs,
                                                  const float* __restrict__ B, const float* __restrict__ M,
                                                  typename Msg<MT>::T* __restrict__ apair) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  const uint32_t v = pair_vox[q];
  const float qv = B[v] - M[v];
  float nn = 1.0f / (1.0f + __expf(qv));
  nn = fminf(fmaxf(nn, kFEps), 1.0f - kFEps);
  Msg<MT>::store(apair, q, __logf(nn));
}

// ---------------------------------------------------------------- ray kernel
// L lanes per ray (L = 1, 2, 4): lane j of a ray handles records j, j+L,
// j+2L, ...; each step covers L consecutive records with a prefix sum over
// the ray's L lanes (DPP quad permutes) and a carried, Kahan-compensated sum
// from earlier steps.  The 64 rays of a slice (one 8x8 pixel tile) are
// spread over L wavefronts; record k of all rays of a slice is contiguous.
template <int CTRL>
__device__ __forceinline__ float qperm(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ int qperm_i(int x) {
  return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, false);
}

template <int L>
struct Group {
  // inclusive prefix over the L lanes of a ray (j = lane index within the ray)
  __device__ __forceinline__ static float scan(float x, int j) {
    if (L == 2) {
      const float y = qperm<0xA0>(x);  // quad_perm [0,0,2,2]
      return j == 1 ? x + y : x;
    } else if (L == 4) {
      float y = qperm<0x90>(x);        // quad_perm [0,0,1,2]
      x = j >= 1 ? x + y : x;
      y = qperm<0x40>(x);              // quad_perm [0,0,0,1]
      return j >= 2 ? x + y : x;
    }
    return x;
  }
  // value of the ray's last lane
  __device__ __forceinline__ static float last(float x) {
    if (L == 2) return qperm<0xF5>(x);  // quad_perm [1,1,3,3]
    if (L == 4) return qperm<0xFF>(x);  // quad_perm [3,3,3,3]
    return x;
  }
  __device__ __forceinline__ static int min(int v) {
    if (L >= 2) v = ::min(v, qperm_i<0xB1>(v));  // quad_perm [1,0,3,2]
    if (L == 4) v = ::min(v, qperm_i<0x4E>(v));  // quad_perm [2,3,0,1]
    return v;
  }
  __device__ __forceinline__ static float sum(float x) {
    if (L >= 2) x = x + qperm<0xB1>(x);
    if (L == 4) x = x + qperm<0x4E>(x);
    return x;
  }
};

struct RayCtx {
  Poly sig, bia;
  float d, z0, inv;
};

// Pass 1: Lambda_k = sum_{i<=k} l_i ln n_i (the reference's visibility
// includes voxel k itself, gvdbkernels.cu:781-783), vpsi_k = exp(Lambda_k)
// psi(s_k), tau_k = vpsi_k (1 - n_k), T = sum tau (+ V_n when the walk ran
// out and d lies beyond, gvdbkernels.cu:1196-1208).  Until the first record
// with vpsi > 0 every earlier tau is exactly 0, so P = 0, mu1 = 0 and the
// message is exactly the clamp floor F_EPS: pass 1 writes those of the
// steps before the one holding that record directly.
// Pass 2 restarts at the step holding that record with the saved sums and
// writes p = mu1 / (mu1 + mu0 + 1e-7), mu1 = P + vpsi, mu0 = P + R / n
// (gvdbkernels.cu:936-951).  U steps of loads are issued before their
// arithmetic.  The file is built without FP contraction, so pass 2
// recomputes pass 1's terms bit for bit.
template <int MI, int MO, int L, int U>
__global__ __launch_bounds__(256) void k_ray_messages(const RayRec* __restrict__ lanes,
                                                      const uint32_t* __restrict__ slice_len, uint32_t nslices,
                                                      const uint2* __restrict__ inc,
                                                      const uint32_t* __restrict__ inc_pair,
                                                      const typename Msg<MI>::T* __restrict__ Xi,
                                                      typename Msg<MO>::T* __restrict__ Xo,
                                                      const float* __restrict__ sigt, const float* __restrict__ biat,
                                                      int lookup_n) {
  constexpr uint32_t kRaysPerBlock = 256 / L;
  // XCD-aware order: consecutive blocks (neighbouring pixel tiles) share an XCD's L2
  const uint32_t per = gridDim.x >> 3;
  const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
  const uint32_t ray = lb * kRaysPerBlock + threadIdx.x / L;  // global ray id = slice * 64 + r
  const int j = (int)(threadIdx.x % L);
  const uint32_t t = ray >> 6;
  if (t >= nslices) return;  // uniform per wavefront: a wavefront never spans slices
  const uint32_t len = slice_len[t];
  if (len == 0) return;
  const RayRec rec = lanes[ray];
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
  const uint32_t* __restrict__ rq = inc_pair + rec.off;
  const uint32_t nsteps = (len + L - 1) / L;

  auto load_rec = [&](uint32_t step) -> uint2 {
    const uint32_t k = step * L + j;
    return k < len ? rl[(size_t)k * kSlice] : make_uint2(kGapVox, 0u);
  };
  auto load_pair = [&](uint32_t step) -> uint32_t {
    const uint32_t k = step * L + j;
    return k < len ? rq[(size_t)k * kSlice] : 0u;
  };
  // voxel -> ray message of the record's pair (gap records read pair 0)
  auto load_msg = [&](const uint2 e, const uint32_t q) -> float {
    return Msg<MI>::load(Xi, e.x != kGapVox ? q : 0u);
  };

  Kahan kL = {0.f, 0.f}, kS = {0.f, 0.f}, kT = {0.f, 0.f};
  Kahan L0 = kL, S0 = kS;
  uint32_t i0 = len;  // first record with vpsi > 0 (same in all lanes of a ray)
  auto step1 = [&](const uint2 e, const float msg, const uint32_t step) {
    const uint32_t k = step * L + j;
    const bool live = e.x != kGapVox;
    const float lnn = live ? msg : 0.0f;
    const float ln = __uint_as_float(e.y);
    const float a = Group<L>::scan(ln * lnn, j);
    const float b = Group<L>::scan(ln, j);
    const float s = rc.z0 + kS.at(b) * rc.inv;
    const float ps = depth_psi(s, rc.d, rc.sig, rc.bia);
    const float vpsi = live ? __expf(kL.at(a)) * ps : 0.0f;
    const bool front = i0 == len;
    const int firstj = Group<L>::min(vpsi > 0.0f ? j : L);
    const bool found = front && firstj < L;
    if (found) {
      i0 = step * L + (uint32_t)firstj;
      L0 = kL;
      S0 = kS;
    }
    const bool before = front && k < i0;  // record in front of the likelihood window
    // steps wholly in front are written here, the step holding i0 by pass 2
    if (front && !found && live) Msg<MO>::store(Xo, e.x, kFEps * ln * Msg<MO>::kOut);
    kT.add(before ? 0.0f : vpsi * (1.0f - __expf(lnn)));
    kL.add(Group<L>::last(a));
    kS.add(Group<L>::last(b));
  };
  uint32_t st = 0;
  for (; st + U <= nsteps; st += U) {
    uint2 e[U];
    uint32_t qv[U];
    float x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      e[u] = load_rec(st + u);
      qv[u] = load_pair(st + u);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = load_msg(e[u], qv[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) step1(e[u], x[u], st + u);
  }
  for (; st < nsteps; ++st) {
    const uint2 e = load_rec(st);
    step1(e, load_msg(e, load_pair(st)), st);
  }
  float T = Group<L>::sum(kT.at(0.0f));
  if (rec.cell_flags & kRanOut) {
    const float sn = rc.z0 + kS.at(0.0f) * rc.inv;
    if (rc.d > sn + rc.bia(sn)) T += __expf(kL.at(0.0f));
  }

  // pass 2 from the step holding i0 (all lanes execute every step; writes
  // and state updates are predicated so the group permutes stay convergent).
  // Records of that step in front of i0 have vpsi = 0 and P = 0, so their p
  // clamps to exactly F_EPS, the value pass 1 gives the records before them.
  // A ray without a record with vpsi > 0 was written wholly by pass 1.
  const uint32_t s0 = i0 == len ? nsteps : i0 / L;
  uint32_t smin = s0;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) smin = ::min(smin, (uint32_t)__shfl_xor((int)smin, o));
  kL = L0;
  kS = S0;
  Kahan kP = {0.f, 0.f};
  auto step2 = [&](const uint2 e, const float msg, const uint32_t step) {
    const bool on = step >= s0;
    const bool live = e.x != kGapVox;
    const float lnn = live ? msg : 0.0f;
    const float ln = __uint_as_float(e.y);
    const float a = Group<L>::scan(ln * lnn, j);
    const float b = Group<L>::scan(ln, j);
    const float s = rc.z0 + kS.at(b) * rc.inv;
    const float ps = depth_psi(s, rc.d, rc.sig, rc.bia);
    const float vpsi = live ? __expf(kL.at(a)) * ps : 0.0f;
    const float nn = __expf(lnn);
    const float tau = vpsi * (1.0f - nn);
    const float pin = Group<L>::scan(tau, j);
    const float P = kP.at(pin - tau);
    const float R = (T - P) - tau;
    if (on && live) {
      const float mu1 = P + vpsi;
      const float mu0 = P + R * __builtin_amdgcn_rcpf(nn);
      float p = mu1 * __builtin_amdgcn_rcpf(mu1 + mu0 + kMsgNormEps);
      p = fminf(fmaxf(p, kFEps), 1.0f - kFEps);
      Msg<MO>::store(Xo, e.x, p * ln * Msg<MO>::kOut);
    }
    const float ta = Group<L>::last(a), tb = Group<L>::last(b), tp = Group<L>::last(pin);
    if (on) {
      kL.add(ta);
      kS.add(tb);
      kP.add(tp);
    }
  };
  st = smin;
  for (; st + U <= nsteps; st += U) {
    uint2 e[U];
    uint32_t qv[U];
    float x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      e[u] = load_rec(st + u);
      qv[u] = load_pair(st + u);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = load_msg(e[u], qv[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) step2(e[u], x[u], st + u);
  }
  for (; st < nsteps; ++st) {
    const uint2 e = load_rec(st);
    step2(e, load_msg(e, load_pair(st)), st);
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
// segment to mid; k_seg_span adds them up.
constexpr uint32_t kNoPair = 0xFFFFFFFFu;

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

// 8 consecutive messages from p0 (a multiple of 8), zero past j1
template <int MT>
__device__ __forceinline__ void load8(const typename Msg<MT>::T* __restrict__ X, uint32_t p0, uint32_t j1,
                                      float (&x)[8]) {
  typedef typename Msg<MT>::T T;
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
                                                 float* __restrict__ S, uint32_t* __restrict__ pf_q,
                                                 float* __restrict__ pf_v, uint32_t* __restrict__ pl_q,
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
  load8<MT>(X, p0, j1, x);
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
    uint32_t pq = kNoPair, lq = kNoPair;
    float pv = 0.0f, lv = 0.0f, mv = 0.0f;
    if (pfmask) {
      pq = tile_h0[w] - 1u;
      pv = pfv0;
    }
    if (ends_here) {
      if (Ft)
        S[qt] = Ut;
      else {
        pq = tile_h0[w] - 1u;  // no head in the tile: the whole tile ends an earlier segment
        pv = Ut;
      }
    } else if (Ft) {
      lq = qt;
      lv = Ut;
    } else {
      mv = Ut;  // the tile lies inside one segment
    }
    pf_q[w] = pq;
    pf_v[w] = pv;
    pl_q[w] = lq;
    pl_v[w] = lv;
    mid[w] = mv;
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

void launch_pair_msgs(const CamPass& a, hipStream_t s) {
  if (a.npairs == 0) return;
  const dim3 grid(grid_for(a.npairs, 256)), block(256);
  switch (msg_in_type(a.msg_type)) {
    case kMsgF16:
      hipLaunchKernelGGL(k_pair_msg<kMsgF16>, grid, block, 0, s, a.pair_vox, a.npairs, a.B, a.M, (__half*)a.apair);
      break;
    case kMsgBF16:
      hipLaunchKernelGGL(k_pair_msg<kMsgBF16>, grid, block, 0, s, a.pair_vox, a.npairs, a.B, a.M,
                         (uint16_t*)a.apair);
      break;
    default:
      hipLaunchKernelGGL(k_pair_msg<kMsgF32>, grid, block, 0, s, a.pair_vox, a.npairs, a.B, a.M, (float*)a.apair);
  }
}

template <int MI, int MO, int L, int U>
void launch_ray_t(const CamPass& a, hipStream_t s) {
  const uint32_t blocks = (a.nslices * kSlice + (256 / L) - 1) / (256 / L);
  const unsigned nb = (blocks + 7) / 8 * 8;  // multiple of the 8 XCDs for the block remapping
  hipLaunchKernelGGL((k_ray_messages<MI, MO, L, U>), dim3(nb), dim3(256), 0, s, a.lanes, a.slice_len, a.nslices,
                     a.inc, a.inc_pair, (const typename Msg<MI>::T*)a.apair, (typename Msg<MO>::T*)a.Xo, a.sig,
                     a.bia, a.lookup_n);
}

template <int MI, int MO>
void launch_ray_sched(const CamPass& a, hipStream_t s) {
  switch (a.ray_sched) {
    case kSchedL1U4: launch_ray_t<MI, MO, 1, 4>(a, s); break;
    case kSchedL1U8: launch_ray_t<MI, MO, 1, 8>(a, s); break;
    case kSchedL2U4: launch_ray_t<MI, MO, 2, 4>(a, s); break;
    case kSchedL4U2: launch_ray_t<MI, MO, 4, 2>(a, s); break;
    case kSchedL4U4: launch_ray_t<MI, MO, 4, 4>(a, s); break;
    default: launch_ray_t<MI, MO, 2, 8>(a, s);
  }
}

void launch_ray_messages(const CamPass& a, hipStream_t s) {
  if (a.nslices == 0) return;
  switch (a.msg_type) {
    case kMsgF16: launch_ray_sched<kMsgF16, kMsgF16>(a, s); break;
    case kMsgBF16: launch_ray_sched<kMsgBF16, kMsgBF16>(a, s); break;
    case kMsgMixedF16: launch_ray_sched<kMsgF32, kMsgF16>(a, s); break;
    case kMsgMixedBF16: launch_ray_sched<kMsgF32, kMsgBF16>(a, s); break;
    default: launch_ray_sched<kMsgF32, kMsgF32>(a, s);
  }
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
