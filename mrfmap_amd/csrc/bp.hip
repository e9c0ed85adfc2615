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
template <>
struct Msg<kMsgF32> {
  typedef float T;
  __device__ static float load(const T* p, uint32_t i) { return p[i]; }
  __device__ static void store(T* p, uint32_t i, float v) { p[i] = v; }
};
template <>
struct Msg<kMsgF16> {
  typedef __half T;
  __device__ static float load(const T* p, uint32_t i) { return __half2float(p[i]); }
  __device__ static void store(T* p, uint32_t i, float v) { p[i] = __float2half_rn(v); }
};
template <>
struct Msg<kMsgBF16> {
  typedef uint16_t T;
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
  const float inv = 1.0f / (3.0f * sig(s));
  const float u = diff * inv;
  return 0.398942f * __expf(-0.5f * u * u) * inv;
}

struct VoxTerm {
  float n;    // clamped P(o = 0) message from the voxel
  float a;    // l * ln n
};

template <int MT>
__device__ __forceinline__ VoxTerm voxel_term(const float* __restrict__ B, const typename Msg<MT>::T* __restrict__ M,
                                              uint32_t v, float len, bool live) {
  VoxTerm t;
  t.n = 1.0f;
  t.a = 0.0f;
  if (live) {
    const float q = B[v] - Msg<MT>::load(M, v);
    float nn = 1.0f / (1.0f + __expf(q));
    nn = fminf(fmaxf(nn, kFEps), 1.0f - kFEps);
    t.n = nn;
    t.a = len * __logf(nn);
  }
  return t;
}

template <int MT>
__global__ __launch_bounds__(256) void k_ray_messages(const RayRec* __restrict__ rays, uint32_t nrays,
                                                      const uint2* __restrict__ inc, const float* __restrict__ B,
                                                      const typename Msg<MT>::T* __restrict__ M,
                                                      float* __restrict__ S, const float* __restrict__ sigt,
                                                      const float* __restrict__ biat, int lookup_n) {
  const int lane = threadIdx.x & 63;
  const uint32_t r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= nrays) return;
  const RayRec rec = rays[r];
  const uint32_t cell = rec.cell_flags & 0x7FFFFFFFu;
  Poly sig, bia;
  sig.n = bia.n = lookup_n;
  sig.c0 = sigt[cell];
  sig.c1 = sigt[cell + 1];
  sig.c2 = lookup_n == 3 ? sigt[cell + 2] : 0.0f;
  bia.c0 = biat[cell];
  bia.c1 = biat[cell + 1];
  bia.c2 = lookup_n == 3 ? biat[cell + 2] : 0.0f;
  const float d = rec.depth;
  const uint2* __restrict__ list = inc + rec.off;
  const uint32_t n = rec.n;

  // Pass 1: Lambda, tau and T.
  float cL = 0.0f, cS = 0.0f, T = 0.0f;
  for (uint32_t base = 0; base < n; base += 64) {
    const uint32_t i = base + lane;
    const bool in = i < n;
    const uint2 e = in ? list[i] : make_uint2(kGapVox, 0u);
    const float len = __uint_as_float(e.y);
    const bool live = e.x != kGapVox;
    const VoxTerm vt = voxel_term<MT>(B, M, e.x, len, live);
    const float L = cL + wave_scan(vt.a);
    const float sl = cS + wave_scan(len);
    const float s = rec.z0 + sl * rec.inv_norm;
    const float tau = live ? __expf(L) * (1.0f - vt.n) * depth_psi(s, d, sig, bia) : 0.0f;
    T += lane63(wave_scan(tau));
    cL = lane63(L);
    cS = lane63(sl);
  }
  if (rec.cell_flags & kRanOut) {
    const float sn = rec.z0 + cS * rec.inv_norm;
    if (d > sn + bia(sn)) T += __expf(cL);
  }

  // Pass 2: outgoing messages p_i and the scatter of p_i * l_i.
  cL = 0.0f;
  cS = 0.0f;
  float cP = 0.0f;
  for (uint32_t base = 0; base < n; base += 64) {
    const uint32_t i = base + lane;
    const bool in = i < n;
    const uint2 e = in ? list[i] : make_uint2(kGapVox, 0u);
    const float len = __uint_as_float(e.y);
    const bool live = e.x != kGapVox;
    const VoxTerm vt = voxel_term<MT>(B, M, e.x, len, live);
    const float L = cL + wave_scan(vt.a);
    const float sl = cS + wave_scan(len);
    const float s = rec.z0 + sl * rec.inv_norm;
    const float vpsi = live ? __expf(L) * depth_psi(s, d, sig, bia) : 0.0f;
    const float tau = vpsi * (1.0f - vt.n);
    const float Pin = wave_scan(tau);
    const float P = cP + (Pin - tau);
    const float R = (T - P) - tau;
    if (live) {
      const float mu1 = P + vpsi;
      const float mu0 = P + R / vt.n;
      float p = mu1 / (mu1 + mu0 + kMsgNormEps);
      p = fminf(fmaxf(p, kFEps), 1.0f - kFEps);
      unsafeAtomicAdd(S + e.x, p * len);
    }
    cL = lane63(L);
    cS = lane63(sl);
    cP += lane63(Pin);
  }
}

__global__ void k_cum_len(const RayRec* __restrict__ rays, uint32_t nrays, const uint2* __restrict__ inc,
                          unsigned long long* __restrict__ cfix) {
  const int lane = threadIdx.x & 63;
  const uint32_t r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (r >= nrays) return;
  const RayRec rec = rays[r];
  for (uint32_t i = lane; i < rec.n; i += 64) {
    const uint2 e = inc[rec.off + i];
    if (e.x == kGapVox) continue;
    const double l = (double)__uint_as_float(e.y) * 4294967296.0;
    atomicAdd(cfix + e.x, (unsigned long long)__double2ll_rn(l));
  }
}

constexpr int kPairThreads = 256;
constexpr int kPairItems = 8;
constexpr int kPairPerBlock = kPairThreads * kPairItems;

__device__ uint32_t block_excl_u32(uint32_t v, uint32_t* sh, uint32_t* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kPairThreads; off <<= 1) {
    uint32_t add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  const uint32_t incl = sh[t];
  *total = sh[kPairThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kPairThreads) void k_pair_count(const unsigned long long* __restrict__ cfix,
                                                              size_t nvox, uint32_t* __restrict__ block_cnt) {
  __shared__ uint32_t sh[kPairThreads];
  const size_t base = (size_t)blockIdx.x * kPairPerBlock + (size_t)threadIdx.x * kPairItems;
  uint32_t c = 0;
  for (int k = 0; k < kPairItems; ++k) {
    const size_t v = base + k;
    if (v < nvox && cfix[v] != 0ull) ++c;
  }
  uint32_t tot;
  block_excl_u32(c, sh, &tot);
  if (threadIdx.x == 0) block_cnt[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kPairThreads) void k_pair_write(unsigned long long* __restrict__ cfix, size_t nvox,
                                                              const uint32_t* __restrict__ block_off,
                                                              uint32_t* __restrict__ pair_vox,
                                                              float* __restrict__ pair_c) {
  __shared__ uint32_t sh[kPairThreads];
  const size_t base = (size_t)blockIdx.x * kPairPerBlock + (size_t)threadIdx.x * kPairItems;
  unsigned long long vals[kPairItems];
  uint32_t c = 0;
  for (int k = 0; k < kPairItems; ++k) {
    const size_t v = base + k;
    vals[k] = v < nvox ? cfix[v] : 0ull;
    if (vals[k] != 0ull) ++c;
  }
  uint32_t tot;
  uint32_t pos = block_off[blockIdx.x] + block_excl_u32(c, sh, &tot);
  for (int k = 0; k < kPairItems; ++k) {
    if (vals[k] != 0ull) {
      pair_vox[pos] = (uint32_t)(base + k);
      pair_c[pos] = (float)((double)vals[k] * (1.0 / 4294967296.0));
      ++pos;
      cfix[base + k] = 0ull;
    }
  }
}

template <int MT>
__global__ void k_epilogue(const uint32_t* __restrict__ pair_vox, const float* __restrict__ pair_c,
                           uint32_t npairs, float* __restrict__ B, typename Msg<MT>::T* __restrict__ M,
                           float* __restrict__ S, uint8_t* __restrict__ upd) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  const uint32_t v = pair_vox[i];
  const float C = pair_c[i];
  const float s = S[v];
  S[v] = 0.0f;
  if (C > kCumLenMin) {
    float a = s / C;
    a = fminf(fmaxf(a, kFEps), 1.0f - kFEps);
    const float m = __logf(a) - __logf(1.0f - a);
    const float mo = Msg<MT>::load(M, v);
    Msg<MT>::store(M, v, m);
    B[v] = (B[v] + m) - mo;
    upd[v] = 1;
  }
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

void launch_ray_messages(const MsgLaunch& a, hipStream_t s) {
  if (a.nrays == 0) return;
  const int waves = 4;
  dim3 grid((a.nrays + waves - 1) / waves), block(64 * waves);
  switch (a.msg_type) {
    case kMsgF16:
      hipLaunchKernelGGL(k_ray_messages<kMsgF16>, grid, block, 0, s, a.rays, a.nrays, a.inc, a.B,
                         (const __half*)a.M, a.S, a.sig, a.bia, a.lookup_n);
      break;
    case kMsgBF16:
      hipLaunchKernelGGL(k_ray_messages<kMsgBF16>, grid, block, 0, s, a.rays, a.nrays, a.inc, a.B,
                         (const uint16_t*)a.M, a.S, a.sig, a.bia, a.lookup_n);
      break;
    default:
      hipLaunchKernelGGL(k_ray_messages<kMsgF32>, grid, block, 0, s, a.rays, a.nrays, a.inc, a.B,
                         (const float*)a.M, a.S, a.sig, a.bia, a.lookup_n);
  }
}

void launch_cum_len(const RayRec* rays, uint32_t nrays, const uint2* inc, unsigned long long* cfix,
                    hipStream_t s) {
  if (nrays == 0) return;
  hipLaunchKernelGGL(k_cum_len, dim3((nrays + 3) / 4), dim3(256), 0, s, rays, nrays, inc, cfix);
}

size_t pair_blocks(size_t nvox) { return (nvox + kPairPerBlock - 1) / kPairPerBlock; }

void launch_build_pairs(unsigned long long* cfix, size_t nvox, uint32_t* block_cnt, uint32_t* block_off,
                        uint64_t* scan_tmp, uint64_t* d_total, uint32_t* pair_vox, float* pair_c,
                        hipStream_t s) {
  const size_t nb = pair_blocks(nvox);
  if (nb == 0) return;
  hipLaunchKernelGGL(k_pair_count, dim3((unsigned)nb), dim3(kPairThreads), 0, s, cfix, nvox, block_cnt);
  launch_exclusive_scan(block_cnt, block_off, nb, scan_tmp, d_total, s);
  if (pair_vox != nullptr)
    hipLaunchKernelGGL(k_pair_write, dim3((unsigned)nb), dim3(kPairThreads), 0, s, cfix, nvox, block_off,
                       pair_vox, pair_c);
}

void launch_epilogue(const uint32_t* pair_vox, const float* pair_c, uint32_t npairs, float* B, void* M,
                     int msg_type, float* S, uint8_t* upd, hipStream_t s) {
  if (npairs == 0) return;
  dim3 grid(grid_for(npairs, 256)), block(256);
  switch (msg_type) {
    case kMsgF16:
      hipLaunchKernelGGL(k_epilogue<kMsgF16>, grid, block, 0, s, pair_vox, pair_c, npairs, B, (__half*)M, S, upd);
      break;
    case kMsgBF16:
      hipLaunchKernelGGL(k_epilogue<kMsgBF16>, grid, block, 0, s, pair_vox, pair_c, npairs, B, (uint16_t*)M, S,
                         upd);
      break;
    default:
      hipLaunchKernelGGL(k_epilogue<kMsgF32>, grid, block, 0, s, pair_vox, pair_c, npairs, B, (float*)M, S, upd);
  }
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
