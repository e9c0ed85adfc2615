// Learned warm start of the BP iteration.
//
// After the first BP iteration of a call every touched voxel (one with
// C_c > 0.1 for at least one camera) gets 32 features; a 32-64-64-1 MLP with
// GELU predicts its converged belief log-odds; beliefs and camera messages
// are moved to the prediction and BP continues from there.  Nothing of this
// exists in the reference, whose BP always starts from the prior with all
// M_c = 0 (GVDBInference.cpp:128-134).
//
// Features of voxel v (LF = -LOGODDS_F_EPS = 9.2102404, GVDBCommon.cuh:6):
//   0  log1p(hits): valid pixels of all cameras whose depth endpoint lies in v
//   1  log1p(sum_c C_c(v)): chord length of all rays through v (voxels)
//   2  mean_c M_c(v) / LF over the cameras with C_c(v) > 0.1
//   3  max_c  M_c(v) / LF over the same cameras
//   4  min(dist to the nearest voxel with hits > 0, 4) / 4, Euclidean in
//      voxels, searched over the 7^3 cube around v
//   5..31  clamp(B(v + o) / LF, -3, 3) for o in {-1,0,1}^3, o_x slowest;
//      voxels outside active bricks or the grid count as the prior
//
// Network (parameters FP32, 6337 floats): W1 64x32, b1 64, W2 64x64, b2 64,
// w3 64, b3 1, all row-major.  y = b3 + w3 . gelu(W2 h1 + b2),
// h1 = gelu(W1 x + b1), gelu(z) = z Phi(z).  The two matrix products run on
// v_mfma_f32_16x16x32_bf16 with FP32 accumulation: x, W1, W2 and h1 are
// rounded to BF16 (nearest even), biases, w3 and all element-wise math stay
// FP32.  The backward pass rounds dz2 and dz1 to BF16 for its three products
// (dW2 = dz2 h1^T, dh1 = W2^T dz2, dW1 = dz1 x^T) and keeps the bias and w3
// gradients FP32.
//
// Layouts of the 16x16x32 MFMA (lane l, g = l / 16, c = l % 16): A lane
// holds A[c][8g + e], B lane holds B[8g + e][c], D lane holds D[4g + r][c],
// e in 0..7, r in 0..3.  A forward tile is 16 samples: hidden x samples
// products (D = W x^T), so a lane owns sample c and hidden units
// 16t + 4g + r.  The K index of layer 2 is permuted so that its B operand is
// the lane's own layer-1 output: K slot 32s + 8g + e stands for hidden unit
// pi(s, g, e) = 16 (2s + e/4) + 4g + e%4, and the W2 operand is loaded with
// the same permutation (a sum over K does not depend on its order).
#include "mrf_kernels.h"

namespace mrf {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr float kLF = 9.2102404f;

__device__ __forceinline__ unsigned short bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_round(float x) { return __uint_as_float((uint32_t)bf16_bits(x) << 16); }
__device__ __forceinline__ bf16x8 to_bf16x8(const u16x8& u) { return __builtin_bit_cast(bf16x8, u); }

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// K slot e of layer-2 step s for lane group g -> hidden unit
__device__ __forceinline__ int perm_k(int s, int g, int e) { return 16 * (2 * s + (e >> 2)) + 4 * g + (e & 3); }

__device__ __forceinline__ float shfl_xor(float v, int m) { return __shfl_xor(v, m, 64); }

// Per-lane copies of the network, BF16 fragments for the MFMA operands.
struct Frags {
  bf16x8 w1[4];      // A of layer 1, tile t: W1[16t + c][8g + e]
  bf16x8 w2[4][2];   // A of layer 2, tile t, step s: W2[16t + c][pi(s, g, e)]
  float b1[16], b2[16], w3[16];  // [4t + r] for hidden unit 16t + 4g + r
  float b3;
};

__device__ void load_frags(const float* __restrict__ P, int g, int c, Frags& f) {
  #pragma unroll
  for (int t = 0; t < 4; ++t) {
    u16x8 u;
    #pragma unroll
    for (int e = 0; e < 8; ++e) u[e] = bf16_bits(P[kMlpW1 + (16 * t + c) * kMlpIn + 8 * g + e]);
    f.w1[t] = to_bf16x8(u);
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
      #pragma unroll
      for (int e = 0; e < 8; ++e) u[e] = bf16_bits(P[kMlpW2 + (16 * t + c) * kMlpHid + perm_k(s, g, e)]);
      f.w2[t][s] = to_bf16x8(u);
    }
    #pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int h = 16 * t + 4 * g + r;
      f.b1[4 * t + r] = P[kMlpB1 + h];
      f.b2[4 * t + r] = P[kMlpB2 + h];
      f.w3[4 * t + r] = P[kMlpW3 + h];
    }
  }
  f.b3 = P[kMlpB3];
}

// A operand of dh1 = W2^T dz2 (tile t over h1 units, step s over h2 units):
// W2[pi(s, g, e)][16t + c]
__device__ void load_w2t(const float* __restrict__ P, int g, int c, bf16x8 (&w2t)[4][2]) {
  #pragma unroll
  for (int t = 0; t < 4; ++t)
    #pragma unroll
    for (int s = 0; s < 2; ++s) {
      u16x8 u;
      #pragma unroll
      for (int e = 0; e < 8; ++e) u[e] = bf16_bits(P[kMlpW2 + perm_k(s, g, e) * kMlpHid + 16 * t + c]);
      w2t[t][s] = to_bf16x8(u);
    }
}

// x fragment (B of layer 1) of one sample: X[i][8g .. 8g + 7], zero if !ok
__device__ __forceinline__ u16x8 load_x(const float* __restrict__ X, uint64_t i, bool ok, int g) {
  u16x8 u;
  if (ok) {
    const float4* p = reinterpret_cast<const float4*>(X + i * kMlpIn + 8 * g);
    const float4 a = p[0], b = p[1];
    u[0] = bf16_bits(a.x); u[1] = bf16_bits(a.y); u[2] = bf16_bits(a.z); u[3] = bf16_bits(a.w);
    u[4] = bf16_bits(b.x); u[5] = bf16_bits(b.y); u[6] = bf16_bits(b.z); u[7] = bf16_bits(b.w);
  } else {
    #pragma unroll
    for (int e = 0; e < 8; ++e) u[e] = 0;
  }
  return u;
}

// Forward of one 16-sample tile.  a1, a2: pre-activations (FP32) of the
// lane's hidden units; h1b: layer-2 B operand; h2: gelu(a2); returns y of the
// lane's sample (the same value in the 4 lanes of that sample).
struct Fwd {
  float a1[16], a2[16], h2[16];
  u16x8 h1u[2];  // layer-2 B operand, K slot e of step s = unit 16 (2s + e/4) + 4g + e%4
  float y;
};

__device__ __forceinline__ void forward(const Frags& f, const u16x8& xu, Fwd& o) {
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  const bf16x8 xb = to_bf16x8(xu);
  f32x4 z1[4];
  #pragma unroll
  for (int t = 0; t < 4; ++t) z1[t] = mfma(f.w1[t], xb, zero);
  #pragma unroll
  for (int t = 0; t < 4; ++t)
    #pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float a = z1[t][r] + f.b1[4 * t + r];
      o.a1[4 * t + r] = a;
      // unit 16t + 4g + r is K slot e = 4 (t % 2) + r of step s = t / 2
      o.h1u[t >> 1][4 * (t & 1) + r] = bf16_bits(gelu(a));
    }
  const bf16x8 h0 = to_bf16x8(o.h1u[0]), h1 = to_bf16x8(o.h1u[1]);
  float part = 0.f;
  #pragma unroll
  for (int t = 0; t < 4; ++t) {
    f32x4 z2 = mfma(f.w2[t][0], h0, zero);
    z2 = mfma(f.w2[t][1], h1, z2);
    #pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float a = z2[r] + f.b2[4 * t + r];
      const float h = gelu(a);
      o.a2[4 * t + r] = a;
      o.h2[4 * t + r] = h;
      part += f.w3[4 * t + r] * h;
    }
  }
  part += shfl_xor(part, 16);
  part += shfl_xor(part, 32);
  o.y = part + f.b3;
}

// ---------------------------------------------------------------- inference
__global__ void __launch_bounds__(256) k_mlp_fwd(const float* __restrict__ P, const float* __restrict__ X,
                                                 uint32_t n, float* __restrict__ y) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
  Frags f;
  load_frags(P, g, c, f);
  const uint32_t ntiles = (n + 15) / 16;
  for (uint32_t tile = wave; tile < ntiles; tile += nwaves) {
    const uint32_t i = tile * 16 + c;
    const bool ok = i < n;
    Fwd o;
    forward(f, load_x(X, i, ok, g), o);
    if (g == 0 && ok) y[i] = o.y;
  }
}

// ---------------------------------------------------------------- gradients
// LDS per wave: h1 and dz2 [64][32], dz1 [64][32], x^T [32][32], BF16 bits,
// the sample index fastest.
constexpr int kTileS = 32;
struct alignas(16) WaveLds {
  unsigned short h1[kMlpHid][kTileS];
  unsigned short dz2[kMlpHid][kTileS];
  unsigned short dz1[kMlpHid][kTileS];
  unsigned short xt[kMlpIn][kTileS];
};
constexpr int kGradWaves = 4;

__device__ __forceinline__ bf16x8 lds_frag(const unsigned short* row, int g) {
  const uint4 v = *reinterpret_cast<const uint4*>(row + 8 * g);
  return __builtin_bit_cast(bf16x8, v);
}

// Samples sel[off + i] (or off + i without sel), i < cnt, in tiles of 32;
// loss 1/2 scale sum (y - t)^2, gradients into partial[wave][kMlpParams] and
// lpart[wave] (the loss sum).  Every wave of the grid writes its row.
__global__ void __launch_bounds__(256) k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                  const float* __restrict__ T, const uint32_t* __restrict__ sel,
                                                  uint32_t off, uint32_t cnt, float scale,
                                                  float* __restrict__ partial, double* __restrict__ lpart) {
  __shared__ WaveLds lds_all[kGradWaves];
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  const int wib = threadIdx.x >> 6;
  WaveLds& L = lds_all[wib];
  const uint32_t wave = blockIdx.x * kGradWaves + wib;
  Frags f;
  load_frags(P, g, c, f);
  bf16x8 w2t[4][2];
  load_w2t(P, g, c, w2t);
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  f32x4 gw2[4][4], gw1[4][2];
  #pragma unroll
  for (int i = 0; i < 4; ++i) {
    #pragma unroll
    for (int j = 0; j < 4; ++j) gw2[i][j] = zero;
    #pragma unroll
    for (int j = 0; j < 2; ++j) gw1[i][j] = zero;
  }
  float gb1[16], gb2[16], gw3[16];
  #pragma unroll
  for (int k = 0; k < 16; ++k) gb1[k] = gb2[k] = gw3[k] = 0.f;
  float gb3 = 0.f, loss = 0.f;
  const uint32_t ntiles = (cnt + kTileS - 1) / kTileS;
  // block-uniform trip count: every wave reaches every barrier
  for (uint32_t base = blockIdx.x * kGradWaves; base < ntiles; base += gridDim.x * kGradWaves) {
    const uint32_t tile = base + wib;
    const bool tile_ok = tile < ntiles;
    #pragma unroll
    for (int hf = 0; hf < 2; ++hf) {
      const uint32_t li = tile * kTileS + hf * 16 + c;
      const bool ok = tile_ok && li < cnt;
      const uint64_t i = ok ? (sel ? (uint64_t)sel[off + li] : (uint64_t)off + li) : 0;
      const u16x8 xu = load_x(X, i, ok, g);
      #pragma unroll
      for (int e = 0; e < 8; ++e) L.xt[8 * g + e][hf * 16 + c] = xu[e];
      Fwd o;
      forward(f, xu, o);
      const float dy = ok ? (o.y - T[i]) * scale : 0.f;
      if (ok && g == 0) loss += (o.y - T[i]) * (o.y - T[i]);
      if (g == 0) gb3 += dy;
      // layer 3 and 2: dz2 = w3 dy gelu'(a2)
      u16x8 du[2];
      #pragma unroll
      for (int t = 0; t < 4; ++t)
        #pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = 4 * t + r;
          gw3[k] += dy * o.h2[k];
          const float dz = f.w3[k] * dy * gelu_d(o.a2[k]);
          gb2[k] += dz;
          const unsigned short b = bf16_bits(dz);
          du[t >> 1][4 * (t & 1) + r] = b;
          L.dz2[16 * t + 4 * g + r][hf * 16 + c] = b;
          L.h1[16 * t + 4 * g + r][hf * 16 + c] = o.h1u[t >> 1][4 * (t & 1) + r];
        }
      const bf16x8 db0 = to_bf16x8(du[0]), db1 = to_bf16x8(du[1]);
      // dh1 = W2^T dz2, dz1 = dh1 gelu'(a1)
      #pragma unroll
      for (int t = 0; t < 4; ++t) {
        f32x4 dh = mfma(w2t[t][0], db0, zero);
        dh = mfma(w2t[t][1], db1, dh);
        #pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = 4 * t + r;
          const float dz = dh[r] * gelu_d(o.a1[k]);
          gb1[k] += dz;
          L.dz1[16 * t + 4 * g + r][hf * 16 + c] = bf16_bits(dz);
        }
      }
    }
    __syncthreads();
    if (tile_ok) {
      // dW2[16i + 4g + r][16j + c] += sum_s dz2[16i + .][s] h1[16j + .][s]
      #pragma unroll
      for (int i = 0; i < 4; ++i) {
        const bf16x8 a = lds_frag(L.dz2[16 * i + c], g);
        #pragma unroll
        for (int j = 0; j < 4; ++j) gw2[i][j] = mfma(a, lds_frag(L.h1[16 * j + c], g), gw2[i][j]);
        const bf16x8 a1 = lds_frag(L.dz1[16 * i + c], g);
        #pragma unroll
        for (int j = 0; j < 2; ++j) gw1[i][j] = mfma(a1, lds_frag(L.xt[16 * j + c], g), gw1[i][j]);
      }
    }
    __syncthreads();
  }
  // reduce the per-sample sums over the 16 lanes of a group
  #pragma unroll
  for (int k = 0; k < 16; ++k)
    #pragma unroll
    for (int m = 1; m < 16; m <<= 1) {
      gb1[k] += shfl_xor(gb1[k], m);
      gb2[k] += shfl_xor(gb2[k], m);
      gw3[k] += shfl_xor(gw3[k], m);
    }
  #pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    gb3 += shfl_xor(gb3, m);
    loss += shfl_xor(loss, m);
  }
  float* out = partial + (size_t)wave * kMlpParams;
  #pragma unroll
  for (int i = 0; i < 4; ++i)
    #pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * i + 4 * g + r;
      #pragma unroll
      for (int j = 0; j < 4; ++j) out[kMlpW2 + row * kMlpHid + 16 * j + c] = gw2[i][j][r];
      #pragma unroll
      for (int j = 0; j < 2; ++j) out[kMlpW1 + row * kMlpIn + 16 * j + c] = gw1[i][j][r];
    }
  if (c == 0)
    #pragma unroll
    for (int t = 0; t < 4; ++t)
      #pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int h = 16 * t + 4 * g + r;
        out[kMlpB1 + h] = gb1[4 * t + r];
        out[kMlpB2 + h] = gb2[4 * t + r];
        out[kMlpW3 + h] = gw3[4 * t + r];
      }
  if (lane == 0) {
    out[kMlpB3] = gb3;
    lpart[wave] = 0.5 * (double)loss;
  }
}

// grad[j] = sum over waves in order; loss[0] += sum of the wave losses x scale
__global__ void k_mlp_reduce(const float* __restrict__ partial, const double* __restrict__ lpart, uint32_t nw,
                             float* __restrict__ grad, double* __restrict__ loss, double scale) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < kMlpParams) {
    float s = 0.f;
    for (uint32_t w = 0; w < nw; ++w) s += partial[(size_t)w * kMlpParams + j];
    grad[j] = s;
  }
  if (j == 0 && loss) {
    double s = 0.0;
    for (uint32_t w = 0; w < nw; ++w) s += lpart[w];
    loss[0] += s * scale;
  }
}

__global__ void k_adam(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= kMlpParams) return;
  const float g = G[j];
  const float mj = b1 * m[j] + (1.f - b1) * g;
  const float vj = b2 * v[j] + (1.f - b2) * g * g;
  m[j] = mj;
  v[j] = vj;
  P[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

// ---------------------------------------------------------------- features
__global__ void k_feat_pairs(const uint32_t* __restrict__ pair_vox, const float* __restrict__ pair_c,
                             const float* __restrict__ M, uint32_t npairs, float* __restrict__ pass,
                             float* __restrict__ msum, float* __restrict__ mmax, uint32_t* __restrict__ nc) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  const uint32_t v = pair_vox[q];
  const float C = pair_c[q];
  pass[v] += C;
  if (C > 0.1f) {
    const float m = M[q];
    msum[v] += m;
    mmax[v] = nc[v] ? fmaxf(mmax[v], m) : m;
    nc[v] += 1u;
  }
}

__global__ void k_feat_flags(const uint32_t* __restrict__ nc, size_t n, uint32_t* __restrict__ flag) {
  const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) flag[v] = nc[v] ? 1u : 0u;
}

constexpr int kHalo = 3;                    // distance search radius (voxels)
constexpr int kBrickW = 8;                   // voxels per brick edge
constexpr int kHW = kBrickW + 2 * kHalo;    // 14
constexpr int kBW = kBrickW + 2;            // 10

// One block of 512 threads per brick slot, one thread per voxel.
__global__ void __launch_bounds__(512) k_features(GridGeom gg, const int32_t* __restrict__ slot,
                                                  const int32_t* __restrict__ slot_lin,
                                                  const uint32_t* __restrict__ hits, const float* __restrict__ pass,
                                                  const float* __restrict__ msum, const float* __restrict__ mmax,
                                                  const uint32_t* __restrict__ nc, const float* __restrict__ B,
                                                  float prior, const uint32_t* __restrict__ row,
                                                  float* __restrict__ X, uint32_t* __restrict__ vox) {
  __shared__ unsigned short hl[kHW * kHW * kHW];
  __shared__ float bl[kBW * kBW * kBW];
  const int s = blockIdx.x;
  const int lin = slot_lin[s];
  const int ox = (lin / (gg.by * gg.bz)) * kBrickW, oy = ((lin / gg.bz) % gg.by) * kBrickW, oz = (lin % gg.bz) * kBrickW;
  for (int idx = threadIdx.x; idx < kHW * kHW * kHW; idx += blockDim.x) {
    const int x = ox - kHalo + idx / (kHW * kHW), y = oy - kHalo + (idx / kHW) % kHW, z = oz - kHalo + idx % kHW;
    unsigned short h = 0;
    if (x >= 0 && y >= 0 && z >= 0 && x < gg.nx && y < gg.ny && z < gg.nz) {
      const int sl = slot[((x >> 3) * gg.by + (y >> 3)) * gg.bz + (z >> 3)];
      if (sl >= 0) {
        const uint32_t hv = hits[(size_t)sl * kBrickVox + ((x & 7) * 64 + (y & 7) * 8 + (z & 7))];
        h = (unsigned short)(hv < 65535u ? hv : 65535u);
      }
    }
    hl[idx] = h;
  }
  for (int idx = threadIdx.x; idx < kBW * kBW * kBW; idx += blockDim.x) {
    const int x = ox - 1 + idx / (kBW * kBW), y = oy - 1 + (idx / kBW) % kBW, z = oz - 1 + idx % kBW;
    float b = prior;
    if (x >= 0 && y >= 0 && z >= 0 && x < gg.nx && y < gg.ny && z < gg.nz) {
      const int sl = slot[((x >> 3) * gg.by + (y >> 3)) * gg.bz + (z >> 3)];
      if (sl >= 0) b = B[(size_t)sl * kBrickVox + ((x & 7) * 64 + (y & 7) * 8 + (z & 7))];
    }
    bl[idx] = b;
  }
  __syncthreads();
  const int t = threadIdx.x;
  const size_t v = (size_t)s * kBrickVox + t;
  const uint32_t n = nc[v];
  if (n == 0) return;
  const int i = t >> 6, j = (t >> 3) & 7, k = t & 7;
  float* xr = X + (size_t)row[v] * kMlpIn;
  const uint32_t hv = hits[v];
  xr[0] = log1pf((float)hv);
  xr[1] = log1pf(pass[v]);
  xr[2] = (msum[v] / (float)n) / kLF;
  xr[3] = mmax[v] / kLF;
  int best = 16;
  for (int a = -kHalo; a <= kHalo; ++a)
    for (int b = -kHalo; b <= kHalo; ++b)
      for (int cc = -kHalo; cc <= kHalo; ++cc) {
        const int d2 = a * a + b * b + cc * cc;
        if (d2 < best && hl[((i + kHalo + a) * kHW + (j + kHalo + b)) * kHW + (k + kHalo + cc)] != 0) best = d2;
      }
  xr[4] = sqrtf((float)best) * 0.25f;
  int f = 5;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int cc = 0; cc < 3; ++cc) {
        const float q = bl[((i + a) * kBW + (j + b)) * kBW + (k + cc)] / kLF;
        xr[f++] = fminf(fmaxf(q, -3.f), 3.f);
      }
  vox[row[v]] = (uint32_t)v;
}

// B <- prediction (kept at the first-iteration value where both are
// saturated on the same side and the first iteration is stronger), and the
// shift spread evenly over the voxel's cameras: delta = (B' - B) / n_c.
__global__ void k_warm_apply(const uint32_t* __restrict__ vox, const float* __restrict__ pred, uint32_t n,
                             const uint32_t* __restrict__ nc, float* __restrict__ B, float* __restrict__ delta) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = vox[i];
  const float b = B[v];
  const float p = pred[i];
  float bn = p;
  if ((p <= -kLF && b <= p) || (p >= kLF && b >= p)) bn = b;
  delta[v] = (bn - b) / (float)nc[v];
  B[v] = bn;
}

__global__ void k_warm_pairs(const uint32_t* __restrict__ pair_vox, const float* __restrict__ pair_c,
                             uint32_t npairs, const float* __restrict__ delta, float* __restrict__ M) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs) return;
  if (pair_c[q] > 0.1f) M[q] += delta[pair_vox[q]];
}

inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

}  // namespace

uint32_t mlp_grad_waves(uint32_t cnt) {
  const uint32_t ntiles = (cnt + kTileS - 1) / kTileS;
  uint32_t nb = (ntiles + kGradWaves - 1) / kGradWaves;
  if (nb > 512) nb = 512;
  if (nb == 0) nb = 1;
  return nb * kGradWaves;
}

void launch_mlp_fwd(const float* P, const float* X, uint32_t n, float* y, hipStream_t s) {
  if (n == 0) return;
  unsigned nb = blocks((n + 15) / 16, 4);
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL(k_mlp_fwd, dim3(nb), dim3(256), 0, s, P, X, n, y);
}

void launch_mlp_grad(const float* P, const float* X, const float* T, const uint32_t* sel, uint32_t off,
                     uint32_t cnt, float scale, float* partial, double* lpart, hipStream_t s) {
  const uint32_t nw = mlp_grad_waves(cnt);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nw / kGradWaves), dim3(64 * kGradWaves), 0, s, P, X, T, sel, off, cnt, scale,
                     partial, lpart);
}

void launch_mlp_reduce(const float* partial, const double* lpart, uint32_t nw, float* grad, double* loss,
                       double scale, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_reduce, dim3(blocks(kMlpParams, 256)), dim3(256), 0, s, partial, lpart, nw, grad, loss,
                     scale);
}

void launch_adam(float* P, const float* G, float* m, float* v, float lr, float b1, float b2, float eps, float c1,
                 float c2, hipStream_t s) {
  hipLaunchKernelGGL(k_adam, dim3(blocks(kMlpParams, 256)), dim3(256), 0, s, P, G, m, v, lr, b1, b2, eps, c1, c2);
}

void launch_feat_pairs(const uint32_t* pair_vox, const float* pair_c, const float* M, uint32_t npairs, float* pass,
                       float* msum, float* mmax, uint32_t* nc, hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_feat_pairs, dim3(blocks(npairs, 256)), dim3(256), 0, s, pair_vox, pair_c, M, npairs, pass,
                     msum, mmax, nc);
}

void launch_feat_flags(const uint32_t* nc, size_t n, uint32_t* flag, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_feat_flags, dim3(blocks(n, 256)), dim3(256), 0, s, nc, n, flag);
}

void launch_features(const GridGeom& g, const int32_t* slot, const int32_t* slot_lin, uint32_t nslots,
                     const uint32_t* hits, const float* pass, const float* msum, const float* mmax, const uint32_t* nc,
                     const float* B, float prior, const uint32_t* row, float* X, uint32_t* vox, hipStream_t s) {
  if (nslots == 0) return;
  hipLaunchKernelGGL(k_features, dim3(nslots), dim3(kBrickVox), 0, s, g, slot, slot_lin, hits, pass, msum, mmax, nc,
                     B, prior, row, X, vox);
}

void launch_warm_apply(const uint32_t* vox, const float* pred, uint32_t n, const uint32_t* nc, float* B,
                       float* delta, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(k_warm_apply, dim3(blocks(n, 256)), dim3(256), 0, s, vox, pred, n, nc, B, delta);
}

void launch_warm_pairs(const uint32_t* pair_vox, const float* pair_c, uint32_t npairs, const float* delta, float* M,
                       hipStream_t s) {
  if (npairs == 0) return;
  hipLaunchKernelGGL(k_warm_pairs, dim3(blocks(npairs, 256)), dim3(256), 0, s, pair_vox, pair_c, npairs, delta, M);
}

}  // namespace mrf
