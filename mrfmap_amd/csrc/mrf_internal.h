// Internal types shared by the HIP kernels and the host orchestration of
// libmrfmap_hip.  Nothing here crosses the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mrf {

// Sensor-model polynomial in FP32 (sigma(z) or bias(z), c2 = 0 for two
// coefficients).
struct Poly {
  float c0, c1, c2;
  __device__ __forceinline__ float operator()(float z) const { return fmaf(z, fmaf(z, c2, c1), c0); }
};

// Kahan-compensated running sum along a ray.
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

// Per-lane sensor model of a ray: lookup cell coefficients (get_depth_sigma /
// get_depth_bias, gvdbkernels.cu:46-55).
__device__ __forceinline__ void load_cell(const float* __restrict__ sigt, const float* __restrict__ biat, int lookup_n,
                                          uint32_t cell, Poly& sig, Poly& bia) {
  sig.c0 = sigt[cell];
  sig.c1 = sigt[cell + 1];
  sig.c2 = lookup_n == 3 ? sigt[cell + 2] : 0.0f;
  bia.c0 = biat[cell];
  bia.c1 = biat[cell + 1];
  bia.c2 = lookup_n == 3 ? biat[cell + 2] : 0.0f;
}

// Constants of the live reference path (include/mrfmap/GVDBCommon.cuh:3-11).
constexpr float kFEps = 1e-4f;         // F_EPS: probability clamp
constexpr float kZMax = 10.0f;         // Z_MAX
constexpr int kLookupPx = 20;          // LOOKUP_PX: sensor-model lookup cell (pixels)
constexpr int kLookupMax = 3096;       // LOOKUP_NUM
// epilogue gate on a camera's summed chord length, C > 0.1f in float
// (gvdbkernels.cu:408), and the same gate applied to the exact sum
constexpr float kCumLenMin = 0.1f;
constexpr double kCumLenGate = (double)kCumLenMin;
constexpr float kMsgNormEps = 1e-7f;   // 0.001 * F_EPS in the message normaliser

// Fixed-point scale of the integer DDA: origins and directions are kept in
// units of 2^-kFix voxels.
constexpr int kFix = 16;
constexpr double kFixScale = 65536.0;

constexpr int kBrickLog = 3;           // 8^3 voxels per brick
constexpr int kBrickVox = 512;
constexpr uint32_t kGapVox = 0xFFFFFFFFu;  // incidence record that only advances z
constexpr uint32_t kTile = 512;           // positions per wavefront in the segment kernels (8 x 64)

constexpr uint32_t kSlice = 64;           // rays per slice (one 8x8 pixel tile, one per lane)

// Per-ray records of one camera's traversal (one per lane of a slice).  A
// ray's records are split into a front part, where the depth likelihood
// psi(s) is exactly 0 in FP32, and a window part; each part is a sliced-ELL
// region of its own: front record k at FrontRec::off + k * kSlice, window
// record k (record nf + k of the ray) at RayRec::off_w + k * kSlice.
struct FrontRec {
  uint32_t off;        // first front record
  uint32_t nf;         // front records (voxels + gap records)
};
struct RayRec {
  uint32_t off_w;      // first window record
  uint32_t nw;         // window records
  float z0;            // planar depth at the entry of the first record (m)
  float inv_norm;      // metres of planar depth per voxel of chord length
  float depth;         // measured depth d (m)
  uint32_t cell_flags; // sensor-model lookup offset (bits 0..29) | kInfTerm | kRanOut
  float ks, kc;        // Kahan sum of the front chord lengths (sum, compensation)
};
constexpr uint32_t kChunk = 16;  // front records per lane in one chunk of the front kernel

// A record belongs to the front while d - (s + bias(s)) > kFrontU * 3 sigma(s)
// at its exit depth s: then psi(s) = 0.398942 exp(-u^2/2) / (3 sigma) has
// u > 13.5, exp(-u^2/2) < 2.7e-40 is an FP32 subnormal and psi < 1e-37 for
// any sigma above 1 mm.  The record's message is the clamp floor F_EPS * l
// (mu1 = P + V psi lies 30 orders of magnitude below the normaliser's 1e-7),
// and its tau, left out of T and of the later records' P, is below the FP32
// rounding of any T that can lift a message above the floor.  With the
// reference's flush-to-zero arithmetic such a tau is exactly 0 (its factor
// 0.398942 exp(-u^2/2) is below 2^-126 from u = 13.15 on).
constexpr float kFrontU = 13.5f;
constexpr uint32_t kRanOut = 0x80000000u;  // ray left the active space before truncation
// the ray ran out and its depth lies beyond the exit: the infinity term applies
// (gvdbkernels.cu:1196-1208), decided in FP64 from the exact exit depth
constexpr uint32_t kInfTerm = 0x40000000u;
constexpr uint32_t kCellMask = 0x3FFFFFFFu;

// Camera geometry in the form the kernels consume (all double so that the
// host oracle and the device agree bit for bit on the fixed-point setup).
struct CamGeom {
  double R[9];      // camera-to-world rotation, row-major
  double t[3];      // camera centre (m)
  double fx, fy, cx, cy;
  int rows, cols;
  int lookup_u;     // ceil(cols / 20)
  int mask848;      // apply the 848x480 border mask
};

struct GridGeom {
  int nx, ny, nz;          // voxel grid (multiple of 8)
  int mx, my, mz;          // map extent in voxels (endpoint gating)
  int bx, by, bz;          // brick grid
  double res;              // voxel size (m)
  double off[3];           // world->voxel offset (voxels)
  int lookup_n;            // 2 or 3 coefficients per cell
};

}  // namespace mrf
