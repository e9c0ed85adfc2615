// Ray traversal for the BP inference path: pixel ray set-up, integer 3-D DDA
// over the brick grid, brick activation and emission of per-ray incidence
// lists.
//
// Replaces GVDB's hierarchical DDA `rayCast` + PREPARE_DDA_LEAF / NEXT_DDA /
// STEP_DDA as driven by render_cumulative_length_kernel and
// render_factors_to_node_kernel (src/cuda/gvdbkernels.cu:1164-1226,
// 1276-1302), and the endpoint activation of GVDBCamera::reprojectAndActivate
// -> GVDBMRFMap::activateSpace (src/GVDBCamera.cpp:70-112,
// src/GVDBMRFMap.cpp:117-135).
//
// Traversal rule (the contract the CPU oracle restates):
//  * ray of pixel (x, y): origin O = t/res + off, direction per metre of planar
//    depth D = R ((x-cx)/fx, (y-cy)/fy, 1) / res, both in voxel units and
//    rounded to fixed point Oi = rint(O 2^16), Di = rint(D 2^16);
//  * the ray is walked from max(0, box entry) through the voxel grid; every
//    voxel boundary crossing is the exact rational z = (k 2^16 - Oi_a) / Di_a,
//    compared by 64-bit cross multiplication, ties step all tied axes at once;
//  * a voxel is emitted when its brick is active, with chord length
//    l = (z_exit - z_entry) |D| evaluated in IEEE double from the rationals;
//    a run of skipped voxels after the first emission becomes one gap record;
//  * at the exit of every active brick the walk stops if
//    z + bias(z) >= d + 3 sigma(z) (src/cuda/gvdbkernels.cu:567-570, 1035-1037);
//    leaving the grid first marks the ray kRanOut (the "infinity" term,
//    src/cuda/gvdbkernels.cu:1196-1208).
// The file is compiled with -ffp-contract=off so each double operation is a
// single correctly rounded IEEE operation, as in the host oracle.

#include "mrf_internal.h"
#include "mrf_kernels.h"

namespace mrf {

namespace {

__host__ __device__ __forceinline__ int tiled_pixels(const CamGeom& c) {
  return ((c.cols + 7) >> 3) * ((c.rows + 7) >> 3) * 64;
}

struct Rat {
  int64_t n, d;  // value n / d with d > 0
};

__device__ __forceinline__ bool rat_lt(const Rat& a, const Rat& b) { return a.n * b.d < b.n * a.d; }

__device__ __forceinline__ int64_t floor_div(int64_t x, int64_t y, bool* exact) {
  int64_t q = x / y, r = x % y;
  if (r != 0 && x < 0) q -= 1;
  *exact = (r == 0);
  return q;
}

struct RaySetup {
  int64_t Oi[3], Di[3];
  double norm;  // |D| in voxels per metre of planar depth
};

// Returns false for rays whose fixed-point set-up would leave the exact range
// of the 64-bit comparisons (|Oi| >= 2^29 or |Di| >= 2^26).
__device__ bool setup_ray(const CamGeom& c, const GridGeom& g, int x, int y, RaySetup& rs) {
  double ux = ((double)x - c.cx) / c.fx;
  double uy = ((double)y - c.cy) / c.fy;
  double n2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    double dw = c.R[a * 3 + 0] * ux;
    dw = dw + c.R[a * 3 + 1] * uy;
    dw = dw + c.R[a * 3 + 2];
    double D = dw / g.res;
    double O = c.t[a] / g.res;
    O = O + g.off[a];
    double oi = rint(O * kFixScale), di = rint(D * kFixScale);
    if (!(fabs(oi) < 536870912.0) || !(fabs(di) < 67108864.0)) return false;
    rs.Oi[a] = (int64_t)oi;
    rs.Di[a] = (int64_t)di;
    double dd = (double)rs.Di[a] / kFixScale;
    n2 = n2 + dd * dd;
  }
  rs.norm = sqrt(n2);
  return rs.norm > 0.0;
}

__device__ __forceinline__ double poly_eval(const float* c, int n, double s) {
  if (n == 3) {
    double p = (s * s) * (double)c[2];
    p = p + s * (double)c[1];
    return p + (double)c[0];
  }
  double p = s * (double)c[1];
  return p + (double)c[0];
}

__device__ __forceinline__ bool pixel_valid(const CamGeom& c, int x, int y, float d) {
  if (!(d > 0.0f)) return false;  // NaN, 0 and negatives are invalid
  if (c.mask848 && c.cols == 848 && c.rows == 480 &&
      (x < 24 || x >= c.cols - 24 || y < 20 || y >= c.rows - 20))
    return false;
  return true;
}

__device__ __forceinline__ uint32_t lookup_offset(const CamGeom& c, int lookup_n, int x, int y) {
  int v = c.rows - 1 - y;  // the reference indexes its tables in the vertically flipped image
  return (uint32_t)((v / kLookupPx) * (c.lookup_u * lookup_n) + (x / kLookupPx) * lookup_n);
}

enum WalkMode { kCount = 1, kEmit = 2, kVisit = 3 };

struct WalkOut {
  uint32_t n;
  float z0;
  bool ran_out;
  double last_exit;  // exact exit depth of the last voxel record (kEmit, kVisit)
};

// Entry of a pixel ray into the voxel grid: the exact entry depth zin (at
// least 0), the voxel there and the step sign of each axis; false when the
// ray misses the grid.
__device__ __forceinline__ bool ray_entry(const GridGeom& g, const RaySetup& rs, Rat& zin, int ix[3], int sg[3]) {
  const int N[3] = {g.nx, g.ny, g.nz};
  zin = {0, 1};
  Rat zout = {0, 1};
  bool has_out = false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t Oi = rs.Oi[a], Di = rs.Di[a], top = (int64_t)N[a] << kFix;
    if (Di == 0) {
      if (Oi < 0 || Oi >= top) return false;
      continue;
    }
    Rat lo, hi;
    if (Di > 0) {
      lo = {-Oi, Di};
      hi = {top - Oi, Di};
    } else {
      lo = {Oi - top, -Di};
      hi = {Oi, -Di};
    }
    if (rat_lt(zin, lo)) zin = lo;
    if (!has_out || rat_lt(hi, zout)) {
      zout = hi;
      has_out = true;
    }
  }
  if (!has_out || !rat_lt(zin, zout)) return false;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t Oi = rs.Oi[a], Di = rs.Di[a];
    bool exact;
    int64_t X = Oi * zin.d + zin.n * Di;
    int64_t q = floor_div(X, zin.d << kFix, &exact);
    if (Di < 0 && exact) q -= 1;
    if (q < 0) q = 0;
    if (q > N[a] - 1) q = N[a] - 1;
    ix[a] = (int)q;
    sg[a] = Di > 0 ? 1 : (Di < 0 ? -1 : 0);
  }
  return true;
}

// Front/window split of one ray during its emission, with the ray kernel's
// own FP32 depth sums: the front ends at the first record whose exit depth s
// has d - (s + bias(s)) <= kFrontU * 3 sigma(s).  The Kahan state of the
// front's length sum is kept: the window continues it.
struct FrontSplit {
  Poly sig, bia;
  float z0, inv_norm, depth;
  Kahan kS;
  uint32_t nf;  // the first window record (valid once open is false)
  bool open;
  __device__ __forceinline__ void record(uint32_t k, float ln) {
    if (!open) return;
    const float s = z0 + kS.at(ln) * inv_norm;
    const float sg = sig(s);
    if (!(sg > 0.0f && depth - (s + bia(s)) > kFrontU * (3.0f * sg))) {
      nf = k;
      open = false;
      return;
    }
    kS.add(ln);
  }
};

// Records of one ray are written out[0], out[kSlice], out[2 kSlice], ...
// (record k of the 64 rays of a pixel tile is contiguous: sliced ELL), the
// first cap of them (the count goes on past cap); vis, when given, splits
// them into front and window (vis->record(k, chord) per record in order,
// vis->z0 the planar depth at the entry of the first).
// kVisit walks to the end of the grid without truncation and hands every
// voxel of an active brick to vis->voxel(id, ix, chord, z entry, z exit) and
// every active brick's exit depth to vis->brick_end(z); a true return from
// either ends the walk.
template <int MODE, class V = int>
__device__ WalkOut walk(const GridGeom& g, const RaySetup& rs, double depth, const float* sig,
                        const float* bia, const int32_t* __restrict__ slot, uint2* __restrict__ out,
                        uint32_t cap, V* vis = nullptr) {
  WalkOut wo = {0u, 0.0f, true};
  const int N[3] = {g.nx, g.ny, g.nz};
  Rat zin;
  int ix[3], sg[3];
  if (!ray_entry(g, rs, zin, ix, sg)) return wo;
  Rat t[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t Oi = rs.Oi[a], Di = rs.Di[a];
    if (Di > 0)
      t[a] = {(((int64_t)ix[a] + 1) << kFix) - Oi, Di};
    else if (Di < 0)
      t[a] = {Oi - ((int64_t)ix[a] << kFix), -Di};
    else
      t[a] = {1, 0};
  }

  const int lookup_n = g.lookup_n;
  Rat zc = zin;
  double zc_d = (double)zc.n / (double)zc.d;
  double last_exit = 0.0;
  bool any = false, skipped = false;
  uint32_t cnt = 0;
  int32_t bslot = 0;
  auto brick_lookup = [&](void) -> int32_t {
    return slot[((ix[0] >> kBrickLog) * g.by + (ix[1] >> kBrickLog)) * g.bz + (ix[2] >> kBrickLog)];
  };
  bslot = brick_lookup();
  // The next crossing of each moving axis is t[a] = n_a / d_a with d_a fixed
  // and n_a growing by 2^kFix per step, so the cross differences
  // X_ab = n_a d_b - n_b d_a (t_a < t_b iff X_ab < 0, equal iff 0) are kept
  // and updated by additions: the same exact comparisons as rat_lt, with no
  // multiplications and no run-time indexing in the loop.
  const bool m0 = sg[0] != 0, m1 = sg[1] != 0, m2 = sg[2] != 0;
  const int64_t s0 = t[0].d << kFix, s1 = t[1].d << kFix, s2 = t[2].d << kFix;
  int64_t n0 = t[0].n, n1 = t[1].n, n2 = t[2].n;
  const int64_t d0 = t[0].d, d1 = t[1].d, d2 = t[2].d;
  const double dd0 = (double)d0, dd1 = (double)d1, dd2 = (double)d2;  // exact (below 2^53)
  int64_t X01 = n0 * d1 - n1 * d0, X02 = n0 * d2 - n2 * d0, X12 = n1 * d2 - n2 * d1;
  const int max_steps = g.nx + g.ny + g.nz + 3;
  for (int step = 0; step < max_steps; ++step) {
    // the nearest crossing; the lowest axis among equals (strict
    // comparisons), in selects rather than branches: axis 1 beats axis 0
    // (or axis 0 does not move), axis 2 beats the better of the two
    const bool b1 = m1 & (!m0 | (X01 > 0));
    const bool b2 = m2 & (b1 ? (X12 > 0) : (!m0 | (X02 > 0)));
    const int best = b2 ? 2 : (b1 ? 1 : 0);
    const Rat zn = {b2 ? n2 : (b1 ? n1 : n0), b2 ? d2 : (b1 ? d1 : d0)};
    // per-voxel depths only where records or visits need them; the counting
    // walk converts the exit depth at brick exits alone
    constexpr bool kDepths = MODE == kEmit || MODE == kVisit;
    double zn_d = 0.0;
    if constexpr (kDepths) zn_d = (double)zn.n / (b2 ? dd2 : (b1 ? dd1 : dd0));
    const bool active = bslot >= 0;
    if (active) {
      if (any && skipped) {
        if constexpr (MODE == kEmit) {
          float gl = (float)((zc_d - last_exit) * rs.norm);
          if (cnt < cap) out[(size_t)cnt * kSlice] = make_uint2(kGapVox, __float_as_uint(gl));
          if (vis) vis->record(cnt, gl);
        }
        ++cnt;
        skipped = false;
      }
      if (!any) {
        wo.z0 = (float)zc_d;
        if constexpr (MODE == kEmit) {
          if (vis) vis->z0 = wo.z0;
        }
        any = true;
      }
      if constexpr (MODE == kEmit || MODE == kVisit) {
        const uint32_t local = (uint32_t)((((ix[0] & 7) << 3) | (ix[1] & 7)) << 3 | (ix[2] & 7));
        const uint32_t vox = (uint32_t)bslot * kBrickVox + local;
        float len = (float)((zn_d - zc_d) * rs.norm);
        if constexpr (MODE == kEmit) {
          if (cnt < cap) out[(size_t)cnt * kSlice] = make_uint2(vox, __float_as_uint(len));
          if (vis) vis->record(cnt, len);
        } else {
          if (vis->voxel(vox, ix, len, zc_d, zn_d)) break;
        }
      }
      ++cnt;
      last_exit = zn_d;
    } else if (any) {
      skipped = true;
    }
    // every axis whose crossing equals the nearest one steps (selects, no branches)
    const bool st0 = m0 & ((best == 0) | ((best == 1) & (X01 == 0)) | ((best == 2) & (X02 == 0)));
    const bool st1 = m1 & ((best == 1) | ((best == 0) & (X01 == 0)) | ((best == 2) & (X12 == 0)));
    const bool st2 = m2 & ((best == 2) | ((best == 0) & (X02 == 0)) | ((best == 1) & (X12 == 0)));
    const int nx[3] = {ix[0] + (st0 ? sg[0] : 0), ix[1] + (st1 ? sg[1] : 0), ix[2] + (st2 ? sg[2] : 0)};
    const int64_t one = (int64_t)1 << kFix;
    n0 += st0 ? one : 0;
    n1 += st1 ? one : 0;
    n2 += st2 ? one : 0;
    X01 += (st0 ? s1 : 0) - (st1 ? s0 : 0);
    X02 += (st0 ? s2 : 0) - (st2 ? s0 : 0);
    X12 += (st1 ? s2 : 0) - (st2 ? s1 : 0);
    const bool leaving = nx[0] < 0 || nx[0] >= N[0] || nx[1] < 0 || nx[1] >= N[1] || nx[2] < 0 ||
                         nx[2] >= N[2];
    const bool brick_change = leaving || ((nx[0] >> kBrickLog) != (ix[0] >> kBrickLog)) ||
                              ((nx[1] >> kBrickLog) != (ix[1] >> kBrickLog)) ||
                              ((nx[2] >> kBrickLog) != (ix[2] >> kBrickLog));
    if (brick_change && active) {
      if constexpr (MODE == kVisit) {
        if (vis->brick_end(zn_d)) break;
      } else {
        const float* sc = sig;
        const float* bc = bia;
        const double zb = kDepths ? zn_d : (double)zn.n / (double)zn.d;
        const double lhs = zb + poly_eval(bc, lookup_n, zb);
        const double rhs = depth + 3.0 * poly_eval(sc, lookup_n, zb);
        if (lhs >= rhs) {
          wo.ran_out = false;
          break;
        }
      }
    }
    if (leaving) break;
    ix[0] = nx[0];
    ix[1] = nx[1];
    ix[2] = nx[2];
    zc = zn;
    zc_d = zn_d;
    if (brick_change) bslot = brick_lookup();
  }
  wo.n = cnt;
  wo.last_exit = last_exit;
  return wo;
}

struct PixelCtx {
  bool ok;
  int x, y;
  float d;
  RaySetup rs;
  uint32_t cell;
};

// Rays are enumerated in 8x8 pixel tiles (tiles row-major, pixels row-major
// inside a tile) so that rays sharing voxels sit next to each other in the
// lists and run together on the GPU.
__device__ __forceinline__ bool tile_pixel(const CamGeom& c, int t, int* x, int* y) {
  const int ntx = (c.cols + 7) >> 3;
  const int tile = t >> 6, l = t & 63;
  *x = (tile % ntx) * 8 + (l & 7);
  *y = (tile / ntx) * 8 + (l >> 3);
  return *x < c.cols && *y < c.rows;
}

__device__ __forceinline__ PixelCtx load_pixel(const CamGeom& c, const GridGeom& g, const float* depth,
                                               int t) {
  PixelCtx p;
  p.ok = false;
  if (!tile_pixel(c, t, &p.x, &p.y)) return p;
  p.d = depth[p.y * c.cols + p.x];
  if (!pixel_valid(c, p.x, p.y, p.d)) return p;
  if (!setup_ray(c, g, p.x, p.y, p.rs)) return p;
  p.cell = lookup_offset(c, g.lookup_n, p.x, p.y);
  p.ok = true;
  return p;
}

// Full-ray activation of one ray, walked over the brick grid: every brick
// the voxel walk would cross up to its truncation is marked.  Brick faces
// are voxel faces, so the bricks are met in the order of the voxel walk and
// each brick exit is the same exact rational (n 2^kFix - Oi) / Di, with the
// same tie rule (all tied axes step at once); the truncation test runs at
// each brick exit as in walk().  The return value bounds the voxel walk's
// record count: every voxel step advances at least one axis by one voxel,
// each axis runs monotonically from the entry voxel to a voxel of the last
// brick, so n <= 1 + sum over axes of |last - entry| with "last" the far
// voxel of that brick along the axis (ties and a stop short of that voxel
// only make n smaller).  0 for a ray that misses the grid.
__device__ uint32_t mark_bricks(const GridGeom& g, const RaySetup& rs, double depth, const float* sig,
                                const float* bia, uint8_t* __restrict__ bflags) {
  Rat zin;
  int ix[3], sg[3];
  if (!ray_entry(g, rs, zin, ix, sg)) return 0;
  const int NB[3] = {g.bx, g.by, g.bz};
  int b[3] = {ix[0] >> kBrickLog, ix[1] >> kBrickLog, ix[2] >> kBrickLog};
  // next brick-face crossing per moving axis: n_a / d_a, n_a growing by
  // 2^(kFix + kBrickLog) per brick; cross differences as in walk()
  int64_t n[3], d[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t Oi = rs.Oi[a], Di = rs.Di[a];
    if (Di > 0) {
      n[a] = (((int64_t)b[a] + 1) << (kFix + kBrickLog)) - Oi;
      d[a] = Di;
    } else if (Di < 0) {
      n[a] = Oi - ((int64_t)b[a] << (kFix + kBrickLog));
      d[a] = -Di;
    } else {
      n[a] = 1;
      d[a] = 0;
    }
  }
  const bool m0 = sg[0] != 0, m1 = sg[1] != 0, m2 = sg[2] != 0;
  const int64_t s0 = d[0] << (kFix + kBrickLog), s1 = d[1] << (kFix + kBrickLog), s2 = d[2] << (kFix + kBrickLog);
  int64_t X01 = n[0] * d[1] - n[1] * d[0], X02 = n[0] * d[2] - n[2] * d[0], X12 = n[1] * d[2] - n[2] * d[1];
  const int lookup_n = g.lookup_n;
  const int max_steps = g.bx + g.by + g.bz + 3;
  for (int step = 0; step < max_steps; ++step) {
    bflags[(b[0] * g.by + b[1]) * g.bz + b[2]] = 1;
    int best = m0 ? 0 : (m1 ? 1 : 2);
    if (m1 && best == 0 && X01 > 0) best = 1;
    if (m2 && best != 2 && (best == 0 ? X02 > 0 : X12 > 0)) best = 2;
    const int64_t zn_n = best == 0 ? n[0] : (best == 1 ? n[1] : n[2]);
    const int64_t zn_dd = best == 0 ? d[0] : (best == 1 ? d[1] : d[2]);
    const double zb = (double)zn_n / (double)zn_dd;
    const double lhs = zb + poly_eval(bia, lookup_n, zb);
    const double rhs = depth + 3.0 * poly_eval(sig, lookup_n, zb);
    if (lhs >= rhs) break;
    const bool st0 = m0 && (best == 0 || (best == 1 ? X01 == 0 : X02 == 0));
    const bool st1 = m1 && (best == 1 || (best == 0 ? X01 == 0 : X12 == 0));
    const bool st2 = m2 && (best == 2 || (best == 0 ? X02 == 0 : X12 == 0));
    const int nb[3] = {b[0] + (st0 ? sg[0] : 0), b[1] + (st1 ? sg[1] : 0), b[2] + (st2 ? sg[2] : 0)};
    if (nb[0] < 0 || nb[0] >= NB[0] || nb[1] < 0 || nb[1] >= NB[1] || nb[2] < 0 || nb[2] >= NB[2]) break;
    if (st0) {
      n[0] += (int64_t)1 << (kFix + kBrickLog);
      X01 += s1;
      X02 += s2;
    }
    if (st1) {
      n[1] += (int64_t)1 << (kFix + kBrickLog);
      X01 -= s0;
      X12 += s2;
    }
    if (st2) {
      n[2] += (int64_t)1 << (kFix + kBrickLog);
      X02 -= s0;
      X12 -= s1;
    }
    b[0] = nb[0];
    b[1] = nb[1];
    b[2] = nb[2];
  }
  uint32_t bound = 1;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (sg[a] > 0) bound += (uint32_t)((b[a] << kBrickLog) + (1 << kBrickLog) - 1 - ix[a]);
    if (sg[a] < 0) bound += (uint32_t)(ix[a] - (b[a] << kBrickLog));
  }
  return bound;
}

// Full-ray activation: marks every brick a ray crosses up to its truncation.
// A ray's own walk activates every brick before and at its stopping brick,
// so its later list walk sees exactly these bricks, with no gaps; bound (per
// tiled pixel, when given) receives the bound of that walk's record count.
__global__ void k_mark_rays(CamGeom c, GridGeom g, const float* __restrict__ depth,
                            const float* __restrict__ sig, const float* __restrict__ bia,
                            uint8_t* __restrict__ bflags, uint32_t* __restrict__ bound) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= tiled_pixels(c)) return;
  PixelCtx p = load_pixel(c, g, depth, pix);
  uint32_t n = 0;
  if (p.ok) n = mark_bricks(g, p.rs, (double)p.d, sig + p.cell, bia + p.cell, bflags);
  if (bound) bound[pix] = n;
}

__global__ void k_mark_endpoints(CamGeom c, GridGeom g, const float* __restrict__ depth,
                                 uint8_t* __restrict__ bflags) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= tiled_pixels(c)) return;
  PixelCtx p = load_pixel(c, g, depth, pix);
  if (!p.ok) return;
  const int M[3] = {g.mx, g.my, g.mz};
  const int N[3] = {g.nx, g.ny, g.nz};
  int lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    double P = (double)p.d * (double)p.rs.Di[a];
    P = (double)p.rs.Oi[a] + P;
    P = P / kFixScale;
    if (!(P >= 0.0 && P < (double)M[a])) return;
    int l = (int)floor(P - 1.0), h = (int)floor(P + 1.0);
    if (l < 0) l = 0;
    if (h > N[a] - 1) h = N[a] - 1;
    lo[a] = l >> kBrickLog;
    hi[a] = h >> kBrickLog;
  }
  for (int i = lo[0]; i <= hi[0]; ++i)
    for (int j = lo[1]; j <= hi[1]; ++j)
      for (int k = lo[2]; k <= hi[2]; ++k) bflags[(i * g.by + j) * g.bz + k] = 1;
}

// Depth endpoints per voxel (a feature of the learned warm start): the voxel
// holding the endpoint of every valid pixel, with the endpoint formula and
// map-box gate of k_mark_endpoints; endpoints in inactive bricks are dropped.
__global__ void k_feat_hits(CamGeom c, GridGeom g, const float* __restrict__ depth,
                            const int32_t* __restrict__ slot, uint32_t* __restrict__ hits) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= tiled_pixels(c)) return;
  PixelCtx p = load_pixel(c, g, depth, pix);
  if (!p.ok) return;
  const int M[3] = {g.mx, g.my, g.mz};
  int v[3];
  for (int a = 0; a < 3; ++a) {
    double P = (double)p.d * (double)p.rs.Di[a];
    P = (double)p.rs.Oi[a] + P;
    P = P / kFixScale;
    if (!(P >= 0.0 && P < (double)M[a])) return;
    v[a] = (int)floor(P);
  }
  const int sl = slot[((v[0] >> kBrickLog) * g.by + (v[1] >> kBrickLog)) * g.bz + (v[2] >> kBrickLog)];
  if (sl < 0) return;
  atomicAdd(&hits[(size_t)sl * kBrickVox + ((v[0] & 7) * 64 + (v[1] & 7) * 8 + (v[2] & 7))], 1u);
}

__global__ void k_count(CamGeom c, GridGeom g, const float* __restrict__ depth,
                        const float* __restrict__ sig, const float* __restrict__ bia,
                        const int32_t* __restrict__ slot, uint32_t* __restrict__ counts) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= tiled_pixels(c)) return;
  PixelCtx p = load_pixel(c, g, depth, pix);
  uint32_t n = 0;
  if (p.ok) n = walk<kCount>(g, p.rs, (double)p.d, sig + p.cell, bia + p.cell, slot, nullptr, 0u).n;
  counts[pix] = n;
}

// Records per slice (max over the tile's 64 rays) and the slice's padded size.
__global__ void k_slice_len(const uint32_t* __restrict__ counts, uint32_t ntiles, uint32_t* __restrict__ slice_len,
                            uint32_t* __restrict__ slice_rec) {
  const uint32_t t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (t >= ntiles) return;
  uint32_t m = counts[t * kSlice + (threadIdx.x & 63)];
  for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o));
  if ((threadIdx.x & 63) == 0) {
    slice_len[t] = m;
    slice_rec[t] = m * kSlice;
  }
}

// Emission of every ray's records into the staging lists, split into front
// and window as they go: front[pix].nf front records, then lanes[pix].nw
// window records; cnt_f / cnt_w per tiled pixel.
__global__ void k_emit(CamGeom c, GridGeom g, const float* __restrict__ depth, const float* __restrict__ sig,
                       const float* __restrict__ bia, const int32_t* __restrict__ slot,
                       const uint32_t* __restrict__ slice_off, const uint32_t* __restrict__ slice_len,
                       uint2* __restrict__ sell, FrontRec* __restrict__ front, RayRec* __restrict__ lanes,
                       uint32_t* __restrict__ cnt_f, uint32_t* __restrict__ cnt_w, uint32_t* __restrict__ overflow) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= tiled_pixels(c)) return;
  const uint32_t tile = (uint32_t)pix >> 6, lane = (uint32_t)pix & 63u;
  uint2* out = sell + slice_off[tile] + lane;
  PixelCtx p = load_pixel(c, g, depth, pix);
  FrontRec f;
  f.off = slice_off[tile] + lane;  // front and window records follow each other until the reslice
  f.nf = 0;
  RayRec r;
  r.off_w = 0;
  r.nw = 0;
  r.ks = 0.0f;
  r.kc = 0.0f;
  r.z0 = 0.0f;
  r.inv_norm = 0.0f;
  r.depth = 0.0f;
  r.cell_flags = 0;
  if (p.ok) {
    FrontSplit sp;
    load_cell(sig, bia, g.lookup_n, p.cell, sp.sig, sp.bia);
    sp.inv_norm = (float)(1.0 / p.rs.norm);
    sp.depth = p.d;
    sp.z0 = 0.0f;
    sp.kS = {0.0f, 0.0f};
    sp.nf = 0;
    sp.open = true;
    const uint32_t cap = slice_len[tile];
    WalkOut wo = walk<kEmit, FrontSplit>(g, p.rs, (double)p.d, sig + p.cell, bia + p.cell, slot, out, cap, &sp);
    uint32_t n = wo.n;
    if (n > cap) {  // the slot's bound did not hold: an error, and the lists stay in bounds
      atomicOr(overflow, 1u);
      n = cap;
    }
    const uint32_t nf = sp.open ? n : min(sp.nf, n);
    f.nf = nf;
    r.nw = n - nf;
    r.ks = sp.kS.s;
    r.kc = sp.kS.c;
    r.z0 = wo.z0;
    r.inv_norm = sp.inv_norm;
    r.depth = p.d;
    // the infinity term's test d > s_n + bias(s_n) at the exact exit depth of
    // the last record, in FP64 as the traversal itself: an FP32 running sum
    // of chord lengths can land on either side when d is within an ulp
    const bool inf = wo.ran_out && (double)p.d > wo.last_exit + poly_eval(bia + p.cell, g.lookup_n, wo.last_exit);
    r.cell_flags = p.cell | (wo.ran_out ? kRanOut : 0u) | (inf ? kInfTerm : 0u);
  }
  // records past the ray's own are never read: no padding
  front[pix] = f;
  lanes[pix] = r;
  cnt_f[pix] = f.nf;
  cnt_w[pix] = r.nw;
}

// ---------------------------------------------------------------- map evaluation
// One pixel ray walked through the current map to the end of the grid (no
// truncation), every voxel of an active brick read as its marginal
// occupancy p (0 where never updated, gvdbkernels.cu:469-476).  Per voxel:
// alpha = -ln(1 - p), Lambda += -alpha l, vis = e^Lambda, weight
// vis_prev - vis (Crispell's visibility model, gvdbkernels.cu:230-236).
//  * log-likelihood (rayLikelihoodBrick + render_likelihood_kernel,
//    gvdbkernels.cu:183-293, 1246-1274): sum of weight x p(z | voxel), with
//    p(z | voxel) = 0 below Z_MIN, (1 - LIKE_EPS) (psi(s) + psi(s_prev)) / 2
//    within 3 sigma of the measured depth, LIKE_EPS elsewhere; plus
//    vis_inf x (d >= Z_MAX ? 1 - LIKE_EPS : LIKE_EPS); the pixel's value is
//    its log; invalid depth gives 0.
//  * expected depth (rayExpectedDepthBrick + render_expected_depth_kernel,
//    gvdbkernels.cu:295-354, 1332-1363): sum of weight x the planar depth of
//    the voxel centre, over 1 - vis_inf + F_EPS; every pixel.
//  * accuracy (render_accuracy_kernel, gvdbkernels.cu:1066-1104): at the
//    voxel with the largest weight per unit length, 1 if its depth is within
//    1 sigma of the measurement else 0; -1 for an invalid pixel, a ray that
//    met nothing, or one that kept vis >= 0.9 and ended short of the
//    measurement (0 if it ended beyond it).
//  * simple accuracy (raySimpleAccuracyBrick + render_simple_accuracy_kernel,
//    gvdbkernels.cu:135-181, 1106-1134): the first voxel with
//    p >= acc_thresh gives 1 (within one voxel size of the measured depth) or
//    2; an active brick ending at or beyond the measured depth first gives 0;
//    otherwise -1.
// The prefix depth s is the planar depth at the voxel's exit; at an active
// brick's end the previous depth carried into the next brick is s + bias(s)
// (the reference passes it on in norm.x).
constexpr float kZMin = 0.1f;
constexpr float kLikeEps = 1e-3f;

__device__ __forceinline__ float poly_f(const float* c, int n, float s) {
  return n == 3 ? (s * s) * c[2] + s * c[1] + c[0] : s * c[1] + c[0];
}

__device__ __forceinline__ float psi_f(float s, float d, const float* sc, const float* bc, int n) {
  const float sg3 = 3.0f * poly_f(sc, n, s);
  const float u = (d - s - poly_f(bc, n, s)) / sg3;
  return 0.398942f * expf(-0.5f * u * u) / sg3;
}

struct EvalVisitor {
  const float* B;
  const uint8_t* upd;
  const float* sc;
  const float* bc;
  int ln;
  float d, res, acc_thresh;
  bool valid;
  double R2[3], t[3], off[3];  // camera z axis in world, centre, world->voxel offset
  float prev_vis, lam, sum, prev_s, norm_x, peak, esum, sacc;
  int peak_ok;
  bool first, sdone;

  __device__ float marginal_of(uint32_t v) const {
    if (!upd[v]) return 0.0f;
    float neg = 1.0f / (1.0f + expf(B[v]));
    neg = fminf(fmaxf(neg, kFEps), 1.0f - kFEps);
    return 1.0f - neg;
  }
  __device__ bool voxel(uint32_t v, const int ix[3], float len, double z0, double z1) {
    const float p = marginal_of(v);
    const float s = (float)z1;
    if (first) {
      prev_s = (float)z0;
      first = false;
    }
    const float alpha = -logf(1.0f - p);
    lam += -alpha * len;
    const float vis = expf(lam);
    const float w = (prev_vis - vis) / len;
    if (valid) {
      const float pdm = s + poly_f(bc, ln, s) - d;
      const float sg = poly_f(sc, ln, s);
      float prob;
      if (s < kZMin)
        prob = 0.0f;
      else if (fabsf(pdm) <= 3.0f * sg)
        prob = (1.0f - kLikeEps) * 0.5f * (psi_f(s, d, sc, bc, ln) + psi_f(prev_s, d, sc, bc, ln));
      else
        prob = kLikeEps;
      if (w > peak) {
        peak = w;
        peak_ok = fabsf(pdm) <= sg ? 1 : 0;
      }
      sum += (prev_vis - vis) * prob;
      norm_x = pdm + d;
      if (!sdone && p >= acc_thresh) {
        sacc = fabsf(s - d) <= res ? 1.0f : 2.0f;
        sdone = true;
      }
    }
    double zc = 0.0;
    for (int a = 0; a < 3; ++a) zc += R2[a] * ((((double)ix[a] + 0.5) - off[a]) * res - t[a]);
    esum += (prev_vis - vis) * (float)zc;
    prev_vis = vis;
    prev_s = s;
    return false;
  }
  __device__ bool brick_end(double z) {
    if (valid) {
      prev_s = norm_x;
      if (!sdone && (float)z >= d) {
        sacc = 0.0f;
        sdone = true;
      }
    }
    return false;
  }
};

__global__ __launch_bounds__(256) void k_eval(CamGeom c, GridGeom g, const float* __restrict__ depth,
                                              const float* __restrict__ sigt, const float* __restrict__ biat,
                                              const int32_t* __restrict__ slot, const float* __restrict__ B,
                                              const uint8_t* __restrict__ upd, float acc_thresh,
                                              float* __restrict__ out_ll, float* __restrict__ out_ed,
                                              float* __restrict__ out_acc, float* __restrict__ out_sacc,
                                              double* __restrict__ part, unsigned long long* __restrict__ counts) {
  __shared__ double sh[256];
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  int x = 0, y = 0;
  const bool in = pix < tiled_pixels(c) && tile_pixel(c, pix, &x, &y);
  float ll = 0.0f, ed = 0.0f, acc = -1.0f, sacc = -1.0f;
  if (in) {
    const float d = depth ? depth[y * c.cols + x] : 0.0f;
    EvalVisitor v;
    v.B = B;
    v.upd = upd;
    const uint32_t cell = lookup_offset(c, g.lookup_n, x, y);
    v.sc = sigt + cell;
    v.bc = biat + cell;
    v.ln = g.lookup_n;
    v.d = d;
    v.res = (float)g.res;
    v.acc_thresh = acc_thresh;
    v.valid = depth != nullptr && pixel_valid(c, x, y, d);
    for (int a = 0; a < 3; ++a) {
      v.R2[a] = c.R[a * 3 + 2];
      v.t[a] = c.t[a];
      v.off[a] = g.off[a];
    }
    v.prev_vis = 1.0f;
    v.lam = v.sum = v.prev_s = v.norm_x = v.peak = v.esum = 0.0f;
    v.sacc = -1.0f;
    v.peak_ok = 0;
    v.first = true;
    v.sdone = false;
    RaySetup rs;
    if (setup_ray(c, g, x, y, rs)) walk<kVisit>(g, rs, 0.0, nullptr, nullptr, slot, nullptr, 0u, &v);
    ed = v.esum / (1.0f - v.prev_vis + kFEps);
    if (v.valid) {
      ll = logf(v.sum + v.prev_vis * (d >= kZMax ? 1.0f - kLikeEps : kLikeEps));
      if (v.peak > 0.0f && v.prev_vis >= 0.9f)
        acc = v.norm_x <= d ? -1.0f : 0.0f;
      else if (v.peak > 0.0f)
        acc = (float)v.peak_ok;
      sacc = v.sacc;
    }
    const int o = y * c.cols + x;
    if (out_ll) out_ll[o] = ll;
    if (out_ed) out_ed[o] = ed;
    if (out_acc) out_acc[o] = acc;
    if (out_sacc) out_sacc[o] = sacc;
  }
  // class counts: accuracy -1, 0, 1 then simple accuracy -1, 0, 1, 2 (pixels of the image only)
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const float want = k < 3 ? (float)(k - 1) : (float)(k - 4);
    const float val = k < 3 ? acc : sacc;
    const unsigned long long m = __ballot(in && val == want);
    if (lane == 0 && m) atomicAdd(counts + k, (unsigned long long)__popcll(m));
  }
  // fixed-order FP64 block sum of the log-likelihoods
  sh[threadIdx.x] = (double)ll;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

}  // namespace

int num_tiled_pixels(const CamGeom& c) { return tiled_pixels(c); }

unsigned eval_blocks(const CamGeom& c) { return (unsigned)((tiled_pixels(c) + 255) / 256); }

void launch_eval(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig, const float* bia,
                 const int32_t* slot, const float* B, const uint8_t* upd, float acc_thresh, float* ll, float* ed,
                 float* acc, float* sacc, double* part, unsigned long long* counts, hipStream_t s) {
  hipLaunchKernelGGL(k_eval, dim3(eval_blocks(c)), dim3(256), 0, s, c, g, depth, sig, bia, slot, B, upd, acc_thresh, ll,
                     ed, acc, sacc, part, counts);
}

void launch_mark_rays(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                      const float* bia, uint8_t* bflags, uint32_t* bound, hipStream_t s) {
  const int n = tiled_pixels(c);
  hipLaunchKernelGGL(k_mark_rays, dim3((n + 127) / 128), dim3(128), 0, s, c, g, depth, sig, bia, bflags, bound);
}

void launch_feat_hits(const CamGeom& c, const GridGeom& g, const float* depth, const int32_t* slot, uint32_t* hits,
                      hipStream_t s) {
  const int n = tiled_pixels(c);
  hipLaunchKernelGGL(k_feat_hits, dim3((n + 255) / 256), dim3(256), 0, s, c, g, depth, slot, hits);
}

void launch_mark_endpoints(const CamGeom& c, const GridGeom& g, const float* depth, uint8_t* bflags,
                           hipStream_t s) {
  const int n = tiled_pixels(c);
  hipLaunchKernelGGL(k_mark_endpoints, dim3((n + 255) / 256), dim3(256), 0, s, c, g, depth, bflags);
}

void launch_count(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                  const float* bia, const int32_t* slot, uint32_t* counts, hipStream_t s) {
  const int n = tiled_pixels(c);
  hipLaunchKernelGGL(k_count, dim3((n + 127) / 128), dim3(128), 0, s, c, g, depth, sig, bia, slot, counts);
}

void launch_slice_len(const uint32_t* counts, uint32_t ntiles, uint32_t* slice_len, uint32_t* slice_rec,
                      hipStream_t s) {
  if (ntiles == 0) return;
  hipLaunchKernelGGL(k_slice_len, dim3((ntiles + 3) / 4), dim3(256), 0, s, counts, ntiles, slice_len, slice_rec);
}

void launch_emit(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig, const float* bia,
                 const int32_t* slot, const uint32_t* slice_off, const uint32_t* slice_len, uint2* sell,
                 FrontRec* front, RayRec* lanes, uint32_t* cnt_f, uint32_t* cnt_w, uint32_t* overflow,
                 hipStream_t s) {
  const int n = tiled_pixels(c);
  hipLaunchKernelGGL(k_emit, dim3((n + 127) / 128), dim3(128), 0, s, c, g, depth, sig, bia, slot, slice_off,
                     slice_len, sell, front, lanes, cnt_f, cnt_w, overflow);
}

}  // namespace mrf
