// Host-side launchers of the HIP kernels of libmrfmap_hip (internal).
#pragma once

#include "mrf_internal.h"

namespace mrf {

enum MsgType { kMsgF32 = 0, kMsgF16 = 1, kMsgBF16 = 2 };

// traverse.hip
void launch_mark_rays(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                      const float* bia, uint8_t* bflags, hipStream_t s);
void launch_mark_endpoints(const CamGeom& c, const GridGeom& g, const float* depth, uint8_t* bflags,
                           hipStream_t s);
void launch_count(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                  const float* bia, const int32_t* slot, uint32_t* counts, uint32_t* has, hipStream_t s);
void launch_emit(const CamGeom& c, const GridGeom& g, const float* depth, const float* sig,
                 const float* bia, const int32_t* slot, const uint32_t* offs, const uint32_t* ray_idx,
                 uint2* inc, RayRec* rays, hipStream_t s);

// scan.hip: exclusive prefix sum of n uint32 values; total is written to
// *d_total (device, uint64).  tmp must hold scan_tmp_words(n) uint64 words.
size_t scan_tmp_words(size_t n);
void launch_exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint64_t* tmp, uint64_t* d_total,
                           hipStream_t s);

// bp.hip
struct MsgLaunch {
  const RayRec* rays;
  uint32_t nrays;
  const uint2* inc;
  const float* B;
  const void* M;     // message array of this camera (MsgType)
  float* S;          // ray->voxel accumulator (pool sized)
  const float* sig;  // sensor-model lookup tables
  const float* bia;
  int lookup_n;
  int msg_type;
};
void launch_ray_messages(const MsgLaunch& a, hipStream_t s);

void launch_cum_len(const RayRec* rays, uint32_t nrays, const uint2* inc, unsigned long long* cfix,
                    hipStream_t s);
// Compacts pool voxels with cfix > 0 into (vox, C) pairs in ascending voxel
// order and clears cfix.  Needs scan scratch for nvox / 2048 blocks.
void launch_build_pairs(unsigned long long* cfix, size_t nvox, uint32_t* block_cnt, uint32_t* block_off,
                        uint64_t* scan_tmp, uint64_t* d_total, uint32_t* pair_vox, float* pair_c,
                        hipStream_t s);
size_t pair_blocks(size_t nvox);

void launch_epilogue(const uint32_t* pair_vox, const float* pair_c, uint32_t npairs, float* B, void* M,
                     int msg_type, float* S, uint8_t* upd, hipStream_t s);
void launch_marginals(const float* B, float* Aprev, size_t nvox, hipStream_t s);
void launch_residual(const float* B, float* Aprev, size_t nvox, unsigned long long* d_max, hipStream_t s);
void launch_alphas(const float* B, const uint8_t* upd, float* out, size_t nvox, hipStream_t s);
void launch_fill_f32(float* p, float v, size_t n, hipStream_t s);
void launch_fill_u8(uint8_t* p, uint8_t v, size_t n, hipStream_t s);
void launch_zero_words(void* p, size_t bytes, hipStream_t s);

}  // namespace mrf
