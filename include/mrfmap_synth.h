/*
 * Synthetic depth-keyframe generator for the MRFMap BP inference path.
 *
 * No counterpart in the reference: the reference ships 13 recorded frames
 * (data/depth_*.bin, data/pose_*.bin) and no generator.  This library makes
 * deterministic scenes of axis-aligned boxes and spheres inside a walled room,
 * camera rings looking at the room centre, and planar-z depth images with the
 * structured-light noise model sigma(z) = 0.0012 + 0.0019 (z - 0.4)^2 m and
 * Bernoulli dropout to 0.  It is host-only C (no GPU), used by the tests, the
 * benchmark and the CPU oracle to create identical inputs.
 *
 * Conventions (shared with libmrfmap_hip):
 *   pose  : 4x4 camera-to-world, row-major, optical frame (x right, y down,
 *           z forward).
 *   depth : rows x cols float32, row-major, metric planar z; 0 = invalid.
 *   ray   : pixel (x = column, y = row) looks along R * ((x-cx)/fx, (y-cy)/fy, 1).
 */
#ifndef MRFMAP_SYNTH_H
#define MRFMAP_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mrf_synth_cfg {
  float room[3];        /* room extent (m): x and y in [-room/2, room/2), z in [0, room_z) */
  int num_boxes;        /* axis-aligned boxes, edges uniform in [box_min, box_max] */
  float box_min, box_max;
  int num_spheres;      /* spheres, radius uniform in [sph_min, sph_max] */
  float sph_min, sph_max;
  float ring_radius;    /* camera ring radius (m) */
  float ring_height;    /* camera height (m) */
  float yaw_jitter_deg; /* uniform yaw jitter (+-) about world z */
  float max_range;      /* depth beyond this is reported as 0 (<= 0: no limit) */
  uint64_t seed;
} mrf_synth_cfg;

typedef struct mrf_synth_scene mrf_synth_scene;

/* Fill cfg with the 4x4x3 m room, 12 boxes 0.2-1 m, 4 spheres r 0.1-0.4 m,
 * 2 m ring at 1.5 m, +-5 deg jitter. */
void mrf_synth_default_room(mrf_synth_cfg* cfg, uint64_t seed);
/* Fill cfg with the 10x10x3 m room, 60 boxes, 4 m ring, 10 m max range. */
void mrf_synth_large_room(mrf_synth_cfg* cfg, uint64_t seed);

int mrf_synth_create(const mrf_synth_cfg* cfg, mrf_synth_scene** out);
void mrf_synth_destroy(mrf_synth_scene* s);

/* Pose of keyframe k of n on the ring (yaw jitter drawn from the scene seed). */
int mrf_synth_ring_pose(const mrf_synth_scene* s, int k, int n, float pose[16]);

/* Render planar-z depth for one camera.  noise != 0 adds the Gaussian noise
 * model and `dropout` fraction of zeros, drawn from `noise_seed`. */
int mrf_synth_render(const mrf_synth_scene* s, const float pose[16], float fx, float fy, float cx,
                     float cy, int rows, int cols, int noise, float dropout, uint64_t noise_seed,
                     float* depth);

/* Ground-truth occupancy of voxel centres: for a grid with resolution res whose
 * voxel (i,j,k) centre is at ((i+.5)res - room_x/2, (j+.5)res - room_y/2, (k+.5)res),
 * writes 1 for centres inside an object (or outside the room) else 0. */
int mrf_synth_occupancy(const mrf_synth_scene* s, float res, const float* centers_xyz, int64_t n,
                        uint8_t* occ);

/* Number of boxes / spheres and their parameters (for tests). */
int mrf_synth_num_objects(const mrf_synth_scene* s, int* boxes, int* spheres);
int mrf_synth_box(const mrf_synth_scene* s, int i, float lo[3], float hi[3]);
int mrf_synth_sphere(const mrf_synth_scene* s, int i, float c[3], float* r);

#ifdef __cplusplus
}
#endif
#endif
