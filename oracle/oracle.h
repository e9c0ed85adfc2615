/*
 * CPU oracle of the MRFMap BP inference path — TEST INFRASTRUCTURE ONLY.
 *
 * A scalar FP64 restatement of the reference's live algorithm, used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker.  Nothing in the product (mrfmap_amd/) links or calls it.
 *
 * Parts and what each restates (paths relative to the reference repository):
 *   traversal   GVDB rayCast/HDDA + brick functions as driven by
 *               src/cuda/gvdbkernels.cu:519-571, 689-1038, 1164-1226,
 *               1276-1302; endpoint activation src/GVDBCamera.cpp:70-112,
 *               src/GVDBMRFMap.cpp:117-135.  GVDB itself is not vendored,
 *               so the traversal rule is this project's (exact rational
 *               crossings, brick-granular truncation) and is pinned by brute
 *               force, closed forms and invariants in tests/.
 *   messages    gvdbkernels.cu:689-1038 (pre/post), 1196-1208 (infinity term),
 *               46-72 (sensor model) — orc_ray_messages.
 *   epilogue    gvdbkernels.cu:457-479.
 *   schedule    src/GVDBInference.cpp:100-178 (Gauss-Seidel over cameras).
 *   readback    src/GVDBMRFMap.cpp:137-208, gvdbkernels.cu:360-373.
 *   octomap-eq  OctoMap's published OcTree::insertPointCloud as configured
 *               by src/GVDBOctomapWrapper.cpp:3-43 (CPU comparator only).
 */
#ifndef MRF_ORACLE_H
#define MRF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
  float K[9];
  int rows, cols;
  float res, prior;
  float dims[3];
  int lookup_n;
  float sigma_poly[3], bias_poly[3]; /* polynomials of every cell when the tables are NULL */
  int activation;                    /* 0 endpoint, 1 full ray */
  int mask_848_border;
  /* optional per-cell tables, ceil(rows/20) x ceil(cols/20) cells of lookup_n
   * coefficients (GVDBParams.h:93-104, 134-162), indexed in the vertically
   * flipped image (gvdbkernels.cu:46-55); every keyframe must then be rows x cols */
  const float* sigma_lookup;
  const float* bias_lookup;
  /* 1: the pre pass restarts visibility at every brick until a tau survives
   * FP32 flush-to-zero, as the reference's (gvdbkernels.cu:708-715);
   * 0: visibility carried along the whole ray in both passes */
  int brick_local_vis;
} orc_params;

typedef struct orc_ctx orc_ctx;

int orc_create(const orc_params* p, orc_ctx** out);
void orc_destroy(orc_ctx* c);
/* pose row-major 4x4 camera-to-world; K row-major 3x3; depth row-major */
int orc_add_camera(orc_ctx* c, const float pose[16], const float K[9], const float* depth, int rows, int cols);
int orc_set_pose(orc_ctx* c, int cam, const float pose[16]);
/* external occupancy map: points n x 3 in voxel units, occupancy n */
int orc_load_map(orc_ctx* c, const float* pts, const float* occ, int64_t n);
/* test hook: beliefs and update flags of the whole pool */
int orc_set_beliefs(orc_ctx* c, const double* B, const uint8_t* upd);
/* map evaluation images (log-likelihood, expected depth, accuracy, simple accuracy) */
int orc_evaluate(orc_ctx* c, const float pose[16], const float K[9], const float* depth, int rows, int cols,
                 double acc_thresh, double* ll, double* ed, double* acc, double* sacc);
/* activation + slot assignment + traversal of every camera */
int orc_prepare(orc_ctx* c);
int64_t orc_num_bricks(orc_ctx* c);
int orc_brick_coords(orc_ctx* c, int32_t* xyz /* n x 3 brick coordinates */);
/* traversal lists in the libmrfmap_hip export format (6 and 2 uint32 words) */
int orc_get_traversal(orc_ctx* c, int cam, int64_t* n_rays, int64_t* n_records, uint32_t* rays,
                      uint32_t* records);
/* FP64 BP.  ids NULL -> all cameras have their messages reset.  Returns
 * iterations run; *residual gets the last max |delta marginal|. */
int orc_perform_inference(orc_ctx* c, const int* ids, int n_ids, int max_iters, double tol, double* residual,
                          int64_t* updates_per_iter);
/* Threads of the BP loop (0: OpenMP's default).  Results depend on the
 * count only through the FP64 summation order of the per-thread sums. */
void orc_set_threads(int n);
int orc_get_threads(void);
/* Iterations from the current beliefs and messages, without the per-call
 * message reset (the warm-started continuation). */
int orc_continue_inference(orc_ctx* c, int max_iters, double tol, double* residual);
/* Warm-start features (32 per touched voxel, pool order) and the warm-start
 * update from given predictions; mrfmap_amd/csrc/warm.hip defines both. */
int orc_features(orc_ctx* c, float* X, uint32_t* vox, int64_t cap, int64_t* n);
int orc_warm_start(orc_ctx* c, const uint32_t* vox, const double* pred, int64_t n);
int orc_get_alphas(orc_ctx* c, double* out /* nbricks x 512 */);
int orc_get_beliefs(orc_ctx* c, double* out);
/* camera cam's summed chord length C per voxel (nbricks x 512; 0 where its
 * rays do not pass) */
int orc_get_cum_lengths(orc_ctx* c, int cam, double* out);

/* Messages of one ray (the per-ray factor of the MRF).
 *   n[i]   : incoming P(o_i = 0) (already clamped by the caller)
 *   len[i] : chord length (voxels); s[i] planar depth at the voxel exit (m)
 *   live[i]: 0 for gap records
 *   psi_inf: value of the "no surface" term when ran_out (reference: 1)
 *   mode 0 : reference (inclusive visibility, length-weighted logs, clamp)
 *   mode 1 : textbook sum-product (exclusive visibility, unit weights, no
 *            clamp, no normaliser epsilon) — equals brute-force enumeration.
 * Writes p[i] (outgoing P(o_i = 1) after normalisation); returns T. */
double orc_ray_messages(int n, const double* nin, const double* len, const double* s, const int* live, double d,
                        const double sig[3], const double bia[3], int lookup_n, int ran_out, double psi_inf,
                        int mode, double* p);
/* The same with the reference's pre-pass visibility restart (mode 0 only):
 * bstart[i] = 1 where record i enters a brick (the first record of the
 * ray, the record after a gap, a record whose voxel lies in another brick
 * than the previous record's).  NULL: visibility carried along the whole ray
 * in both passes (orc_ray_messages). */
double orc_ray_messages_ex(int n, const double* nin, const double* len, const double* s, const int* live,
                           const int* bstart, double d, const double sig[3], const double bia[3], int lookup_n,
                           int ran_out, double psi_inf, int mode, double* p);
/* Brute force over all 2^n occupancy configurations of the textbook factor. */
int orc_ray_messages_bruteforce(int n, const double* nin, const double* psi, double psi_inf, double* p);
double orc_depth_psi(double s, double d, const double sig[3], const double bia[3], int lookup_n);

/* Traversal of a single ray (no activation: every brick active), for
 * closed-form and brute-force tests.  O and D are in voxel units (D per
 * metre of planar depth), grid n^3 voxels.  Writes up to cap voxels
 * (x,y,z), chord lengths (double) and exit depths; returns the count. */
int orc_walk_ray(const double O[3], const double D[3], const int N[3], double depth, int trunc,
                 const double sig[3], const double bia[3], int lookup_n, int cap, int32_t* vox, double* len,
                 double* zexit, int* ran_out);

/* OctoMap-equivalent insertion (CPU comparator, not OctoMap itself). */
typedef struct oct_map oct_map;
int oct_create(double res, const double bbx_min[3], const double bbx_max[3], double occ_thresh, oct_map** out);
void oct_destroy(oct_map* m);
/* points n x 3 world metres, origin = sensor position; returns ray-voxel updates */
int64_t oct_insert_point_cloud(oct_map* m, const float* pts, int64_t n, const double origin[3]);
int64_t oct_num_nodes(oct_map* m);
/* log-odds at world point (NAN if unknown) */
double oct_logodds_at(oct_map* m, const double p[3]);

#ifdef __cplusplus
}
#endif
#endif
