/*
 * libmrfmap_hip — C ABI of the MI355X BP inference path of MRFMap.
 *
 * The library replaces the GVDB/CUDA hot path behind the reference's
 * GVDBInference object (include/mrfmap/GVDBInference.h:20-103): keyframes
 * (pose + depth) are added, perform_inference runs loopy belief propagation
 * over the per-ray forward-sensor factors on one GPU, and the occupancy
 * marginals are read back brick by brick.  All entry points are synchronous
 * on return, take plain pointers and sizes, never throw, and return 0 on
 * success or a negative MRF_E_* code (text via mrf_last_error).  A context is
 * not thread-safe: one caller thread per context, as in the reference
 * (GVDBInference.cpp:201 has no lock either).
 *
 * Conventions:
 *   pose  : 4x4 camera-to-world in the optical frame (x right, y down,
 *           z forward), float[16], row-major unless MRF_COL_MAJOR is given
 *           (Eigen::MatrixXf default storage is column-major).
 *   depth : rows x cols float32 metric planar depth; NaN, inf, 0 and
 *           negatives mark invalid pixels (GVDBImage.cpp:32 maps NaN to -1;
 *           gvdbkernels.cu:1170-1173 skips d <= 0).  Row-major unless
 *           MRF_COL_MAJOR.
 *   map   : voxel v = p/res + (dims_x/2, dims_y/2, 0)/res
 *           (GVDBCamera.cpp:13-18); bricks are 8^3 voxels; brick rows are
 *           returned in slot order, voxels inside a brick in x*64 + y*8 + z
 *           order (gvdbkernels.cu:372).
 */
#ifndef MRFMAP_HIP_H
#define MRFMAP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRF_OK 0
#define MRF_E_INVALID (-1) /* bad argument or shape */
#define MRF_E_HIP (-2)     /* HIP runtime error */
#define MRF_E_NOMEM (-3)   /* device allocation failed */
#define MRF_E_SHAPE (-4)   /* output buffer shape mismatch (GVDBMRFMap.cpp:174-177 throws) */
#define MRF_E_RANGE (-5)   /* camera id out of range */
#define MRF_E_STATE (-6)   /* call not valid in the current state */
#define MRF_E_LIMIT (-7)   /* a size limit of the implementation is exceeded */

#define MRF_ROW_MAJOR 0
#define MRF_COL_MAJOR 1

/* How perform_inference launches a BP iteration. */
#define MRF_LAUNCH_GRAPH 0 /* replay of a captured HIP graph (the default; eager if capture fails) */
#define MRF_LAUNCH_TIMED 1 /* kernels launched one by one with HIP events around each camera's
                              ray and epilogue kernels (per-kernel timings in mrf_stats) */

/* Which bricks join the MRF. */
#define MRF_ACT_ENDPOINT 0 /* bricks within 1 voxel of a depth endpoint (GVDBMRFMap.cpp:117-135) */
#define MRF_ACT_FULL_RAY 1 /* every brick a truncated ray crosses */

/* Storage type of the per-incidence messages of a camera pass: the
 * voxel->ray log message ln n and the ray->voxel message p*l.  Scans and
 * accumulation are FP32 (Kahan-compensated along rays) in every mode; the
 * per-(camera, voxel) messages M_c and the beliefs stay FP32. */
#define MRF_MSG_F32 0        /* both FP32 */
#define MRF_MSG_F16 1        /* both FP16 */
#define MRF_MSG_BF16 2       /* both BF16 */
#define MRF_MSG_MIXED_F16 3  /* ln n FP32, p*l FP16 */
#define MRF_MSG_MIXED_BF16 4 /* ln n FP32, p*l BF16 */

/* Visibility of the pre pass.  The reference's pre pass runs once per active
 * brick and restores its visibility carry only when the accumulated tau is
 * non-zero (gvdbkernels.cu:708-715); its FP32 arithmetic flushes denormals
 * (--use_fast_math), so in front of a surface, where every tau underflows,
 * each brick starts again from V = 1, and T and the infinity term are summed
 * with these brick-local visibilities.  The post pass always carries its
 * visibility along the whole ray (gvdbkernels.cu:844-856). */
#define MRF_VIS_CARRY 0       /* the pre pass carries its visibility along the whole ray, as the post pass */
#define MRF_VIS_BRICK_LOCAL 1 /* the reference's: restart at every brick entry until a tau reaches 2^-126
                                 (the default of mrf_params_default) */

typedef struct mrf_ctx mrf_ctx;

/* Mirrors gvdb_params_t (include/mrfmap/GVDBParams.h:11-215). */
typedef struct mrf_params {
  float K[9];                  /* intrinsics, row-major 3x3 */
  int rows, cols;              /* image size */
  float res;                   /* voxel size (m) */
  float prior;                 /* prior occupancy probability */
  float dims[3];               /* map extent (m) */
  float occ_thresh, acc_thresh;
  int max_iters;               /* BP iterations when perform_inference gets max_iters <= 0 */
  int lookup_n;                /* polynomial coefficients per lookup cell (2 or 3) */
  float const_sigmasq_poly[3]; /* sigma(z) = p0 + p1 z + p2 z^2 (used as sigma, not sigma^2) */
  float const_bias_poly[3];    /* bias(z) likewise */
  const float* depth_sigma_lookup; /* optional ceil(rows/20)*ceil(cols/20)*lookup_n table */
  const float* depth_bias_lookup;  /* optional, same shape; NULL -> constant polynomials */
  int activation;              /* MRF_ACT_* */
  int msg_dtype;               /* MRF_MSG_* */
  int mask_848_border;         /* blank 24/20 px borders of 848x480 frames (GVDBImage.cpp:35-40) */
  int brick_local_vis;         /* MRF_VIS_*: visibility of the pre pass (the sum T of a ray's taus) */
} mrf_params;

/* Per-call statistics of mrf_perform_inference. */
typedef struct mrf_stats {
  int iters;                   /* BP iterations run */
  int converged;               /* residual fell below tol */
  double residual;             /* max |delta marginal| of the last iteration (FP64 reduction) */
  int64_t updates_per_iter;    /* ray-msg updates per iteration (voxel incidences) */
  int64_t pairs_per_iter;      /* (camera, touched voxel) pairs per iteration */
  int64_t rays;                /* rays with at least one incidence */
  int64_t list_records;        /* incidence records incl. gap records */
  double t_geometry_ms;        /* activation + traversal + pair build (0 if cached) */
  double t_iters_ms;           /* wall time of all iterations (device events) */
  double t_ray_kernel_ms;      /* summed device time of the ray-message kernels; -1 unless the
                                  launch mode is MRF_LAUNCH_TIMED */
  double t_epilogue_ms;        /* summed device time of the epilogue kernels; same */
  double t_total_ms;           /* wall time of the whole call */
  int ray_sched;               /* ray-kernel schedule in use (records in flight: front x window) */
  double t_autotune_ms;        /* time spent choosing it (first call of a context) */
  int64_t window_updates_per_iter; /* updates whose message depends on the ray; the others lie where
                                      the depth likelihood is exactly 0 in FP32 and carry the constant
                                      message F_EPS * l, summed per voxel once per traversal */
  int graph;                   /* 1: iterations replayed from a captured HIP graph, 0: launched one by one
                                  (MRF_LAUNCH_TIMED, MRF_NO_GRAPH=1, or capture unavailable) */
  int warm;                    /* learned warm start: 0 not used, 1 applied, 2 applied and undone
                                  because the residual grew (plain BP from the first iteration on) */
  int64_t warm_voxels;         /* voxels the warm start predicted */
  double t_warm_ms;            /* features + MLP + warm-start update, wall time */
  double ref_error;            /* max-abs marginal difference to the mrf_set_reference map after the
                                  last iteration, -1 without a reference */
  int64_t static_front_updates; /* front updates whose visibility terms stop changing after the first two
                                   iterations of a call (their pair is a front-only pair at a voxel no
                                   camera reaches with window records); from the third iteration on
                                   their stored sums are used instead of re-reading them */
  int64_t front_slots;         /* record slots of the sliced front regions (each slice padded to its
                                  longest ray), summed over the cameras */
  int64_t window_slots;        /* the same for the window regions */
  /* summed device times of each kernel of the camera passes, -1 unless the
   * launch mode is MRF_LAUNCH_TIMED: voxel->ray messages (k_pair_msg), front
   * visibility sums (k_front_sums), window passes (k_ray_window), segment
   * sums (k_seg_sum), per-pair update (k_seg_update) */
  double t_pair_msg_ms, t_front_ms, t_window_ms, t_seg_sum_ms, t_seg_update_ms;
} mrf_stats;

/* gvdb_params_t() defaults (GVDBParams.h:32-55). */
void mrf_params_default(mrf_params* p);

/* Replaces GVDBInference(create_context, use_gl) + the process-global
 * gvdb_params_ (GVDBInference.h:36-44, GVDBMRFMap.cpp:23-93). */
int mrf_create(const mrf_params* p, mrf_ctx** out);
/* Replaces ~GVDBInference (GVDBInference.h:46-61). */
void mrf_destroy(mrf_ctx* ctx);

/* Replaces GVDBInference::add_camera_with_depth (GVDBInference.cpp:58-64). */
int mrf_add_camera_with_depth(mrf_ctx* ctx, const float pose[16], const float* depth, int rows, int cols,
                              int layout, uint32_t* cam_id);
/* Same with per-camera intrinsics (K row-major 3x3) for uneven keyframes; no
 * counterpart in the reference, whose GVDBImage shares one gvdb_params_.K. */
int mrf_add_camera_with_depth_k(mrf_ctx* ctx, const float pose[16], const float K[9], const float* depth,
                                int rows, int cols, int layout, uint32_t* cam_id);
/* Replaces GVDBInference::set_pose (GVDBInference.cpp:66-82): the camera's
 * rays are re-traversed at the next perform_inference. */
int mrf_set_pose(mrf_ctx* ctx, uint32_t cam_id, const float pose[16], int layout);
int mrf_num_cameras(mrf_ctx* ctx, uint32_t* n);

/* Replaces GVDBInference::perform_inference[_ids] (GVDBInference.cpp:84-178).
 * ids == NULL runs every camera.  max_iters <= 0 uses params.max_iters.
 * tol > 0 stops early once the max marginal change of an iteration is below
 * tol.  Per-camera messages are reset at every call, beliefs persist
 * (GVDBInference.cpp:128-134). */
int mrf_perform_inference(mrf_ctx* ctx, const uint32_t* ids, int n_ids, int max_iters, double tol,
                          mrf_stats* out);

/* Number of active bricks (rows of get_occupied_bricks / get_alphas). */
int mrf_num_bricks(mrf_ctx* ctx, int64_t* n);
/* Replaces GVDBMRFMap::get_occupied_bricks(0) (GVDBMRFMap.cpp:137-152):
 * brick centres in voxel units, n x 3 row-major. */
int mrf_get_occupied_bricks(mrf_ctx* ctx, float* centers, int64_t n);
/* Replaces GVDBMRFMap::get_alphas (GVDBMRFMap.cpp:154-208): rows x cols must
 * be num_bricks x 512 (else MRF_E_SHAPE); voxels never updated read 0. */
int mrf_get_alphas(mrf_ctx* ctx, float* out, int64_t rows, int64_t cols);
/* Belief log-odds B per voxel, same shape/order as mrf_get_alphas.  No
 * counterpart (the reference keeps B in a half2 atlas channel). */
int mrf_get_beliefs(mrf_ctx* ctx, float* out, int64_t rows, int64_t cols);
/* Camera cam_id's summed chord length C per voxel, same shape/order as
 * mrf_get_alphas (0 where its rays do not pass): the reference's CUM_LENS
 * channel (GVDBCommon.cuh:37-44) after that camera's cumulative-length pass
 * (gvdbkernels.cu:519-571, 1276-1302), read back as get_channel_kernel does
 * (gvdbkernels.cu:360-373).  Exact sums rounded to float; C > 0.1 holds
 * exactly where it holds for the exact sum.  MRF_E_RANGE for a bad id. */
int mrf_get_cum_lengths(mrf_ctx* ctx, uint32_t cam_id, float* out, int64_t rows, int64_t cols);

/* Map evaluation from a camera: every pixel's ray is walked through the
 * current marginals to the end of the map.  Replaces
 * GVDBInference::compute_likelihood, get_likelihood_image[_at_pose],
 * get_expected_depth_image and get_accuracy_image_at_pose
 * (GVDBInference.cpp:180-232, GVDBInference.h:86-102), and the simple
 * accuracy of GVDBMapLikelihoodEstimator; kernels gvdbkernels.cu:135-354,
 * 1066-1134, 1246-1274, 1332-1363, reductions gvdbhelper.cu:10-63.  Images
 * are rows x cols float, row-major, NULL to skip:
 *   loglik          log p(z | map) per pixel, 0 where the depth is invalid;
 *   expected_depth  sum of ray weights x voxel depth / (1 - vis_inf + F_EPS);
 *   accuracy        1 / 0 when the ray's weight peak lies within / beyond
 *                   1 sigma of the measurement, -1 invalid or no peak;
 *   simple_accuracy first voxel with marginal >= acc_thresh within one voxel
 *                   size of the measurement 1, elsewhere 2, none before it 0,
 *                   invalid -1. */
typedef struct mrf_eval {
  double log_likelihood;      /* sum of the per-pixel log-likelihoods (sum_likelihood) */
  double accuracy;            /* #accuracy 1 / (#pixels - #accuracy -1) (compute_accuracy) */
  double simple_accuracy;     /* (#simple 1 + #simple -1) / #pixels (compute_tp_fp_tn_fn) */
  int64_t accuracy_counts[3]; /* pixels with accuracy -1, 0, 1 */
  int64_t simple_counts[4];   /* pixels with simple accuracy -1 (tn), 0 (fn), 1 (tp), 2 (fp) */
} mrf_eval;
int mrf_evaluate_camera(mrf_ctx* ctx, uint32_t cam_id, float* loglik, float* expected_depth, float* accuracy,
                        float* simple_accuracy, mrf_eval* out);
/* Same from any pose and depth image (K NULL: the params' intrinsics; depth
 * NULL: only expected_depth is meaningful). */
int mrf_evaluate_at_pose(mrf_ctx* ctx, const float pose[16], const float K[9], const float* depth, int rows,
                         int cols, int layout, float* loglik, float* expected_depth, float* accuracy,
                         float* simple_accuracy, mrf_eval* out);

/* Loads an external occupancy map (GVDBMapLikelihoodEstimator::load_map,
 * GVDBMapLikelihoodEstimator.cpp:168-240, fed by
 * GVDBOctomapWrapper::push_to_gvdb_volume, GVDBOctomapWrapper.cpp:45-74):
 * points n x 3 row-major in voxel units (world / res + (dims_x/2, dims_y/2, 0)),
 * occupancy probabilities n.  Bricks within one voxel of each point inside
 * the map box are activated; the voxel holding a point gets belief
 * logit(occupancy) (the last point of a voxel wins) and reads back as that
 * marginal, clamped to [F_EPS, 1 - F_EPS]. */
int mrf_load_map(mrf_ctx* ctx, const float* points, const float* occupancies, int64_t n);
/* Thresholds of the map evaluation (GVDBMapLikelihoodEstimator::set_acc_thresh). */
int mrf_set_eval_thresholds(mrf_ctx* ctx, float occ_thresh, float acc_thresh);
/* ---- Learned warm start (no counterpart in the reference, whose BP always
 * starts from the prior with all camera messages 0, GVDBInference.cpp:128-134).
 *
 * After the first BP iterations every voxel with C_c > 0.1 for some camera
 * gets MRF_MLP_FEATURES features: 0 log1p(depth endpoints in the voxel),
 * 1 log1p(summed chord length of all rays through it), 2 and 3 the mean and
 * the max over cameras of the camera messages M_c / 9.2102404, 4 the
 * Euclidean distance (voxels) to the nearest voxel holding an endpoint,
 * capped at 4, / 4, and 5..31 the beliefs of the 3x3x3 block around it
 * (x slowest) / 9.2102404 clamped to [-3, 3], the prior outside active bricks.
 * A 32-64-64-1 GELU MLP (BF16 matrix cores, FP32 accumulation) predicts the
 * converged belief log-odds; the belief moves there and every camera
 * message of the voxel by the same share of the change, so
 * B = logit(prior) + sum_c M_c keeps holding.  MLP parameters are
 * MRF_MLP_PARAMS floats: W1 64x32, b1 64, W2 64x64, b2 64, w3 64, b3 1,
 * matrices row-major. */
#define MRF_MLP_FEATURES 32
#define MRF_MLP_PARAMS 6337

typedef struct mrf_train_opts {
  int epochs;      /* passes over the samples */
  int batch;       /* samples per Adam step */
  float lr;        /* Adam step size */
  float beta1, beta2, eps;
  float lr_decay;  /* lr multiplied by this after every epoch */
  uint64_t seed;   /* sample order */
} mrf_train_opts;

void mrf_train_opts_default(mrf_train_opts* o);
/* perform_inference with the warm start applied after warm_after plain
 * iterations (features of that state; the network must have been trained on
 * features taken after as many iterations).  With growth > 0 the warm start
 * is undone -- beliefs and messages return to their state before it and
 * plain BP continues -- when, from the second warm iteration on, a residual
 * exceeds both growth x the previous iteration's residual and 0.05. */
int mrf_perform_inference_warm(mrf_ctx* ctx, const float* params, int warm_after, int max_iters, double tol,
                               double growth, mrf_stats* out);
/* Features of the current state (call after perform_inference): *n rows of
 * MRF_MLP_FEATURES floats (row-major) and their pool indices
 * (slot * 512 + voxel, the order of mrf_get_alphas), in pool order.  With X
 * and vox null only *n is set. */
int mrf_features(mrf_ctx* ctx, float* X, uint32_t* vox, int64_t cap, int64_t* n);
/* (pool index, predicted log-odds) of the last warm start or feature call. */
int mrf_warm_predictions(mrf_ctx* ctx, uint32_t* vox, float* pred, int64_t cap, int64_t* n);
/* The MLP on the GPU for host buffers: y[n] = f(X[n x 32]); gradient of
 * 1/(2n) sum (y - t)^2 into grad[MRF_MLP_PARAMS]; Adam training of params
 * in place (epoch_loss[epochs]: mean loss over each epoch, may be null).
 * Deterministic: the same inputs give the same bits. */
int mrf_mlp_forward(mrf_ctx* ctx, const float* params, const float* X, int64_t n, float* y);
int mrf_mlp_grad(mrf_ctx* ctx, const float* params, const float* X, const float* t, int64_t n, float* grad,
                 double* loss);
int mrf_mlp_train(mrf_ctx* ctx, float* params, const float* X, const float* t, int64_t n, const mrf_train_opts* o,
                  double* epoch_loss);

/* mrf_mlp_grad without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving or training loops, harnesses that forbid hipMalloc).
 * params [MRF_MLP_PARAMS], X [n x 32], t [n], grad [MRF_MLP_PARAMS] and loss
 * [1] (may be NULL) are DEVICE pointers of the context's GPU; the result is
 * bit-identical to mrf_mlp_grad on the same values.  The per-wave partial
 * gradients live in a fixed workspace of the library (12.4 MiB per GPU, held
 * until the process ends), so n <= MRF_MLP_NOALLOC_MAX_SAMPLES (else
 * MRF_E_LIMIT).  The work is enqueued on `stream` (a hipStream_t; NULL: the
 * context's stream) and not waited for.  Every call on one GPU shares that
 * workspace: calls that overlap on one GPU (two streams, two contexts) race. */
#define MRF_MLP_NOALLOC_MAX_SAMPLES 16384
int mrf_mlp_grad_noalloc(mrf_ctx* ctx, const float* params, const float* X, const float* t, int64_t n, float* grad,
                         double* loss, void* stream);

/* Reference marginals (rows = num_bricks, cols = 512, the layout of
 * mrf_get_alphas; call after mrf_prepare) for measuring convergence to a
 * known map, e.g. an FP64 or tightly converged solution: after every
 * iteration the GPU reduces max |alpha - ref| into mrf_stats.ref_error, and
 * with stop_at > 0 perform_inference returns (converged = 1) once it is
 * <= stop_at.  rows = 0 removes the reference.  No counterpart in the
 * reference. */
int mrf_set_reference(mrf_ctx* ctx, const float* alphas, int64_t rows, int64_t cols, double stop_at);
/* Per-iteration residual and reference error (-1 without a reference) of
 * the last perform_inference call: *n iterations, the first min(cap, *n)
 * copied (either array may be null). */
int mrf_get_trace(mrf_ctx* ctx, double* resid, double* err, int cap, int* n);

/* MRF_LAUNCH_GRAPH or MRF_LAUNCH_TIMED (no counterpart: GVDB launched its
 * kernels one by one).  The environment variable MRF_NO_GRAPH=1 makes
 * MRF_LAUNCH_TIMED the default of new contexts.  Results are bit-identical. */
int mrf_set_launch_mode(mrf_ctx* ctx, int mode);

/* Traversal lists of one camera (for parity tests and tools).  No
 * counterpart: GVDB recomputed the DDA in every pass.  Ray records are
 * 6 x uint32 {off, n, z0 bits, inv_norm bits, depth bits, cell | ran_out<<31};
 * incidence records are 2 x uint32 {voxel or 0xFFFFFFFF for a gap, length bits}.
 * Call with NULL buffers to get the sizes.  Valid after perform_inference
 * (or mrf_prepare). */
int mrf_prepare(mrf_ctx* ctx);
int mrf_get_traversal(mrf_ctx* ctx, uint32_t cam_id, int64_t* n_rays, int64_t* n_records, uint32_t* rays,
                      uint32_t* records);

const char* mrf_last_error(mrf_ctx* ctx);

/* Returns to the HIP runtime the large device blocks (32 MiB and up) that
 * the library keeps after contexts free them, for reuse by later contexts
 * (the runtime takes seconds to provide tens of GB again after such a
 * free).  No counterpart in the reference; for callers that share the
 * device with other allocators.  The cache holds at most 3/4 of the
 * device's memory and empties itself when an allocation fails;
 * MRF_NO_DEVICE_CACHE=1 in the environment switches it off. */
int mrf_release_device_cache(void);

#ifdef __cplusplus
}
#endif
#endif
