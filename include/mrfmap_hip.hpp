/*
 * Header-only C++ wrapper of libmrfmap_hip with the method names of the
 * reference's GVDBInference (include/mrfmap/GVDBInference.h:36-102), for C++
 * callers such as a viewer worker loop (Viewer.cpp:56-57, 961-966).
 *
 * Matrices are plain pointers: an Eigen::MatrixXf is column-major, so pass
 * m.data() with MRF_COL_MAJOR (the default here); a row-major
 * Eigen::Matrix<float, Dynamic, Dynamic, RowMajor> goes with MRF_ROW_MAJOR.
 * Errors throw mrfmap_hip::Error (a std::runtime_error) where the reference
 * throws or would fail; set_pose keeps the reference's silent no-op for an
 * unknown camera id (GVDBInference.cpp:66-82) and reports it by its return
 * value.  Not thread-safe, like the reference: one caller thread per object.
 */
#ifndef MRFMAP_HIP_HPP
#define MRFMAP_HIP_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "mrfmap_hip.h"

namespace mrfmap_hip {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void check(int rc, mrf_ctx* c, const char* where) {
  if (rc != MRF_OK) {
    const char* msg = c ? mrf_last_error(c) : "";
    throw Error(rc, std::string(where) + ": error " + std::to_string(rc) + (msg && *msg ? std::string(": ") + msg : ""));
  }
}

/* Gives the large device blocks that destroyed contexts left with the
 * library back to the HIP runtime (mrf_release_device_cache). */
inline void release_device_cache() { check(mrf_release_device_cache(), nullptr, "mrf_release_device_cache"); }

/* gvdb_params_t with the reference defaults (GVDBParams.h:32-55). */
struct Params : mrf_params {
  Params() { mrf_params_default(this); }
};

class GVDBInference {
 public:
  explicit GVDBInference(const mrf_params& p) { check(mrf_create(&p, &ctx_), nullptr, "mrf_create"); }
  ~GVDBInference() { mrf_destroy(ctx_); }
  GVDBInference(const GVDBInference&) = delete;
  GVDBInference& operator=(const GVDBInference&) = delete;

  /* add_camera_with_depth(pose 4x4, depth rows x cols) (GVDBInference.cpp:58-64); returns the camera id */
  uint32_t add_camera_with_depth(const float* pose, const float* depth, int rows, int cols,
                                 int layout = MRF_COL_MAJOR) {
    uint32_t id = 0;
    check(mrf_add_camera_with_depth(ctx_, pose, depth, rows, cols, layout, &id), ctx_, "add_camera_with_depth");
    return id;
  }
  /* same with this camera's own intrinsics (row-major 3x3) */
  uint32_t add_camera_with_depth(const float* pose, const float* K, const float* depth, int rows, int cols,
                                 int layout = MRF_COL_MAJOR) {
    uint32_t id = 0;
    check(mrf_add_camera_with_depth_k(ctx_, pose, K, depth, rows, cols, layout, &id), ctx_, "add_camera_with_depth");
    return id;
  }
  /* set_pose(id, pose) (GVDBInference.cpp:66-82): false for an unknown id */
  bool set_pose(uint32_t id, const float* pose, int layout = MRF_COL_MAJOR) {
    const int rc = mrf_set_pose(ctx_, id, pose, layout);
    if (rc == MRF_E_RANGE) return false;
    check(rc, ctx_, "set_pose");
    return true;
  }
  /* perform_inference() (GVDBInference.cpp:84-90): every camera, params.max_iters
   * iterations unless max_iters > 0; tol > 0 stops at that marginal change */
  mrf_stats perform_inference(int max_iters = 0, double tol = 0.0) {
    mrf_stats s;
    check(mrf_perform_inference(ctx_, nullptr, 0, max_iters, tol, &s), ctx_, "perform_inference");
    return s;
  }
  /* perform_inference_ids(ids, dry_run) (GVDBInference.cpp:100-178); dry_run
   * only builds the traversal */
  mrf_stats perform_inference_ids(const std::vector<uint32_t>& ids, bool dry_run = false, int max_iters = 0,
                                  double tol = 0.0) {
    mrf_stats s = {};
    if (dry_run) {
      check(mrf_prepare(ctx_), ctx_, "perform_inference_ids");
      return s;
    }
    check(mrf_perform_inference(ctx_, ids.data(), (int)ids.size(), max_iters, tol, &s), ctx_,
          "perform_inference_ids");
    return s;
  }
  uint32_t num_cameras() {
    uint32_t n = 0;
    check(mrf_num_cameras(ctx_, &n), ctx_, "num_cameras");
    return n;
  }
  int64_t num_bricks() {
    int64_t n = 0;
    check(mrf_num_bricks(ctx_, &n), ctx_, "num_bricks");
    return n;
  }
  /* get_occupied_bricks(0) (GVDBMRFMap.cpp:137-152): n x 3 brick centres in
   * voxel units, row-major */
  std::vector<float> get_occupied_bricks(int lev = 0) {
    (void)lev;
    const int64_t n = num_bricks();
    std::vector<float> c((size_t)n * 3);
    check(mrf_get_occupied_bricks(ctx_, c.data(), n), ctx_, "get_occupied_bricks");
    return c;
  }
  /* get_alphas(out) (GVDBMRFMap.cpp:154-208): out must be num_bricks x 512,
   * row-major; a mismatch throws as the reference does */
  void get_alphas(float* out, int64_t rows, int64_t cols) {
    check(mrf_get_alphas(ctx_, out, rows, cols), ctx_, "get_alphas");
  }
  std::vector<float> get_alphas() {
    const int64_t n = num_bricks();
    std::vector<float> a((size_t)n * 512);
    get_alphas(a.data(), n, 512);
    return a;
  }
  /* camera id's CUM_LENS channel (GVDBCommon.cuh:37-44): its summed chord
   * length per voxel, num_bricks x 512, 0 where its rays do not pass */
  std::vector<float> get_cum_lengths(uint32_t id) {
    const int64_t n = num_bricks();
    std::vector<float> a((size_t)n * 512);
    check(mrf_get_cum_lengths(ctx_, id, a.data(), n, 512), ctx_, "get_cum_lengths");
    return a;
  }
  /* compute_likelihood(id) / get_likelihood_image(id, img) (GVDBInference.cpp:180-193):
   * sum of the per-pixel log-likelihoods; img (rows x cols, row-major) may be null */
  float compute_likelihood(uint32_t id, float* img = nullptr) {
    mrf_eval e;
    check(mrf_evaluate_camera(ctx_, id, img, nullptr, nullptr, nullptr, &e), ctx_, "compute_likelihood");
    return (float)e.log_likelihood;
  }
  /* get_expected_depth_image(id, img) (GVDBInference.cpp:226-231) */
  void get_expected_depth_image(uint32_t id, float* img) {
    check(mrf_evaluate_camera(ctx_, id, nullptr, img, nullptr, nullptr, nullptr), ctx_, "get_expected_depth_image");
  }
  /* get_likelihood_image_at_pose / get_accuracy_image_at_pose: every image and
   * reduction of one evaluation from any pose (K null: the params' intrinsics) */
  mrf_eval evaluate_at_pose(const float* pose, const float* depth, int rows, int cols, float* loglik,
                            float* expected_depth, float* accuracy, float* simple_accuracy, const float* K = nullptr,
                            int layout = MRF_COL_MAJOR) {
    mrf_eval e;
    check(mrf_evaluate_at_pose(ctx_, pose, K, depth, rows, cols, layout, loglik, expected_depth, accuracy,
                               simple_accuracy, &e),
          ctx_, "evaluate_at_pose");
    return e;
  }
  mrf_ctx* handle() { return ctx_; }

 private:
  mrf_ctx* ctx_ = nullptr;
};

}  // namespace mrfmap_hip

#endif
