// Runs one BP inference through the C++ surface of include/mrfmap_hip.hpp
// (the GVDBInference method names) on a workload read from a file, and
// writes the marginals and brick centres.  Built in-tree by the package
// build (mrfmap_amd/build.py); tests/test_gpu_cpp.py writes the input and
// checks the output against the Python path and the oracle.
//
// input (little-endian): int32 ncams, rows, cols, iters, activation,
// lookup_n, brick_local_vis; float32 res, prior, dims[3], sigma poly[3],
// bias poly[3]; per camera float32 pose[16] (row-major), K[9] (row-major),
// depth[rows * cols].
// output: int64 num_bricks, float32 alphas[num_bricks * 512], float32
// brick centres[num_bricks * 3], int32 iterations run.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "mrfmap_hip.hpp"

namespace {

template <typename T>
bool read(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input output\n", argv[0]);
    return 2;
  }
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int32_t hdr[7];
  float fp[11];
  if (!read(in, hdr, 7) || !read(in, fp, 11)) return 2;
  const int ncams = hdr[0], rows = hdr[1], cols = hdr[2];
  mrfmap_hip::Params p;
  p.rows = rows;
  p.cols = cols;
  p.max_iters = hdr[3];
  p.activation = hdr[4];
  p.lookup_n = hdr[5];
  p.brick_local_vis = hdr[6];
  p.res = fp[0];
  p.prior = fp[1];
  for (int a = 0; a < 3; ++a) {
    p.dims[a] = fp[2 + a];
    p.const_sigmasq_poly[a] = fp[5 + a];
    p.const_bias_poly[a] = fp[8 + a];
  }
  std::vector<float> pose(16), K(9), depth((size_t)rows * cols);
  try {
    bool first = true;
    std::vector<std::vector<float>> poses, Ks, depths;
    for (int c = 0; c < ncams; ++c) {
      if (!read(in, pose.data(), 16) || !read(in, K.data(), 9) || !read(in, depth.data(), depth.size())) {
        std::fprintf(stderr, "input shorter than its header says\n");
        return 2;
      }
      if (first) {
        for (int i = 0; i < 9; ++i) p.K[i] = K[i];
        first = false;
      }
      poses.push_back(pose);
      Ks.push_back(K);
      depths.push_back(depth);
    }
    const bool at_end = std::fgetc(in) == EOF;  // nothing left over: the layout matched
    std::fclose(in);
    if (!at_end) {
      std::fprintf(stderr, "input longer than its header says\n");
      return 2;
    }
    mrf_stats st;
    std::vector<float> a, centres;
    int64_t nb = 0;
    uint32_t ncams_ctx = 0;
    {
      mrfmap_hip::GVDBInference inf(p);
      for (int c = 0; c < ncams; ++c)
        inf.add_camera_with_depth(poses[c].data(), Ks[c].data(), depths[c].data(), rows, cols, MRF_ROW_MAJOR);
      st = inf.perform_inference();
      a = inf.get_alphas();
      centres = inf.get_occupied_bricks();
      nb = inf.num_bricks();
      ncams_ctx = inf.num_cameras();
    }
    mrfmap_hip::release_device_cache();  // the context's kept device blocks back to the runtime
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    const int32_t iters = st.iters;
    const bool ok = std::fwrite(&nb, sizeof(nb), 1, out) == 1 && std::fwrite(a.data(), 4, a.size(), out) == a.size() &&
                    std::fwrite(centres.data(), 4, centres.size(), out) == centres.size() &&
                    std::fwrite(&iters, 4, 1, out) == 1;
    std::fclose(out);
    std::printf("cameras %u bricks %lld iterations %d\n", ncams_ctx, (long long)nb, st.iters);
    return ok ? 0 : 2;
  } catch (const mrfmap_hip::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
}
