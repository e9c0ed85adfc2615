// C ABI of libmrfmap_hip: context, keyframes, geometry build and the BP
// schedule.  Host orchestration only; the kernels live in traverse.hip,
// bp.hip and scan.hip.
//
// Schedule (GVDBInference::perform_inference_ids, src/GVDBInference.cpp:100-178):
//   reset M_c of the requested cameras;
//   repeat max_iters times (or until the FP64 residual drops below tol):
//     for every camera c in order (Gauss-Seidel):
//       ray kernel (pre + infinity term + post fused)  -> S
//       epilogue over c's (voxel, C) pairs             -> B, M_c
//     residual = max |marginal(B) - marginal(B_prev)|
// The per-camera summed chord length C (render_cumulative_length_kernel,
// recomputed every camera pass in the reference) depends on geometry only
// and is built once per traversal.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <unordered_map>
#include <chrono>
#include <random>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mrfmap_hip.h"
#include "mrf_kernels.h"

using namespace mrf;

namespace {

struct Camera {
  float pose[16];
  float K[9];
  int rows = 0, cols = 0;
  float* d_depth = nullptr;
  FrontRec* d_front = nullptr;     // one per lane of every slice (tiled pixel order)
  RayRec* d_lanes = nullptr;       // one per lane of every slice (tiled pixel order)
  uint32_t* d_len_f = nullptr;     // front records per lane of each slice
  uint32_t* d_len_w = nullptr;     // window records per lane of each slice
  uint32_t nslices = 0;
  uint32_t nrays = 0;              // rays with at least one record
  uint2* d_inc = nullptr;          // sliced records {pair | gap, length}: front region, then window region
  uint32_t* d_walkn = nullptr;     // full-ray activation: bound of the records per tiled pixel
  uint8_t* d_lnn = nullptr;        // per pair: voxel -> ray message log2 n (input-side message type)
  uint2* d_dyn = nullptr;          // front-only pairs at voxels where some camera has window records: {pair, voxel}
  uint32_t ndyn = 0;
  float* d_partial = nullptr;      // front partial sums (64 per chunk), kept between iterations
  uint4* d_dyn_desc = nullptr;        // descriptors of the front chunks whose partial sums keep changing
  uint32_t ndyn_chunks = 0;
  uint32_t* d_fdmask = nullptr;       // per front chunk and lane: its changing records (bits; in d_partial)
  float* d_fstatic = nullptr;         // per front chunk and lane: sum of its other records (in d_partial)
  uint2* d_dfront = nullptr;          // the changing chunks' changing records, compacted per lane (in d_dcr)
  uint2* d_dcr = nullptr;             // per changing chunk: {first row, rows} in d_dfront
  uint2* d_fspan = nullptr;           // per lane of every slice: first front record, front records to sum
  uint64_t nstatic = 0;            // live records in such chunks
  uint64_t nrec = 0, base_w = 0;   // padded records; first record of the window region
  uint64_t nupd = 0;               // live incidences
  uint64_t nwin = 0;               // live window incidences (window positions)
  uint64_t nlist = 0;              // records without the slice padding
  uint32_t* d_win_idx = nullptr;   // per window position (voxel order): the record's slot in the window region
  uint32_t* d_pair_vox = nullptr;
  float* d_pair_c = nullptr;       // summed chord length per pair
  float* d_pair_sf = nullptr;      // summed constant front messages per pair
  float* d_M = nullptr;            // camera messages M_c per pair (FP32)
  uint32_t npairs = 0, nwp = 0;    // pairs; pairs with window incidences (numbered first)
  uint32_t* d_wseg_off = nullptr;  // nwp + 1 window segment offsets
  uint64_t* d_heads = nullptr;     // window segment head bitmask
  uint32_t* d_tile_h0 = nullptr;   // window heads before each kTile tile
  uint32_t* d_chunk_off = nullptr; // nslices + 1: first front chunk of each slice
  uint4* d_chunk_desc = nullptr;
  uint32_t nchunks = 0;
  uint64_t* d_bstart = nullptr;    // MRF_VIS_BRICK_LOCAL: brick-entry word per window row of every slice
  uint8_t* d_tail = nullptr;       // MRF_VIS_BRICK_LOCAL: per lane, front records in the first window brick
};

}  // namespace

struct mrf_ctx {
  mrf_params p;
  std::vector<float> sig_tab, bia_tab;
  bool user_tables = false;
  int tab_rows = 0, tab_cols = 0;
  float *d_sig = nullptr, *d_bia = nullptr;
  GridGeom g;
  int nbricks = 0;
  float prior_logodds = 0.f;
  hipStream_t stream = nullptr;

  std::vector<int32_t> h_slot;   // brick directory: linear brick index -> slot or -1
  std::vector<int32_t> slot_lin; // slot -> linear brick index
  int32_t* d_slot = nullptr;
  uint8_t* d_bflags = nullptr;
  size_t cap_slots = 0;

  float* d_B = nullptr;
  float* d_A = nullptr;   // marginal at the end of the previous iteration
  uint8_t* d_upd = nullptr;
  float* d_out = nullptr; // readback staging
  void* d_Xo = nullptr;   // segment array of ray->voxel messages of the camera being processed
  size_t Xo_bytes = 0;
  // per-pair / per-tile scratch of a camera pass
  float* d_Spair = nullptr;
  size_t cap_pairs = 0;
  float *d_pf_v = nullptr, *d_pl_v = nullptr, *d_mid = nullptr;
  size_t cap_tiles = 0;
  // segment-build scratch (per record), per-pair scratch and the emitted lists
  uint32_t *d_keys_in = nullptr, *d_keys_out = nullptr, *d_vals_in = nullptr;  // window sort (values out: win_idx)
  uint32_t *d_x0 = nullptr, *d_x1 = nullptr;  // window head flags and their scan
  size_t cap_sort = 0;
  uint32_t* d_vflag = nullptr;  // 4 pool-sized tables: touched, window-touched and their scans
  size_t cap_vflag = 0;
  uint32_t* d_acc = nullptr;  // 6 words per pair: three 64-bit fixed-point chord-length sums
  size_t cap_acc = 0;
  uint2* d_sell0 = nullptr;
  size_t cap_sell0 = 0;
  // map evaluation scratch: 4 images + depth, FP64 block sums, class counts
  float* d_eval = nullptr;
  size_t cap_eval = 0;
  double* d_eval_part = nullptr;
  size_t cap_eval_part = 0;
  unsigned long long* d_eval_counts = nullptr;
  void* d_sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;

  // scratch
  uint32_t *d_counts = nullptr, *d_has = nullptr, *d_offs = nullptr, *d_rayidx = nullptr;
  size_t cap_pix = 0;
  uint64_t* d_scan_tmp = nullptr;
  size_t cap_scan = 0;
  void* d_fscan_tmp = nullptr;  // flag-scan scratch
  size_t fscan_bytes = 0;
  uint64_t* d_total = nullptr;  // [0] scan total, [1] residual bits, [2..9] counters, flags and totals
  uint32_t* d_all_pairs = nullptr;  // [0] CamPass::all_pairs, [1] CamPass::msg_all

  std::vector<Camera> cams;
  bool geom_dirty = true;
  int ray_sched = -1;      // chosen RaySched, -1 until autotuned
  // captured BP iteration: valid while no buffer it uses is reallocated
  uint64_t epoch = 0;      // bumped whenever geometry or pool buffers change
  hipGraphExec_t gexec = nullptr;
  uint64_t gexec_key = 0;
  int launch_mode = MRF_LAUNCH_GRAPH;
  bool graph_ok = true;    // false once a capture failed
  bool sched_forced = false;
  double autotune_ms = 0.0;
  // learned warm start: per-voxel feature accumulators, feature rows,
  // predictions, the fallback snapshot, MLP parameters and training scratch
  float *d_wpass = nullptr, *d_wmsum = nullptr, *d_wmmax = nullptr, *d_wdelta = nullptr;
  uint32_t *d_wnc = nullptr, *d_whits = nullptr, *d_wflag = nullptr, *d_wrow = nullptr;
  size_t cap_wvox = 0;
  int32_t* d_slot_lin = nullptr;
  size_t cap_slot_lin = 0;
  float* d_wX = nullptr;
  uint32_t* d_wvox = nullptr;
  float* d_wpred = nullptr;
  size_t cap_wrows = 0;
  uint32_t wrows = 0;  // rows of the last feature extraction
  // reference marginals for convergence measurements (mrf_set_reference)
  float* d_ref = nullptr;
  size_t ref_n = 0;
  double ref_stop = 0.0;
  unsigned long long* d_rerr = nullptr;
  std::vector<double> trace_resid, trace_err;  // per iteration of the last call
  void* scratch = nullptr;  // this device's copy of the fixed workspace (mrf_mlp_grad_noalloc)
  float* d_mlp = nullptr;
  float *d_snapB = nullptr, *d_snapM = nullptr;
  size_t cap_snapB = 0, cap_snapM = 0;
  float *d_tX = nullptr, *d_tT = nullptr, *d_tG = nullptr, *d_tM = nullptr, *d_tV = nullptr, *d_tPart = nullptr;
  uint32_t* d_tSel = nullptr;
  double *d_tLpart = nullptr, *d_tLoss = nullptr;
  size_t cap_trows = 0, cap_tpart = 0;
  std::string err;
};

namespace {

// Fixed workspace of mrf_mlp_grad_noalloc, the entry point that allocates
// nothing: the per-wave partial gradients (kMlpParams floats per wave) and
// loss sums (one double per wave) of k_mlp_grad for at most
// kNoAllocMaxSamples samples, i.e. mlp_grad_waves(kNoAllocMaxSamples) = 512
// waves.  Every GPU the library runs on holds its own copy until the process
// ends; host code reaches it only through the address hipGetSymbolAddress
// returns on that device (mrf_ctx::scratch), never through the array's name.
constexpr uint32_t kNoAllocMaxSamples = 16384;
constexpr size_t kNoAllocWaves = 512;
constexpr size_t kNoAllocPartBytes = kNoAllocWaves * (size_t)kMlpParams * sizeof(float);
constexpr size_t kNoAllocBytes = kNoAllocPartBytes + kNoAllocWaves * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// This device's address of the workspace, looked up once per device with
// that device current (mrf_create runs on the device its context uses).
hipError_t scratch_address(int dev, void** out) {
  *out = nullptr;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

int fail(mrf_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  (void)hipGetLastError();  // reported here; must not resurface in a later call's launch check
  return code;
}

#define HIPCHK(ctx, expr)                                                                         \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return fail(ctx, MRF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));             \
  } while (0)

// Process-wide cache of large device blocks.  A context's lists take tens
// of GB (85 GB for 64 keyframes); when they are freed and a new context
// allocates as much again, the runtime takes seconds to provide the memory
// (a 64-keyframe build: 0.86 s in a fresh process, 4.3-4.5 s after a
// context of that size was closed, whose frees took 2.4 s).  Blocks of at
// least kCacheMin bytes are therefore kept on free (after the device has
// finished with them) and handed out again for requests of up to their size
// and at least 4/5 of it.  The cache keeps at most 3/4 of the device's
// memory, gives everything back when an allocation fails, and is emptied by
// mrf_release_device_cache(); MRF_NO_DEVICE_CACHE=1 switches it off.
struct BlockCache {
  std::mutex mu;
  std::unordered_map<void*, std::pair<size_t, int>> live;  // large blocks in use: bytes, device
  std::multimap<std::pair<int, size_t>, void*> idle;        // (device, bytes) -> kept block
  size_t kept = 0;
};
constexpr size_t kCacheMin = size_t(32) << 20;

BlockCache& block_cache() {
  static BlockCache* bc = new BlockCache;  // never destroyed: frees may run during static destruction
  return *bc;
}

bool cache_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("MRF_NO_DEVICE_CACHE");
    return !(e && e[0] && e[0] != '0');
  }();
  return on;
}

// the cached blocks of device dev (all devices for dev < 0) back to the runtime; mu held
void release_idle(BlockCache& bc, int dev) {
  for (auto it = bc.idle.begin(); it != bc.idle.end();) {
    if (dev >= 0 && it->first.first != dev) {
      ++it;
      continue;
    }
    (void)hipFree(it->second);
    bc.kept -= it->first.second;
    it = bc.idle.erase(it);
  }
}

hipError_t dev_malloc(void** p, size_t bytes) {
  *p = nullptr;
  if (bytes < kCacheMin || !cache_enabled()) return hipMalloc(p, bytes);
  int dev = 0;
  (void)hipGetDevice(&dev);
  BlockCache& bc = block_cache();
  std::lock_guard<std::mutex> lock(bc.mu);
  auto it = bc.idle.lower_bound({dev, bytes});
  if (it != bc.idle.end() && it->first.first == dev && it->first.second - it->first.second / 5 <= bytes) {
    *p = it->second;
    bc.live[*p] = {it->first.second, dev};
    bc.kept -= it->first.second;
    bc.idle.erase(it);
    return hipSuccess;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && !bc.idle.empty()) {  // give the kept blocks back and try again
    (void)hipGetLastError();
    release_idle(bc, dev);
    e = hipMalloc(p, bytes);
  }
  if (e == hipSuccess) bc.live[*p] = {bytes, dev};
  return e;
}

void dev_free(void* p) {
  if (!p) return;
  BlockCache& bc = block_cache();
  {
    std::lock_guard<std::mutex> lock(bc.mu);
    auto it = bc.live.find(p);
    if (it != bc.live.end()) {
      const size_t bytes = it->second.first;
      const int dev = it->second.second;
      bc.live.erase(it);
      static size_t cap = 0;
      if (!cap) {
        size_t free_b = 0, total_b = 0;
        cap = hipMemGetInfo(&free_b, &total_b) == hipSuccess ? total_b / 4 * 3 : size_t(64) << 30;
      }
      if (bc.kept + bytes <= cap) {
        // the block's last uses end before another allocation can take it
        int cur = dev;
        (void)hipGetDevice(&cur);
        if (cur != dev) (void)hipSetDevice(dev);
        (void)hipDeviceSynchronize();
        if (cur != dev) (void)hipSetDevice(cur);
        bc.idle.insert({{dev, bytes}, p});
        bc.kept += bytes;
        return;
      }
    }
  }
  (void)hipFree(p);
}

template <typename T>
int dalloc(mrf_ctx* c, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = dev_malloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(c, MRF_E_NOMEM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
  }
  return MRF_OK;
}

template <typename T>
void dfree(T** p) {
  dev_free((void*)*p);
  *p = nullptr;
}

void to_row_major(const float* in, int rows, int cols, int layout, std::vector<float>& out) {
  out.resize((size_t)rows * cols);
  if (layout == MRF_COL_MAJOR) {
    for (int r = 0; r < rows; ++r)
      for (int c = 0; c < cols; ++c) out[(size_t)r * cols + c] = in[(size_t)c * rows + r];
  } else {
    std::memcpy(out.data(), in, out.size() * sizeof(float));
  }
}

CamGeom make_geom(const mrf_ctx* c, const Camera& cam) {
  CamGeom g;
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) g.R[i * 3 + j] = (double)cam.pose[i * 4 + j];
    g.t[i] = (double)cam.pose[i * 4 + 3];
  }
  g.fx = cam.K[0];
  g.fy = cam.K[4];
  g.cx = cam.K[2];
  g.cy = cam.K[5];
  g.rows = cam.rows;
  g.cols = cam.cols;
  g.lookup_u = (cam.cols + kLookupPx - 1) / kLookupPx;
  g.mask848 = c->p.mask_848_border;
  return g;
}

// (Re)build the sensor-model tables so that every camera's lookup cells exist.
int ensure_tables(mrf_ctx* c, int rows, int cols) {
  const int ln = c->p.lookup_n;
  if (c->user_tables) {
    if (rows != c->p.rows || cols != c->p.cols)
      return fail(c, MRF_E_INVALID, "per-pixel lookup tables require every image to be rows x cols of the params");
    return MRF_OK;
  }
  if (c->d_sig && rows <= c->tab_rows && cols <= c->tab_cols) return MRF_OK;
  c->tab_rows = std::max(rows, c->tab_rows);
  c->tab_cols = std::max(cols, c->tab_cols);
  const size_t cells = (size_t)((c->tab_rows + kLookupPx - 1) / kLookupPx) * ((c->tab_cols + kLookupPx - 1) / kLookupPx);
  c->sig_tab.assign(cells * ln, 0.f);
  c->bia_tab.assign(cells * ln, 0.f);
  for (size_t i = 0; i < cells; ++i)
    for (int j = 0; j < ln; ++j) {
      c->sig_tab[i * ln + j] = c->p.const_sigmasq_poly[j];
      c->bia_tab[i * ln + j] = c->p.const_bias_poly[j];
    }
  dfree(&c->d_sig);
  dfree(&c->d_bia);
  ++c->epoch;  // a captured iteration holds the old table pointers
  int rc;
  if ((rc = dalloc(c, &c->d_sig, c->sig_tab.size())) || (rc = dalloc(c, &c->d_bia, c->bia_tab.size()))) return rc;
  HIPCHK(c, hipMemcpy(c->d_sig, c->sig_tab.data(), c->sig_tab.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->d_bia, c->bia_tab.data(), c->bia_tab.size() * 4, hipMemcpyHostToDevice));
  return MRF_OK;
}

int ensure_pixel_scratch(mrf_ctx* c, size_t npix) {
  if (npix <= c->cap_pix) return MRF_OK;
  dfree(&c->d_counts);
  dfree(&c->d_has);
  dfree(&c->d_offs);
  dfree(&c->d_rayidx);
  int rc;
  if ((rc = dalloc(c, &c->d_counts, npix)) || (rc = dalloc(c, &c->d_has, npix)) ||
      (rc = dalloc(c, &c->d_offs, npix)) || (rc = dalloc(c, &c->d_rayidx, npix)))
    return rc;
  c->cap_pix = npix;
  return MRF_OK;
}

int ensure_scan(mrf_ctx* c, size_t n) {
  const size_t w = scan_tmp_words(n);
  if (w <= c->cap_scan) return MRF_OK;
  dfree(&c->d_scan_tmp);
  int rc = dalloc(c, &c->d_scan_tmp, w);
  if (rc) return rc;
  c->cap_scan = w;
  return MRF_OK;
}

uint64_t read_total(mrf_ctx* c) {
  uint64_t v = 0;
  (void)hipMemcpyAsync(&v, c->d_total, sizeof(v), hipMemcpyDeviceToHost, c->stream);
  (void)hipStreamSynchronize(c->stream);
  return v;
}

// Grow the voxel pool to hold nslots bricks, preserving B and the update flags.
int ensure_pool(mrf_ctx* c, size_t nslots) {
  if (nslots <= c->cap_slots) return MRF_OK;
  size_t cap = std::max(nslots, c->cap_slots + c->cap_slots / 2);
  const size_t nv = cap * kBrickVox, old = c->cap_slots * kBrickVox;
  float *B = nullptr, *A = nullptr, *O = nullptr;
  uint8_t* U = nullptr;
  int rc;
  if ((rc = dalloc(c, &B, nv)) || (rc = dalloc(c, &A, nv)) || (rc = dalloc(c, &O, nv)) || (rc = dalloc(c, &U, nv))) {
    dfree(&B); dfree(&A); dfree(&O); dfree(&U);
    return rc;
  }
  launch_fill_f32(B, c->prior_logodds, nv, c->stream);
  launch_fill_u8(U, 0, nv, c->stream);
  if (old) {
    HIPCHK(c, hipMemcpyAsync(B, c->d_B, old * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(U, c->d_upd, old, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(&c->d_B); dfree(&c->d_A); dfree(&c->d_out); dfree(&c->d_upd);
  c->d_B = B; c->d_A = A; c->d_out = O; c->d_upd = U;
  c->cap_slots = cap;
  ++c->epoch;
  return MRF_OK;
}

int ensure_sort(mrf_ctx* c, size_t n, int bits) {
  if (n > c->cap_sort) {
    n += n / 4;  // headroom: the next camera is likely a little larger, and a free waits for the device
    dfree(&c->d_keys_in); dfree(&c->d_keys_out); dfree(&c->d_vals_in);
    dfree(&c->d_x0); dfree(&c->d_x1);
    c->cap_sort = 0;
    int rc;
    const size_t nx = n + 1024;
    if ((rc = dalloc(c, &c->d_keys_in, nx)) || (rc = dalloc(c, &c->d_keys_out, nx)) ||
        (rc = dalloc(c, &c->d_vals_in, nx)) ||
        (rc = dalloc(c, &c->d_x0, nx)) || (rc = dalloc(c, &c->d_x1, nx)))
      return rc;
    c->cap_sort = n;
  }
  const size_t tb = window_sort_temp_bytes((uint32_t)n, bits);
  if (tb > c->sort_tmp_bytes) {
    dev_free(c->d_sort_tmp);
    c->d_sort_tmp = nullptr;
    c->sort_tmp_bytes = 0;
    if (dev_malloc(&c->d_sort_tmp, tb) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, MRF_E_NOMEM, "sort scratch allocation failed");
    }
    c->sort_tmp_bytes = tb;
  }
  return ensure_scan(c, n);
}

// Exclusive scan of 0/1 flags; the total lands in d_total[0].
int flag_scan(mrf_ctx* c, const uint32_t* in, uint32_t* out, size_t n, uint64_t* total = nullptr) {
  if (!total) total = c->d_total;
  if (n >= 0x7FFFFFFFull) {  // beyond the single-pass scan's item count: the general scan
    int rc = ensure_scan(c, n);
    if (rc) return rc;
    launch_exclusive_scan(in, out, n, c->d_scan_tmp, total, c->stream);
    return MRF_OK;
  }
  const size_t tb = flag_scan_temp_bytes(n);
  if (tb > c->fscan_bytes || !c->d_fscan_tmp) {
    dev_free(c->d_fscan_tmp);
    c->d_fscan_tmp = nullptr;
    c->fscan_bytes = 0;
    if (dev_malloc(&c->d_fscan_tmp, std::max<size_t>(tb, 256)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, MRF_E_NOMEM, "flag-scan scratch allocation failed");
    }
    c->fscan_bytes = std::max<size_t>(tb, 256);
  }
  HIPCHK(c, launch_flag_scan(in, out, n, c->d_fscan_tmp, c->fscan_bytes, total, c->stream));
  return MRF_OK;
}

// Exclusive scan of n uint64 values into out, total into *total.
int u64_scan(mrf_ctx* c, const uint64_t* in, uint64_t* out, size_t n, uint64_t* total) {
  if (n >= 0x7FFFFFFFull) return fail(c, MRF_E_LIMIT, "more than 2^31 bricks in the pool");
  const size_t tb = u64_scan_temp_bytes(n);
  if (tb > c->fscan_bytes || !c->d_fscan_tmp) {
    dev_free(c->d_fscan_tmp);
    c->d_fscan_tmp = nullptr;
    c->fscan_bytes = 0;
    if (dev_malloc(&c->d_fscan_tmp, std::max<size_t>(tb, 256)) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, MRF_E_NOMEM, "scan scratch allocation failed");
    }
    c->fscan_bytes = std::max<size_t>(tb, 256);
  }
  HIPCHK(c, launch_u64_scan(in, out, n, c->d_fscan_tmp, c->fscan_bytes, total, c->stream));
  return MRF_OK;
}

int ensure_vflags(mrf_ctx* c, size_t nvox) {
  if (nvox <= c->cap_vflag && c->d_vflag) return MRF_OK;
  dfree(&c->d_vflag);
  c->cap_vflag = 0;
  int rc = dalloc(c, &c->d_vflag, 4 * nvox);
  if (rc) return rc;
  c->cap_vflag = nvox;
  return ensure_scan(c, nvox);
}

int ensure_acc(mrf_ctx* c, size_t np) {
  if (np <= c->cap_acc && c->d_acc) return MRF_OK;
  np += np / 4;
  dfree(&c->d_acc);
  c->cap_acc = 0;
  int rc = dalloc(c, &c->d_acc, 6 * np);
  if (rc) return rc;
  c->cap_acc = np;
  return MRF_OK;
}

int ensure_sell0(mrf_ctx* c, size_t n) {
  if (n <= c->cap_sell0 && c->d_sell0) return MRF_OK;
  n += n / 4;
  dfree(&c->d_sell0);
  c->cap_sell0 = 0;
  int rc = dalloc(c, &c->d_sell0, n);
  if (rc) return rc;
  c->cap_sell0 = n;
  return MRF_OK;
}

void free_camera_lists(Camera& cam) {
  dfree(&cam.d_partial);  // holds d_fstatic and d_fdmask too
  cam.d_fstatic = nullptr;
  cam.d_fdmask = nullptr;
  dfree(&cam.d_dyn_desc);
  dfree(&cam.d_dcr);  // holds d_dfront too
  cam.d_dfront = nullptr;
  dfree(&cam.d_fspan);
  dfree(&cam.d_lnn);
  dfree(&cam.d_dyn);
  cam.ndyn = 0;
  dfree(&cam.d_front);
  dfree(&cam.d_lanes);
  dfree(&cam.d_len_f);
  dfree(&cam.d_len_w);
  dfree(&cam.d_inc);
  dfree(&cam.d_win_idx);
  dfree(&cam.d_pair_vox);
  dfree(&cam.d_pair_c);
  dfree(&cam.d_pair_sf);
  dfree(&cam.d_M);
  dfree(&cam.d_wseg_off);
  dfree(&cam.d_heads);
  dfree(&cam.d_tile_h0);
  dfree(&cam.d_chunk_off);
  dfree(&cam.d_chunk_desc);
  dfree(&cam.d_bstart);
  dfree(&cam.d_tail);
  cam.nchunks = 0;
  cam.nrays = 0;
  cam.nslices = 0;
  cam.nrec = cam.base_w = cam.nupd = cam.nwin = cam.nlist = 0;
  cam.npairs = cam.nwp = 0;
}


int bit_length(uint64_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

// Traversal lists, pairs and the window segment layout of one camera.
int traverse_camera(mrf_ctx* c, Camera& cam) {
  // the camera's messages follow its voxels through the re-traversal
  // (d_A, pool sized, is free until the next inference recomputes it)
  const size_t nvox_pool = c->slot_lin.size() * (size_t)kBrickVox;
  const bool keep_msgs = cam.d_M && cam.npairs && c->d_A && nvox_pool;
  if (keep_msgs) {
    launch_fill_f32(c->d_A, 0.0f, nvox_pool, c->stream);
    launch_scatter_pairs(cam.d_pair_vox, cam.d_M, cam.npairs, c->d_A, c->stream);
  }
  free_camera_lists(cam);
  const CamGeom cg = make_geom(c, cam);
  const size_t npix = (size_t)num_tiled_pixels(cg);
  const uint32_t ntiles = (uint32_t)(npix / kSlice);
  hipStream_t s = c->stream;
  int rc;
  if ((rc = ensure_pixel_scratch(c, npix)) || (rc = ensure_scan(c, npix))) return rc;
  // 1. every ray's records, in order, into one sliced staging region sized
  //    by a bound of each ray's count (full-ray activation: the bound of the
  //    activation walk; endpoint activation: the count): per tile the
  //    largest of its 64 rays (d_has), slice sizes (d_rayidx) -> offsets (d_offs)
  const uint32_t* per_ray = cam.d_walkn;
  if (!per_ray) {
    launch_count(cg, c->g, cam.d_depth, c->d_sig, c->d_bia, c->d_slot, c->d_counts, s);
    per_ray = c->d_counts;
  }
  launch_slice_len(per_ray, ntiles, c->d_has, c->d_rayidx, s);
  launch_exclusive_scan(c->d_rayidx, c->d_offs, ntiles, c->d_scan_tmp, c->d_total, s);
  const uint64_t nrec0 = read_total(c);
  HIPCHK(c, hipGetLastError());
  if (nrec0 >= 0x7FFFFFF0ull) return fail(c, MRF_E_LIMIT, "more than 2^31 incidence records in one camera");
  if ((rc = ensure_sell0(c, std::max<uint64_t>(nrec0, 1))) || (rc = dalloc(c, &cam.d_front, npix)) ||
      (rc = dalloc(c, &cam.d_lanes, npix)))
    return rc;
  uint32_t* overflow = reinterpret_cast<uint32_t*>(c->d_total + 6);
  launch_zero_words(overflow, 8, s);
  // 2. the emission splits each ray into front and window (counts into
  //    d_counts / d_rayidx); then the front region (sizes d_has -> offsets
  //    d_offs) and the window region (sizes d_counts -> offsets d_has) of the
  //    final record buffer
  launch_emit(cg, c->g, cam.d_depth, c->d_sig, c->d_bia, c->d_slot, c->d_offs, c->d_has, c->d_sell0, cam.d_front,
              cam.d_lanes, c->d_counts, c->d_rayidx, overflow, s);
  if ((rc = dalloc(c, &cam.d_len_f, ntiles)) || (rc = dalloc(c, &cam.d_len_w, ntiles)) ||
      (rc = dalloc(c, &cam.d_chunk_off, (size_t)ntiles + 1)))
    return rc;
  // ... and the front chunks, ceil(len_f / kChunk) per slice that has window
  // records (sizes d_counts -> chunk_off); the three totals in one readback
  uint64_t* tot3 = c->d_total + 2;
  launch_slice_len(c->d_counts, ntiles, cam.d_len_f, c->d_has, s);
  launch_exclusive_scan(c->d_has, c->d_offs, ntiles, c->d_scan_tmp, tot3, s);
  launch_slice_len(c->d_rayidx, ntiles, cam.d_len_w, c->d_counts, s);
  launch_exclusive_scan(c->d_counts, c->d_has, ntiles, c->d_scan_tmp, tot3 + 1, s);
  launch_chunk_counts(cam.d_len_f, cam.d_len_w, ntiles, c->d_counts, s);
  launch_exclusive_scan(c->d_counts, cam.d_chunk_off, (size_t)ntiles + 1, c->d_scan_tmp, tot3 + 2, s);
  uint64_t t3[3] = {0, 0, 0};
  HIPCHK(c, hipMemcpyAsync(t3, tot3, sizeof(t3), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  const uint64_t nrec_f = t3[0], nrec_w = t3[1];
  cam.nchunks = (uint32_t)t3[2];
  const uint64_t nrec = nrec_f + nrec_w;
  HIPCHK(c, hipGetLastError());
  if (nrec >= 0x7FFFFFF0ull) return fail(c, MRF_E_LIMIT, "more than 2^31 incidence records in one camera");
  cam.nrec = nrec;
  cam.base_w = nrec_f;
  cam.nslices = ntiles;
  if ((rc = dalloc(c, &cam.d_inc, nrec))) return rc;
  // 3. the records into their regions; the voxels of live records (any) and
  //    of live window records (wany), and the counts of both
  const uint64_t nvox = (uint64_t)c->slot_lin.size() * kBrickVox;
  if ((rc = ensure_vflags(c, std::max<uint64_t>(nvox, 1)))) return rc;
  uint8_t* any = reinterpret_cast<uint8_t*>(c->d_vflag);  // byte flags
  uint8_t* wany = any + nvox;
  uint32_t* pid = c->d_vflag + (2 * nvox + 3) / 4;  // the pair id of every touched voxel
  // per brick of the pool: counts of the two flags, then their exclusive scan
  const uint32_t nbr = (uint32_t)(nvox / kBrickVox);
  uint64_t* bcnt = reinterpret_cast<uint64_t*>(pid + nvox);
  uint64_t* boff = bcnt + nbr;
  // (any and wany are all zero here: cleared before the first camera, and
  // behind by the pair table of each camera)
  HIPCHK(c, hipMemsetAsync(c->d_total + 2, 0, 32, s));
  launch_reslice(cam.d_front, cam.d_lanes, (uint32_t)npix, c->d_sell0, c->d_offs, cam.d_len_f, c->d_has, cam.d_len_w,
                 (uint32_t)nrec_f, cam.d_inc, any, wany, (unsigned long long*)(c->d_total + 2), s);
  if (c->p.brick_local_vis == MRF_VIS_BRICK_LOCAL) {
    // brick entries of the window records, while the records hold voxel ids;
    // rows past the last slice read as zero words (the window kernel's
    // batches may look up to 16 rows past a slice's longest ray)
    const size_t wrows = (size_t)(nrec_w / kSlice);
    if ((rc = dalloc(c, &cam.d_bstart, wrows + 32)) || (rc = dalloc(c, &cam.d_tail, npix))) return rc;
    launch_zero_words(cam.d_bstart + wrows, 32 * sizeof(uint64_t), s);
    launch_brick_starts(cam.d_front, cam.d_lanes, cam.d_len_w, ntiles, cam.d_inc, (uint32_t)nrec_f, cam.d_bstart,
                        cam.d_tail, s);
  }
  // one allocation for the partial sums, the static sums and the masks of
  // changing records (per chunk and lane each)
  const size_t nparts = (size_t)cam.nchunks * kSlice;
  if ((rc = dalloc(c, &cam.d_chunk_desc, cam.nchunks)) || (rc = dalloc(c, &cam.d_partial, 3 * nparts)) ||
      (rc = dalloc(c, &cam.d_fspan, (size_t)ntiles * kSlice)))
    return rc;
  cam.d_fstatic = cam.d_partial + nparts;
  cam.d_fdmask = reinterpret_cast<uint32_t*>(cam.d_partial + 2 * nparts);
  launch_chunk_desc(cam.d_chunk_off, cam.d_len_f, ntiles, cam.d_chunk_desc, s);
  launch_front_span(cam.d_front, cam.d_lanes, ntiles * kSlice, cam.d_fspan, s);
  // 4. (camera, voxel) pairs: window pairs first, each group in voxel order.
  //    One readback: [0..3] the reslice counters (live records, live window
  //    records, records with gap records, rays with records), [4] the
  //    emission's overflow flag, [6] the pairs | the window pairs << 32
  launch_brick_flag_counts(any, wany, nbr, bcnt, s);
  if ((rc = u64_scan(c, bcnt, boff, nbr, c->d_total + 8))) return rc;
  uint64_t t8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIPCHK(c, hipMemcpyAsync(t8, c->d_total + 2, sizeof(t8), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipStreamSynchronize(s));
  if (t8[4]) return fail(c, MRF_E_LIMIT, "traversal: a ray has more records than its bound");
  const uint64_t np64 = t8[6] & 0xFFFFFFFFull;
  // the ray kernels address a camera's per-pair messages with 32-bit byte offsets
  if (np64 >= (1ull << 30)) return fail(c, MRF_E_LIMIT, "more than 2^30 (camera, voxel) pairs in one camera");
  const uint32_t np = (uint32_t)np64;
  cam.nupd = t8[0];
  const uint64_t nwin = t8[1];
  cam.nwin = nwin;
  cam.nlist = t8[2];
  cam.nrays = (uint32_t)t8[3];
  const uint32_t nwp = (uint32_t)(t8[6] >> 32);
  cam.npairs = np;
  cam.nwp = nwp;
  if ((rc = dalloc(c, &cam.d_pair_vox, np)) || (rc = dalloc(c, &cam.d_pair_c, np)) ||
      (rc = dalloc(c, &cam.d_pair_sf, np)) || (rc = dalloc(c, &cam.d_M, np)) ||
      (rc = dalloc(c, &cam.d_lnn, (size_t)std::max<uint32_t>(np, 1) * msg_bytes(msg_in_type(c->p.msg_dtype)))) ||
      (rc = ensure_acc(c, std::max<uint32_t>(np, 1))))
    return rc;
  launch_pair_table(any, wany, boff, nbr, nwp, pid, cam.d_pair_vox, s);
  // 5. per pair: the front parts of C; records point at their pairs from
  //    here on.  The same pass writes the window sort's input
  const int wbits = std::max(1, bit_length(nwp));
  if ((rc = ensure_sort(c, std::max<uint64_t>(nrec_w, 1), wbits))) return rc;
  // win_idx holds the sorted values of the whole window region; its first
  // nwin (the live records) are the segment layout
  if ((rc = dalloc(c, &cam.d_wseg_off, (size_t)nwp + 1)) ||
      (rc = dalloc(c, &cam.d_win_idx, std::max<uint64_t>(nrec_w, 1))))
    return rc;
  uint32_t* acc_w = c->d_acc;
  uint32_t* acc_f = c->d_acc + 2 * (size_t)np;
  uint32_t* acc_r = c->d_acc + 4 * (size_t)np;
  launch_pair_lengths(cam.d_inc, nrec, nrec_f, pid, acc_f, acc_r, np, c->d_keys_in, c->d_vals_in, nwp, s);
  // 6. window segments: the window records stably sorted by pair
  HIPCHK(c, sort_window((uint32_t)nrec_w, wbits, c->d_keys_in, c->d_keys_out, c->d_vals_in, cam.d_win_idx,
                        c->d_sort_tmp, c->sort_tmp_bytes, s));
  uint32_t* whead = c->d_x0;
  uint32_t* hw = c->d_x1;
  if (nwin) {
    launch_window_heads(c->d_keys_out, (uint32_t)nwin, whead, s);
    if ((rc = flag_scan(c, whead, hw, nwin))) return rc;
  }
  launch_window_offsets(c->d_keys_out, whead, (uint32_t)nwin, nwp, cam.d_wseg_off, s);
  // C with its window part summed per segment, and the front messages
  launch_pair_finish(cam.d_inc + nrec_f, c->d_keys_out, cam.d_win_idx, (uint32_t)nwin, nwp, acc_w, acc_f, acc_r, np,
                     cam.d_pair_c, cam.d_pair_sf, s);
  // 7. window segment head bitmask and per-tile head counts
  const size_t nwords_w = nwin / 64 + 1, ntiles_w = nwin / kTile + 1;
  if ((rc = dalloc(c, &cam.d_heads, nwords_w)) || (rc = dalloc(c, &cam.d_tile_h0, ntiles_w))) return rc;
  launch_pack_flags(whead, hw, (uint32_t)nwin, cam.d_heads, cam.d_tile_h0, s);
  if (keep_msgs)
    launch_gather_pairs(cam.d_pair_vox, c->d_A, np, cam.d_M, s);
  else
    launch_zero_words(cam.d_M, (size_t)np * sizeof(float), s);
  HIPCHK(c, hipStreamSynchronize(s));
  HIPCHK(c, hipGetLastError());
  dfree(&cam.d_walkn);  // after the sync: a free waits for the device
  return MRF_OK;
}

CamPass cam_pass(mrf_ctx* c, size_t ci);

int prepare(mrf_ctx* c) {
  if (!c->geom_dirty) return MRF_OK;
  ++c->epoch;  // camera lists and per-pass scratch are rebuilt below
  int rc;
  // 1. activation
  HIPCHK(c, hipMemsetAsync(c->d_bflags, 0, (size_t)c->nbricks, c->stream));
  for (Camera& cam : c->cams) {
    const CamGeom cg = make_geom(c, cam);
    if (c->p.activation == MRF_ACT_FULL_RAY) {
      const size_t npix = (size_t)num_tiled_pixels(cg);
      dfree(&cam.d_walkn);
      if ((rc = dalloc(c, &cam.d_walkn, npix))) return rc;
      launch_mark_rays(cg, c->g, cam.d_depth, c->d_sig, c->d_bia, c->d_bflags, cam.d_walkn, c->stream);
    } else
      launch_mark_endpoints(cg, c->g, cam.d_depth, c->d_bflags, c->stream);
  }
  std::vector<uint8_t> flags((size_t)c->nbricks);
  HIPCHK(c, hipMemcpyAsync(flags.data(), c->d_bflags, flags.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipGetLastError());
  // 2. slots for new bricks, ascending linear brick index
  for (int lin = 0; lin < c->nbricks; ++lin) {
    if (flags[lin] && c->h_slot[lin] < 0) {
      c->h_slot[lin] = (int32_t)c->slot_lin.size();
      c->slot_lin.push_back(lin);
    }
  }
  if ((rc = ensure_pool(c, c->slot_lin.size()))) return rc;
  // the directory is uploaded by every build, grown or not: the device copy
  // is not initialised when the context is created (4 B per brick of the grid)
  HIPCHK(c, hipMemcpyAsync(c->d_slot, c->h_slot.data(), c->h_slot.size() * 4, hipMemcpyHostToDevice, c->stream));
  // 3. traversal lists and pairs of every camera; the voxel flags of the
  //    pair numbering start cleared (each camera clears its own behind)
  {
    const size_t nvox = c->slot_lin.size() * (size_t)kBrickVox;
    if ((rc = ensure_vflags(c, std::max<size_t>(nvox, 1)))) return rc;
    launch_zero_words(c->d_vflag, (2 * nvox + 3) / 4 * 4, c->stream);  // the byte flags any, wany
  }
  for (Camera& cam : c->cams)
    if ((rc = traverse_camera(c, cam))) return rc;
  // 3b. voxels where some camera has window records; per camera, its
  //     front-only pairs there (their log2 n keeps changing after the first
  //     iterations of a call, the others' stops)
  {
    const size_t nvox = c->slot_lin.size() * (size_t)kBrickVox;
    size_t nflag = std::max<size_t>(nvox, 1);  // flags of the front-only pairs, then of the chunks
    for (const Camera& cam : c->cams) nflag = std::max<size_t>(nflag, cam.nchunks);
    if ((rc = ensure_vflags(c, nflag))) return rc;
    uint32_t* W = c->d_vflag;
    uint32_t* flags = W + nflag;   // per front-only pair: at a voxel of W
    uint32_t* pos = flags + nflag;
    uint32_t* cflags = pos + nflag;  // per front chunk: some log2 n of it keeps changing
    launch_zero_words(W, nvox * 4, c->stream);
    for (const Camera& cam : c->cams) launch_mark_window_voxels(cam.d_pair_vox, cam.nwp, W, c->stream);
    for (size_t ci = 0; ci < c->cams.size(); ++ci) {
      Camera& cam = c->cams[ci];
      const uint32_t nrest = cam.npairs - cam.nwp;
      // the front-only pairs at voxels of W, as a list (sized by its bound);
      // the front chunks whose log2 n keep changing: per chunk and lane the
      // mask of its changing records, per chunk the rows of its compact block
      // (crow) and a flag (cflags, scanned into pos); one readback of the
      // kept records [2], the list's length [8] and the changing chunks [9]
      dfree(&cam.d_dyn);
      if ((rc = dalloc(c, &cam.d_dyn, std::max<uint32_t>(nrest, 1)))) return rc;
      if ((rc = ensure_sort(c, std::max<size_t>(cam.nchunks, 1), 1))) return rc;
      uint32_t* crow = c->d_x0;  // sort scratch of at least nchunks words: rows per chunk
      launch_zero_words(c->d_total + 2, 8, c->stream);
      launch_zero_words(c->d_total + 8, 16, c->stream);
      if (nrest) {
        launch_dyn_flags(cam.d_pair_vox + cam.nwp, nrest, W, flags, c->stream);
        if ((rc = flag_scan(c, flags, pos, nrest, c->d_total + 8))) return rc;
        launch_dyn_pairs(flags, pos, nrest, cam.nwp, cam.d_pair_vox, cam.d_dyn, c->stream);
      }
      launch_chunk_static(cam_pass(c, ci), flags, cflags, cam.d_fdmask, crow, (unsigned long long*)(c->d_total + 2),
                          c->stream);
      if (cam.nchunks && (rc = flag_scan(c, cflags, pos, cam.nchunks, c->d_total + 9))) return rc;
      uint64_t t8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      HIPCHK(c, hipMemcpyAsync(t8, c->d_total + 2, sizeof(t8), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      cam.nstatic = t8[0];
      cam.ndyn = (uint32_t)t8[6];
      const uint64_t ndc = t8[7];
      dfree(&cam.d_dyn_desc);
      if ((rc = dalloc(c, &cam.d_dyn_desc, std::max<uint64_t>(ndc, 1)))) return rc;
      cam.ndyn_chunks = (uint32_t)ndc;
      launch_dyn_compact(cflags, pos, cam.nchunks, 0, flags, c->stream);  // chunk ids, over the used pair flags
      launch_gather_desc(cam.d_chunk_desc, flags, (uint32_t)ndc, cam.d_dyn_desc, c->stream);
      // the compact blocks of the changing chunks: rows per entry, their
      // exclusive scan (the first rows), the copies of the changing records
      uint64_t nrows = 0;
      if (ndc) {
        launch_dyn_rows(cam.d_dyn_desc, (uint32_t)ndc, crow, cflags, c->stream);
        if ((rc = flag_scan(c, cflags, pos, ndc))) return rc;
        nrows = read_total(c);
      }
      dfree(&cam.d_dcr);  // one allocation: dcr, then the blocks
      cam.d_dfront = nullptr;
      if ((rc = dalloc(c, &cam.d_dcr, ndc + nrows * kSlice))) return rc;
      cam.d_dfront = cam.d_dcr + ndc;
      launch_fill_dfront(cam_pass(c, ci), crow, pos, cam.d_dfront, cam.d_dcr, c->stream);
    }
    HIPCHK(c, hipGetLastError());
  }
  // 4. per-pass scratch sized for the largest camera
  size_t xneed = 1, pneed = 1, tneed = 1;
  for (const Camera& cam : c->cams) {
    xneed = std::max<size_t>(xneed, cam.nrec - cam.base_w);  // window region, padding included
    pneed = std::max<size_t>(pneed, cam.npairs);
    tneed = std::max<size_t>(tneed, (cam.nwin + kTile - 1) / kTile);
  }
  if (pneed > c->cap_pairs) {
    dfree(&c->d_Spair);
    c->cap_pairs = 0;
    if ((rc = dalloc(c, &c->d_Spair, pneed))) return rc;
    c->cap_pairs = pneed;
  }
  if (tneed > c->cap_tiles) {
    dfree(&c->d_pf_v); dfree(&c->d_pl_v); dfree(&c->d_mid);
    c->cap_tiles = 0;
    if ((rc = dalloc(c, &c->d_pf_v, tneed)) || (rc = dalloc(c, &c->d_pl_v, tneed)) || (rc = dalloc(c, &c->d_mid, tneed)))
      return rc;
    c->cap_tiles = tneed;
  }
  const size_t xout = xneed * msg_bytes(msg_out_type(c->p.msg_dtype));
  if (xout > c->Xo_bytes) {
    dev_free(c->d_Xo);
    c->d_Xo = nullptr;
    c->Xo_bytes = 0;
    if (dev_malloc(&c->d_Xo, xout) != hipSuccess) {
      (void)hipGetLastError();
      return fail(c, MRF_E_NOMEM, "segment array allocation failed");
    }
    c->Xo_bytes = xout;
  }
  c->geom_dirty = false;
  return MRF_OK;
}

int add_camera(mrf_ctx* c, const float pose[16], const float K[9], const float* depth, int rows, int cols,
               int layout, uint32_t* cam_id) {
  if (!pose || !depth || rows <= 0 || cols <= 0 || (layout != MRF_ROW_MAJOR && layout != MRF_COL_MAJOR))
    return fail(c, MRF_E_INVALID, "add_camera: invalid argument");
  if (K[0] <= 0 || K[4] <= 0) return fail(c, MRF_E_INVALID, "add_camera: focal lengths must be positive");
  int rc = ensure_tables(c, rows, cols);
  if (rc) return rc;
  Camera cam;
  if (layout == MRF_COL_MAJOR) {
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) cam.pose[i * 4 + j] = pose[j * 4 + i];
  } else {
    std::memcpy(cam.pose, pose, sizeof(cam.pose));
  }
  std::memcpy(cam.K, K, sizeof(cam.K));
  cam.rows = rows;
  cam.cols = cols;
  std::vector<float> rm;
  to_row_major(depth, rows, cols, layout, rm);
  if ((rc = dalloc(c, &cam.d_depth, rm.size()))) return rc;
  HIPCHK(c, hipMemcpy(cam.d_depth, rm.data(), rm.size() * 4, hipMemcpyHostToDevice));
  c->cams.push_back(cam);
  c->geom_dirty = true;
  if (cam_id) *cam_id = (uint32_t)(c->cams.size() - 1);
  return MRF_OK;
}

CamPass cam_pass(mrf_ctx* c, size_t ci) {
  Camera& cam = c->cams[ci];
  CamPass a;
  a.front = cam.d_front;
  a.lanes = cam.d_lanes;
  a.len_f = cam.d_len_f;
  a.len_w = cam.d_len_w;
  a.nslices = cam.nslices;
  a.inc = cam.d_inc;
  a.win_idx = cam.d_win_idx;
  a.base_w = (uint32_t)cam.base_w;
  a.chunk_off = cam.d_chunk_off;
  a.chunk_desc = cam.d_chunk_desc;
  a.fspan = cam.d_fspan;
  a.nchunks = cam.nchunks;
  a.partial = cam.d_partial;
  a.dyn_desc = cam.d_dyn_desc;
  a.ndyn_chunks = cam.ndyn_chunks;
  a.fdmask = cam.d_fdmask;
  a.fstatic = cam.d_fstatic;
  a.dfront = cam.d_dfront;
  a.dcr = cam.d_dcr;
  a.pair_vox = cam.d_pair_vox;
  a.pair_c = cam.d_pair_c;
  a.pair_sf = cam.d_pair_sf;
  a.npairs = cam.npairs;
  a.nwp = cam.nwp;
  a.nwin = (uint32_t)cam.nwin;
  a.wseg_off = cam.d_wseg_off;
  a.heads = cam.d_heads;
  a.tile_h0 = cam.d_tile_h0;
  a.apair = cam.d_lnn;
  a.dyn = cam.d_dyn;
  a.ndyn = cam.ndyn;
  a.msg_all = c->d_all_pairs + 1;
  a.S = c->d_Spair;
  a.pf_v = c->d_pf_v;
  a.pl_v = c->d_pl_v;
  a.mid = c->d_mid;
  a.B = c->d_B;
  a.M = cam.d_M;
  a.Xo = c->d_Xo;
  a.upd = c->d_upd;
  a.all_pairs = c->d_all_pairs;
  a.sig = c->d_sig;
  a.bia = c->d_bia;
  a.lookup_n = c->p.lookup_n;
  a.msg_type = c->p.msg_dtype;
  a.ray_sched = c->ray_sched < 0 ? kSchedF8W4 : c->ray_sched;
  a.brick_local = c->p.brick_local_vis == MRF_VIS_BRICK_LOCAL;
  a.bstart = cam.d_bstart;
  a.tail = cam.d_tail;
  return a;
}

// Pick the ray-kernel schedule (lanes per ray x loads in flight) by timing
// each candidate twice on the camera with the most incidences.  Only the
// pass's message scratch is touched, not beliefs.
int autotune(mrf_ctx* c) {
  if (c->cams.empty()) return MRF_OK;
  auto t0 = std::chrono::steady_clock::now();
  size_t best_cam = 0;
  for (size_t i = 1; i < c->cams.size(); ++i)
    if (c->cams[i].nupd > c->cams[best_cam].nupd) best_cam = i;
  hipEvent_t e0, e1;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  float best = 1e30f;
  int pick = kSchedF8W4;
  for (int sched = 0; sched < kNumSched; ++sched) {
    CamPass a = cam_pass(c, best_cam);
    a.ray_sched = sched;
    float tmin = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      launch_pair_msgs(a, c->stream);
      HIPCHK(c, hipEventRecord(e0, c->stream));
      launch_ray_messages(a, c->stream);
      HIPCHK(c, hipEventRecord(e1, c->stream));
      HIPCHK(c, hipEventSynchronize(e1));
      float ms = 0;
      HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
      tmin = std::min(tmin, ms);
    }
    if (tmin < best) {
      best = tmin;
      pick = sched;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(c, hipGetLastError());
  c->ray_sched = pick;
  c->autotune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return MRF_OK;
}

// Evaluate the current map from camera geometry cg (depth already on the
// device, or null), images copied back to the non-null outputs.
int evaluate(mrf_ctx* c, const CamGeom& cg, const float* h_depth, const float* d_depth, float* out_ll, float* out_ed,
             float* out_acc, float* out_sacc, mrf_eval* out) {
  int rc;
  if ((rc = prepare(c))) return rc;
  if (!c->d_B) return fail(c, MRF_E_STATE, "evaluate: the map is empty (no keyframe added)");
  const size_t n = (size_t)cg.rows * cg.cols;
  const size_t need = 5 * n;
  if (need > c->cap_eval) {
    dfree(&c->d_eval);
    c->cap_eval = 0;
    if ((rc = dalloc(c, &c->d_eval, need))) return rc;
    c->cap_eval = need;
  }
  const unsigned nb = eval_blocks(cg);
  if (nb > c->cap_eval_part) {
    dfree(&c->d_eval_part);
    c->cap_eval_part = 0;
    if ((rc = dalloc(c, &c->d_eval_part, nb))) return rc;
    c->cap_eval_part = nb;
  }
  if (!c->d_eval_counts && (rc = dalloc(c, &c->d_eval_counts, 8))) return rc;
  float* img[4] = {c->d_eval, c->d_eval + n, c->d_eval + 2 * n, c->d_eval + 3 * n};
  const float* dep = d_depth;
  if (h_depth) {
    float* up = c->d_eval + 4 * n;
    HIPCHK(c, hipMemcpyAsync(up, h_depth, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    dep = up;
  }
  launch_zero_words(c->d_eval_counts, 8 * sizeof(unsigned long long), c->stream);
  launch_eval(cg, c->g, dep, c->d_sig, c->d_bia, c->d_slot, c->d_B, c->d_upd, c->p.acc_thresh, img[0], img[1], img[2],
              img[3], c->d_eval_part, c->d_eval_counts, c->stream);
  HIPCHK(c, hipGetLastError());
  float* outs[4] = {out_ll, out_ed, out_acc, out_sacc};
  for (int k = 0; k < 4; ++k)
    if (outs[k]) HIPCHK(c, hipMemcpyAsync(outs[k], img[k], n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  std::vector<double> part(nb);
  unsigned long long cnt[8];
  HIPCHK(c, hipMemcpyAsync(part.data(), c->d_eval_part, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(cnt, c->d_eval_counts, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (out) {
    double ll = 0.0;
    for (double v : part) ll += v;
    out->log_likelihood = ll;
    for (int k = 0; k < 3; ++k) out->accuracy_counts[k] = (int64_t)cnt[k];
    for (int k = 0; k < 4; ++k) out->simple_counts[k] = (int64_t)cnt[3 + k];
    const double den = (double)n - (double)cnt[0];
    out->accuracy = den > 0 ? (double)cnt[2] / den : 0.0;
    out->simple_accuracy = n ? ((double)cnt[5] + (double)cnt[3]) / (double)n : 0.0;
  }
  return MRF_OK;
}

// One BP iteration on c->stream: every camera's pass (Gauss-Seidel), then
// the FP64 residual into d_total[1].  ev (kCamEvents per camera), when given,
// times each of a camera's kernels (pair messages, front sums, window,
// segment sums, update); a captured sequence records no events (the runtime
// cannot time events recorded by graph nodes).
constexpr int kCamEvents = 6;

hipError_t enqueue_iteration(mrf_ctx* c, const hipEvent_t* ev) {
  const size_t ncam = c->cams.size();
  const size_t nvox = c->slot_lin.size() * (size_t)kBrickVox;
  hipError_t e;
  for (size_t ci = 0; ci < ncam; ++ci) {
    const CamPass a = cam_pass(c, ci);
    const hipEvent_t* ec = ev ? ev + kCamEvents * ci : nullptr;
    if (ec && (e = hipEventRecord(ec[0], c->stream)) != hipSuccess) return e;
    launch_pair_msgs(a, c->stream);
    if (ec && (e = hipEventRecord(ec[1], c->stream)) != hipSuccess) return e;
    if (ec) {
      launch_ray_messages(a, c->stream, 1);
      if ((e = hipEventRecord(ec[2], c->stream)) != hipSuccess) return e;
      launch_ray_messages(a, c->stream, 2);
    } else {
      launch_ray_messages(a, c->stream);
    }
    if (ec && (e = hipEventRecord(ec[3], c->stream)) != hipSuccess) return e;
    if (ec) {
      launch_epilogue(a, c->stream, 1);
      if ((e = hipEventRecord(ec[4], c->stream)) != hipSuccess) return e;
      launch_epilogue(a, c->stream, 2);
    } else {
      launch_epilogue(a, c->stream);
    }
    if (ec && (e = hipEventRecord(ec[5], c->stream)) != hipSuccess) return e;
  }
  launch_zero_words(c->d_total + 1, 8, c->stream);  // a kernel, not a memset: the sequence may be captured
  launch_residual(c->d_B, c->d_A, nvox, (unsigned long long*)(c->d_total + 1), c->stream);
  return hipGetLastError();
}

void drop_graph(mrf_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
}

// Capture the iteration into a graph (once per buffer epoch and schedule).
// Any failure switches the context to eager launches for good.
void ensure_graph(mrf_ctx* c) {
  const uint64_t key = (c->epoch << 8) ^ (uint64_t)(c->ray_sched + 1) ^ ((uint64_t)c->cams.size() << 40);
  if (!c->graph_ok || (c->gexec && c->gexec_key == key)) return;
  drop_graph(c);
  hipGraph_t g = nullptr;
  bool ok = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    const hipError_t e1 = enqueue_iteration(c, nullptr);
    const hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    ok = e1 == hipSuccess && e2 == hipSuccess && g;
  }
  if (ok) ok = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0) == hipSuccess;
  if (g) (void)hipGraphDestroy(g);
  if (!ok) {
    (void)hipGetLastError();
    c->gexec = nullptr;
    c->graph_ok = false;
    drop_graph(c);
    return;
  }
  c->gexec_key = key;
}

// ---- learned warm start
int ensure_mlp(mrf_ctx* c) { return c->d_mlp ? MRF_OK : dalloc(c, &c->d_mlp, (size_t)kMlpParams); }

// Warm-start features of every touched voxel of the current state (after an
// iteration) into d_wX / d_wvox, in pool order; *n_out rows.
int features_device(mrf_ctx* c, uint32_t* n_out) {
  const size_t nslots = c->slot_lin.size(), nvox = nslots * (size_t)kBrickVox;
  if (nvox > 0xFFFFFFFFull) return fail(c, MRF_E_LIMIT, "features: more than 2^32 voxels in the pool");
  int rc;
  if (nvox > c->cap_wvox) {
    dfree(&c->d_wpass); dfree(&c->d_wmsum); dfree(&c->d_wmmax); dfree(&c->d_wdelta);
    dfree(&c->d_wnc); dfree(&c->d_whits); dfree(&c->d_wflag); dfree(&c->d_wrow);
    c->cap_wvox = 0;
    if ((rc = dalloc(c, &c->d_wpass, nvox)) || (rc = dalloc(c, &c->d_wmsum, nvox)) ||
        (rc = dalloc(c, &c->d_wmmax, nvox)) || (rc = dalloc(c, &c->d_wdelta, nvox)) ||
        (rc = dalloc(c, &c->d_wnc, nvox)) || (rc = dalloc(c, &c->d_whits, nvox)) ||
        (rc = dalloc(c, &c->d_wflag, nvox)) || (rc = dalloc(c, &c->d_wrow, nvox)))
      return rc;
    c->cap_wvox = nvox;
  }
  if (nslots > c->cap_slot_lin) {
    dfree(&c->d_slot_lin);
    c->cap_slot_lin = 0;
    if ((rc = dalloc(c, &c->d_slot_lin, nslots))) return rc;
    c->cap_slot_lin = nslots;
  }
  if ((rc = ensure_scan(c, nvox))) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_slot_lin, c->slot_lin.data(), nslots * sizeof(int32_t), hipMemcpyHostToDevice,
                           c->stream));
  launch_zero_words(c->d_wpass, nvox * 4, c->stream);
  launch_zero_words(c->d_wmsum, nvox * 4, c->stream);
  launch_zero_words(c->d_wnc, nvox * 4, c->stream);
  launch_zero_words(c->d_whits, nvox * 4, c->stream);
  for (const Camera& cam : c->cams) {
    launch_feat_hits(make_geom(c, cam), c->g, cam.d_depth, c->d_slot, c->d_whits, c->stream);
    launch_feat_pairs(cam.d_pair_vox, cam.d_pair_c, cam.d_M, cam.npairs, c->d_wpass, c->d_wmsum, c->d_wmmax,
                      c->d_wnc, c->stream);
  }
  launch_feat_flags(c->d_wnc, nvox, c->d_wflag, c->stream);
  if ((rc = flag_scan(c, c->d_wflag, c->d_wrow, nvox))) return rc;
  const uint64_t n = read_total(c);
  if (n > c->cap_wrows) {
    dfree(&c->d_wX); dfree(&c->d_wvox); dfree(&c->d_wpred);
    c->cap_wrows = 0;
    if ((rc = dalloc(c, &c->d_wX, (size_t)n * kMlpIn)) || (rc = dalloc(c, &c->d_wvox, (size_t)n)) ||
        (rc = dalloc(c, &c->d_wpred, (size_t)n)))
      return rc;
    c->cap_wrows = n;
  }
  launch_features(c->g, c->d_slot, c->d_slot_lin, (uint32_t)nslots, c->d_whits, c->d_wpass, c->d_wmsum, c->d_wmmax,
                  c->d_wnc, c->d_B, c->prior_logodds, c->d_wrow, c->d_wX, c->d_wvox, c->stream);
  HIPCHK(c, hipGetLastError());
  c->wrows = (uint32_t)n;
  *n_out = (uint32_t)n;
  return MRF_OK;
}

// Copies of the beliefs and every camera's messages (to = true: take the
// snapshot, false: restore it).
int snapshot(mrf_ctx* c, bool take) {
  const size_t nvox = c->slot_lin.size() * (size_t)kBrickVox;
  size_t npairs = 0;
  for (const Camera& cam : c->cams) npairs += cam.npairs;
  int rc;
  if (take && nvox > c->cap_snapB) {
    dfree(&c->d_snapB);
    c->cap_snapB = 0;
    if ((rc = dalloc(c, &c->d_snapB, nvox))) return rc;
    c->cap_snapB = nvox;
  }
  if (take && npairs > c->cap_snapM) {
    dfree(&c->d_snapM);
    c->cap_snapM = 0;
    if ((rc = dalloc(c, &c->d_snapM, npairs ? npairs : 1))) return rc;
    c->cap_snapM = npairs;
  }
  const hipMemcpyKind k = hipMemcpyDeviceToDevice;
  if (take)
    HIPCHK(c, hipMemcpyAsync(c->d_snapB, c->d_B, nvox * 4, k, c->stream));
  else
    HIPCHK(c, hipMemcpyAsync(c->d_B, c->d_snapB, nvox * 4, k, c->stream));
  size_t off = 0;
  for (Camera& cam : c->cams) {
    if (cam.npairs) {
      if (take)
        HIPCHK(c, hipMemcpyAsync(c->d_snapM + off, cam.d_M, (size_t)cam.npairs * 4, k, c->stream));
      else
        HIPCHK(c, hipMemcpyAsync(cam.d_M, c->d_snapM + off, (size_t)cam.npairs * 4, k, c->stream));
    }
    off += cam.npairs;
  }
  return MRF_OK;
}

// Training/check buffers for n samples.
int ensure_train(mrf_ctx* c, size_t n, uint32_t batch) {
  int rc;
  if (n > c->cap_trows) {
    dfree(&c->d_tX); dfree(&c->d_tT); dfree(&c->d_tSel);
    c->cap_trows = 0;
    if ((rc = dalloc(c, &c->d_tX, n * kMlpIn)) || (rc = dalloc(c, &c->d_tT, n)) || (rc = dalloc(c, &c->d_tSel, n)))
      return rc;
    c->cap_trows = n;
  }
  if (!c->d_tG) {
    if ((rc = dalloc(c, &c->d_tG, (size_t)kMlpParams)) || (rc = dalloc(c, &c->d_tM, (size_t)kMlpParams)) ||
        (rc = dalloc(c, &c->d_tV, (size_t)kMlpParams)) || (rc = dalloc(c, &c->d_tLoss, (size_t)1)))
      return rc;
  }
  const size_t nw = mlp_grad_waves(batch);
  if (nw > c->cap_tpart) {
    dfree(&c->d_tPart); dfree(&c->d_tLpart);
    c->cap_tpart = 0;
    if ((rc = dalloc(c, &c->d_tPart, nw * kMlpParams)) || (rc = dalloc(c, &c->d_tLpart, nw))) return rc;
    c->cap_tpart = nw;
  }
  return ensure_mlp(c);
}

}  // namespace

extern "C" {

void mrf_params_default(mrf_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  const float hfov = 0.01f;
  p->res = 0.01f;
  p->occ_thresh = 0.1f;
  p->acc_thresh = 0.5f;
  p->prior = 0.01f;
  p->max_iters = 1;
  p->dims[0] = p->dims[1] = p->dims[2] = 1.0f;
  p->rows = 1;
  p->cols = 1;
  p->K[0] = p->K[4] = (p->cols / 2.0f) / tanf(hfov / 2.0f);
  p->K[2] = p->cols * 0.5f;
  p->K[5] = p->rows * 0.5f;
  p->K[8] = 1.0f;
  p->lookup_n = 3;
  p->const_sigmasq_poly[0] = 1.5f * p->res;
  p->activation = MRF_ACT_ENDPOINT;
  p->msg_dtype = MRF_MSG_F32;
  p->mask_848_border = 1;
  p->brick_local_vis = MRF_VIS_BRICK_LOCAL;  // the reference's pre pass
}

int mrf_create(const mrf_params* p, mrf_ctx** out) {
  if (!p || !out) return MRF_E_INVALID;
  *out = nullptr;
  if (!(p->res > 0) || !(p->prior > 0 && p->prior < 1) || (p->lookup_n != 2 && p->lookup_n != 3) ||
      p->activation < 0 || p->activation > 1 || p->msg_dtype < 0 || p->msg_dtype > 4 ||
      (p->brick_local_vis != MRF_VIS_CARRY && p->brick_local_vis != MRF_VIS_BRICK_LOCAL))
    return MRF_E_INVALID;
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || cnt == 0) {
    (void)hipGetLastError();
    return MRF_E_HIP;
  }
  mrf_ctx* c = new mrf_ctx();
  c->p = *p;
  c->p.depth_sigma_lookup = nullptr;
  c->p.depth_bias_lookup = nullptr;
  int mv[3];
  for (int a = 0; a < 3; ++a) {
    mv[a] = (int)std::lround((double)p->dims[a] / (double)p->res);
    if (mv[a] <= 0 || mv[a] > 2048) {
      delete c;
      return MRF_E_LIMIT;
    }
  }
  GridGeom& g = c->g;
  g.mx = mv[0];
  g.my = mv[1];
  g.mz = mv[2];
  g.nx = (mv[0] + 7) & ~7;
  g.ny = (mv[1] + 7) & ~7;
  g.nz = (mv[2] + 7) & ~7;
  g.bx = g.nx >> kBrickLog;
  g.by = g.ny >> kBrickLog;
  g.bz = g.nz >> kBrickLog;
  g.res = (double)p->res;
  g.off[0] = 0.5 * (double)mv[0];
  g.off[1] = 0.5 * (double)mv[1];
  g.off[2] = 0.0;
  g.lookup_n = p->lookup_n;
  c->nbricks = g.bx * g.by * g.bz;
  c->prior_logodds = logf(p->prior) - logf(1.0f - p->prior);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return MRF_E_HIP;
  }
  {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || scratch_address(dev, &c->scratch) != hipSuccess) {
      (void)hipGetLastError();
      mrf_destroy(c);
      return MRF_E_HIP;
    }
  }
  c->h_slot.assign((size_t)c->nbricks, -1);
  if (const char* env = getenv("MRF_NO_GRAPH"))
    if (env[0] && env[0] != '0') c->launch_mode = MRF_LAUNCH_TIMED;
  if (const char* env = getenv("MRF_RAY_SCHED")) {
    const int v = atoi(env);
    if (v >= 0 && v < kNumSched) {
      c->ray_sched = v;
      c->sched_forced = true;
    }
  }
  int rc;
  if ((rc = dalloc(c, &c->d_slot, (size_t)c->nbricks)) || (rc = dalloc(c, &c->d_bflags, (size_t)c->nbricks)) ||
      (rc = dalloc(c, &c->d_total, 16)) || (rc = dalloc(c, &c->d_all_pairs, 2))) {
    mrf_destroy(c);
    return rc;
  }
  {
    const uint32_t flags[2] = {1u, 2u};  // defined before the autotuner's first pair pass
    if (hipMemcpy(c->d_all_pairs, flags, sizeof(flags), hipMemcpyHostToDevice) != hipSuccess) {
      mrf_destroy(c);
      return MRF_E_HIP;
    }
  }
  if (p->depth_sigma_lookup && p->depth_bias_lookup) {
    const size_t cells = (size_t)((p->rows + kLookupPx - 1) / kLookupPx) * ((p->cols + kLookupPx - 1) / kLookupPx);
    c->sig_tab.assign(p->depth_sigma_lookup, p->depth_sigma_lookup + cells * p->lookup_n);
    c->bia_tab.assign(p->depth_bias_lookup, p->depth_bias_lookup + cells * p->lookup_n);
    c->user_tables = true;
    c->tab_rows = p->rows;
    c->tab_cols = p->cols;
    if ((rc = dalloc(c, &c->d_sig, c->sig_tab.size())) || (rc = dalloc(c, &c->d_bia, c->bia_tab.size()))) {
      mrf_destroy(c);
      return rc;
    }
    if (hipMemcpy(c->d_sig, c->sig_tab.data(), c->sig_tab.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_bia, c->bia_tab.data(), c->bia_tab.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
      mrf_destroy(c);
      return MRF_E_HIP;
    }
  }
  *out = c;
  return MRF_OK;
}

void mrf_destroy(mrf_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (Camera& cam : c->cams) {
    free_camera_lists(cam);
    dfree(&cam.d_depth);
    dfree(&cam.d_walkn);
  }
  dfree(&c->d_sig); dfree(&c->d_bia); dfree(&c->d_slot); dfree(&c->d_bflags);
  dfree(&c->d_B); dfree(&c->d_A); dfree(&c->d_upd); dfree(&c->d_out);
  dev_free(c->d_Xo);
  dev_free(c->d_sort_tmp);
  dfree(&c->d_Spair);
  dfree(&c->d_pf_v); dfree(&c->d_pl_v); dfree(&c->d_mid);
  dfree(&c->d_keys_in); dfree(&c->d_keys_out); dfree(&c->d_vals_in);
  dfree(&c->d_x0); dfree(&c->d_x1);
  dfree(&c->d_vflag); dfree(&c->d_acc); dfree(&c->d_sell0);
  dfree(&c->d_eval); dfree(&c->d_eval_part); dfree(&c->d_eval_counts);
  dfree(&c->d_wpass); dfree(&c->d_wmsum); dfree(&c->d_wmmax); dfree(&c->d_wdelta);
  dfree(&c->d_wnc); dfree(&c->d_whits); dfree(&c->d_wflag); dfree(&c->d_wrow); dfree(&c->d_slot_lin);
  dfree(&c->d_wX); dfree(&c->d_wvox); dfree(&c->d_wpred); dfree(&c->d_mlp);
  dfree(&c->d_snapB); dfree(&c->d_snapM); dfree(&c->d_ref); dfree(&c->d_rerr);
  dfree(&c->d_tX); dfree(&c->d_tT); dfree(&c->d_tG); dfree(&c->d_tM); dfree(&c->d_tV); dfree(&c->d_tPart);
  dfree(&c->d_tSel); dfree(&c->d_tLpart); dfree(&c->d_tLoss);
  drop_graph(c);
  dfree(&c->d_counts); dfree(&c->d_has); dfree(&c->d_offs); dfree(&c->d_rayidx);
  dfree(&c->d_scan_tmp); dfree(&c->d_total);
  dev_free(c->d_fscan_tmp); dfree(&c->d_all_pairs);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int mrf_add_camera_with_depth(mrf_ctx* c, const float pose[16], const float* depth, int rows, int cols, int layout,
                              uint32_t* cam_id) {
  if (!c) return MRF_E_INVALID;
  if (rows != c->p.rows || cols != c->p.cols)
    return fail(c, MRF_E_INVALID, "depth image must be params.rows x params.cols");
  return add_camera(c, pose, c->p.K, depth, rows, cols, layout, cam_id);
}

int mrf_add_camera_with_depth_k(mrf_ctx* c, const float pose[16], const float K[9], const float* depth, int rows,
                                int cols, int layout, uint32_t* cam_id) {
  if (!c || !K) return MRF_E_INVALID;
  return add_camera(c, pose, K, depth, rows, cols, layout, cam_id);
}

int mrf_set_pose(mrf_ctx* c, uint32_t id, const float pose[16], int layout) {
  if (!c || !pose) return MRF_E_INVALID;
  if (id >= c->cams.size()) return fail(c, MRF_E_RANGE, "set_pose: camera id out of range");
  // remove the camera's rays: voxels whose summed chord length over the other
  // cameras' traversals is below 0.05 return to the prior
  // (GVDBInference.cpp:66-82, reset_empty_voxels, gvdbkernels.cu:498-517);
  // d_A serves as the per-voxel sum (every inference recomputes it first)
  const size_t nv = c->slot_lin.size() * (size_t)kBrickVox;
  if (nv && c->d_A) {
    launch_fill_f32(c->d_A, 0.0f, nv, c->stream);
    for (size_t ci = 0; ci < c->cams.size(); ++ci)
      if (ci != id) launch_cum_add(c->cams[ci].d_pair_vox, c->cams[ci].d_pair_c, c->cams[ci].npairs, c->d_A, c->stream);
    launch_reset_empty(c->d_A, nv, c->prior_logodds, c->d_B, c->d_upd, c->stream);
    const Camera& mc = c->cams[id];
    if (mc.d_M) launch_reset_pair_msgs(mc.d_pair_vox, mc.npairs, c->d_A, mc.d_M, c->stream);
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
  }
  Camera& cam = c->cams[id];
  if (layout == MRF_COL_MAJOR) {
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) cam.pose[i * 4 + j] = pose[j * 4 + i];
  } else {
    std::memcpy(cam.pose, pose, sizeof(cam.pose));
  }
  c->geom_dirty = true;
  return MRF_OK;
}

int mrf_num_cameras(mrf_ctx* c, uint32_t* n) {
  if (!c || !n) return MRF_E_INVALID;
  *n = (uint32_t)c->cams.size();
  return MRF_OK;
}

int mrf_prepare(mrf_ctx* c) {
  if (!c) return MRF_E_INVALID;
  return prepare(c);
}

// One perform_inference call; with warm != nullptr the learned warm start
// is applied after warm_after iterations and undone when a later residual
// exceeds growth x the previous one (growth > 0) and 0.05.
static int run_inference(mrf_ctx* c, const uint32_t* ids, int n_ids, int max_iters, double tol, const float* warm,
                         int warm_after, double growth, mrf_stats* st) {
  if (!c) return MRF_E_INVALID;
  auto t0 = std::chrono::steady_clock::now();
  mrf_stats s;
  std::memset(&s, 0, sizeof(s));
  std::vector<uint32_t> sel;
  if (ids) {
    for (int i = 0; i < n_ids; ++i) {
      if (ids[i] >= c->cams.size()) return fail(c, MRF_E_RANGE, "perform_inference: camera id out of range");
      sel.push_back(ids[i]);
    }
  } else {
    for (uint32_t i = 0; i < c->cams.size(); ++i) sel.push_back(i);
  }
  if (max_iters <= 0) max_iters = c->p.max_iters;
  const bool dirty = c->geom_dirty;
  int rc = prepare(c);
  if (rc) return rc;
  auto t1 = std::chrono::steady_clock::now();
  if (dirty) s.t_geometry_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (c->ray_sched < 0 && (rc = autotune(c))) return rc;
  const size_t nvox = c->slot_lin.size() * (size_t)kBrickVox;
  // reset M_c of the selected cameras (GVDBInference.cpp:128-134); the
  // first iteration refreshes every pair's message
  for (uint32_t id : sel) launch_zero_words(c->cams[id].d_M, (size_t)c->cams[id].npairs * sizeof(float), c->stream);
  launch_fill_u32(c->d_all_pairs, 1u, 1, c->stream);      // every pair's update
  launch_fill_u32(c->d_all_pairs + 1, 2u, 1, c->stream);  // every pair's log2 n, two iterations
  launch_marginals(c->d_B, c->d_A, nvox, c->stream);
  for (const Camera& cam : c->cams) {
    s.updates_per_iter += (int64_t)cam.nupd;
    s.window_updates_per_iter += (int64_t)cam.nwin;
    s.static_front_updates += (int64_t)cam.nstatic;
    s.front_slots += (int64_t)cam.base_w;
    s.window_slots += (int64_t)(cam.nrec - cam.base_w);
    s.pairs_per_iter += cam.npairs;
    s.rays += cam.nrays;
    s.list_records += (int64_t)cam.nlist;
  }
  const size_t ncam = c->cams.size();
  // the iteration is replayed from a captured graph unless per-kernel timing is asked for
  if (c->launch_mode == MRF_LAUNCH_GRAPH) ensure_graph(c);
  const bool graph = c->launch_mode == MRF_LAUNCH_GRAPH && c->gexec != nullptr;
  // events: [0, 1] around the iteration, then kCamEvents per camera unless graph-launched
  std::vector<hipEvent_t> ev(graph ? 2 : 2 + kCamEvents * ncam, nullptr);
  struct Events {
    std::vector<hipEvent_t>& v;
    ~Events() {
      for (auto& e : v)
        if (e) (void)hipEventDestroy(e);
    }
  } ev_guard{ev};
  for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
  double tray = 0, tepi = 0, titer = 0;
  double tk[kCamEvents - 1] = {0, 0, 0, 0, 0};  // pair messages, front sums, window, segment sums, update
  int it = 0;
  double resid = 0;
  bool conv = false;
  if (warm) {
    if ((rc = ensure_mlp(c))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_mlp, warm, kMlpParams * sizeof(float), hipMemcpyHostToDevice, c->stream));
  }
  int warm_state = 0;  // 0: none, 1: applied, 2: applied and undone
  double r_plain = 0, r_prev = 0;
  if (c->ref_n && c->ref_n != nvox)
    return fail(c, MRF_E_STATE, "perform_inference: the map changed since mrf_set_reference");
  double rerr = -1.0;
  c->trace_resid.clear();
  c->trace_err.clear();
  for (; it < max_iters; ++it) {
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    if (graph)
      HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    else
      HIPCHK(c, enqueue_iteration(c, ev.data() + 2));
    HIPCHK(c, hipEventRecord(ev[1], c->stream));
    launch_fill_u32(c->d_all_pairs, 0u, 1, c->stream);  // front-only messages are in M now
    launch_countdown(c->d_all_pairs + 1, c->stream);
    uint64_t bits = 0, ebits = 0;
    if (c->ref_n) {
      launch_ref_error(c->d_A, c->d_upd, c->d_ref, nvox, c->d_rerr, c->stream);
      HIPCHK(c, hipMemcpyAsync(&ebits, c->d_rerr, 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(&bits, c->d_total + 1, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipGetLastError());
    std::memcpy(&resid, &bits, 8);
    if (c->ref_n) std::memcpy(&rerr, &ebits, 8);
    c->trace_resid.push_back(resid);
    c->trace_err.push_back(rerr);
    float ms = 0;
    if (!graph) {
      for (size_t ci = 0; ci < ncam; ++ci) {
        const hipEvent_t* ec = ev.data() + 2 + kCamEvents * ci;
        for (int k = 0; k + 1 < kCamEvents; ++k) {
          HIPCHK(c, hipEventElapsedTime(&ms, ec[k], ec[k + 1]));
          tk[k] += ms;
        }
        HIPCHK(c, hipEventElapsedTime(&ms, ec[1], ec[3]));
        tray += ms;
        HIPCHK(c, hipEventElapsedTime(&ms, ec[3], ec[5]));
        tepi += ms;
      }
    }
    HIPCHK(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
    titer += ms;
    bool check = true;
    if (warm && it == warm_after - 1) {
      // features of the first-iteration state, MLP, warm-start update
      const auto w0 = std::chrono::steady_clock::now();
      uint32_t n = 0;
      if ((rc = features_device(c, &n))) return rc;
      launch_mlp_fwd(c->d_mlp, c->d_wX, n, c->d_wpred, c->stream);
      if ((rc = snapshot(c, true))) return rc;
      launch_warm_apply(c->d_wvox, c->d_wpred, n, c->d_wnc, c->d_B, c->d_wdelta, c->stream);
      for (const Camera& cam : c->cams)
        launch_warm_pairs(cam.d_pair_vox, cam.d_pair_c, cam.npairs, c->d_wdelta, cam.d_M, c->stream);
      launch_marginals(c->d_B, c->d_A, nvox, c->stream);
      launch_fill_u32(c->d_all_pairs, 1u, 1, c->stream);  // every message was moved
      launch_fill_u32(c->d_all_pairs + 1, 2u, 1, c->stream);
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipGetLastError());
      s.t_warm_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
      s.warm_voxels = n;
      warm_state = 1;
      r_plain = resid;
      check = false;
    } else if (warm_state == 1 && it > warm_after && growth > 0 && resid > growth * r_prev && resid > 0.05) {
      // the residual grows: back to the plain state before the warm start
      // (the restored front-only messages are their constants already; the
      // stored log2 n of the warm iterations are recomputed once)
      if ((rc = snapshot(c, false))) return rc;
      launch_marginals(c->d_B, c->d_A, nvox, c->stream);
      launch_fill_u32(c->d_all_pairs + 1, 1u, 1, c->stream);
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipGetLastError());
      warm_state = 2;
      resid = r_plain;
      check = false;
    }
    r_prev = resid;
    if (check && ((tol > 0 && resid < tol) || (c->ref_n && c->ref_stop > 0 && rerr <= c->ref_stop))) {
      conv = true;
      ++it;
      break;
    }
  }
  s.iters = it;
  s.graph = graph ? 1 : 0;
  s.converged = conv ? 1 : 0;
  s.residual = resid;
  s.t_ray_kernel_ms = graph ? -1.0 : tray;
  s.t_epilogue_ms = graph ? -1.0 : tepi;
  s.t_pair_msg_ms = graph ? -1.0 : tk[0];
  s.t_front_ms = graph ? -1.0 : tk[1];
  s.t_window_ms = graph ? -1.0 : tk[2];
  s.t_seg_sum_ms = graph ? -1.0 : tk[3];
  s.t_seg_update_ms = graph ? -1.0 : tk[4];
  s.t_iters_ms = titer;
  s.t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  s.ray_sched = c->ray_sched;
  s.t_autotune_ms = c->autotune_ms;
  s.warm = warm_state;
  s.ref_error = rerr;
  if (st) *st = s;
  return MRF_OK;
}

int mrf_perform_inference(mrf_ctx* c, const uint32_t* ids, int n_ids, int max_iters, double tol, mrf_stats* st) {
  return run_inference(c, ids, n_ids, max_iters, tol, nullptr, 0, 0.0, st);
}

int mrf_perform_inference_warm(mrf_ctx* c, const float* params, int warm_after, int max_iters, double tol,
                               double growth, mrf_stats* st) {
  if (!c) return MRF_E_INVALID;
  if (!params) return fail(c, MRF_E_INVALID, "perform_inference_warm: no MLP parameters");
  if (warm_after < 1) return fail(c, MRF_E_INVALID, "perform_inference_warm: warm_after must be >= 1");
  return run_inference(c, nullptr, 0, max_iters, tol, params, warm_after, growth, st);
}

int mrf_features(mrf_ctx* c, float* X, uint32_t* vox, int64_t cap, int64_t* n) {
  if (!c || !n) return MRF_E_INVALID;
  if (c->geom_dirty) return fail(c, MRF_E_STATE, "features: run perform_inference first");
  uint32_t rows = 0;
  int rc = features_device(c, &rows);
  if (rc) return rc;
  *n = rows;
  if (!X && !vox) return MRF_OK;
  if (cap < (int64_t)rows) return fail(c, MRF_E_SHAPE, "features: buffers hold fewer rows than there are voxels");
  if (X) HIPCHK(c, hipMemcpyAsync(X, c->d_wX, (size_t)rows * kMlpIn * 4, hipMemcpyDeviceToHost, c->stream));
  if (vox) HIPCHK(c, hipMemcpyAsync(vox, c->d_wvox, (size_t)rows * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRF_OK;
}

int mrf_warm_predictions(mrf_ctx* c, uint32_t* vox, float* pred, int64_t cap, int64_t* n) {
  if (!c || !n) return MRF_E_INVALID;
  *n = c->wrows;
  if (!vox && !pred) return MRF_OK;
  if (cap < (int64_t)c->wrows) return fail(c, MRF_E_SHAPE, "warm_predictions: buffers too small");
  if (c->wrows) {
    if (vox) HIPCHK(c, hipMemcpyAsync(vox, c->d_wvox, (size_t)c->wrows * 4, hipMemcpyDeviceToHost, c->stream));
    if (pred) HIPCHK(c, hipMemcpyAsync(pred, c->d_wpred, (size_t)c->wrows * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRF_OK;
}

int mrf_get_trace(mrf_ctx* c, double* resid, double* err, int cap, int* n) {
  if (!c || !n) return MRF_E_INVALID;
  *n = (int)c->trace_resid.size();
  const int m = std::min(cap, *n);
  for (int i = 0; i < m; ++i) {
    if (resid) resid[i] = c->trace_resid[i];
    if (err) err[i] = c->trace_err[i];
  }
  return MRF_OK;
}

int mrf_set_reference(mrf_ctx* c, const float* alphas, int64_t rows, int64_t cols, double stop_at) {
  if (!c) return MRF_E_INVALID;
  if (rows == 0) {
    c->ref_n = 0;
    c->ref_stop = 0.0;
    return MRF_OK;
  }
  if (!alphas || cols != kBrickVox) return fail(c, MRF_E_SHAPE, "set_reference: reference must be bricks x 512");
  if (c->geom_dirty) return fail(c, MRF_E_STATE, "set_reference: call mrf_prepare first");
  const size_t nvox = c->slot_lin.size() * (size_t)kBrickVox;
  if ((size_t)rows * kBrickVox != nvox) return fail(c, MRF_E_SHAPE, "set_reference: rows != num_bricks");
  int rc;
  if (!c->d_rerr && (rc = dalloc(c, &c->d_rerr, (size_t)1))) return rc;
  if (nvox > c->ref_n || !c->d_ref) {
    dfree(&c->d_ref);
    c->ref_n = 0;
    if ((rc = dalloc(c, &c->d_ref, nvox))) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(c->d_ref, alphas, nvox * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->ref_n = nvox;
  c->ref_stop = stop_at;
  return MRF_OK;
}

int mrf_mlp_forward(mrf_ctx* c, const float* params, const float* X, int64_t n, float* y) {
  if (!c || !params || (n > 0 && (!X || !y)) || n < 0) return MRF_E_INVALID;
  if (n > 0xFFFFFFFFll) return fail(c, MRF_E_LIMIT, "mlp_forward: more than 2^32 samples");
  int rc = ensure_train(c, (size_t)(n ? n : 1), 1);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_mlp, params, kMlpParams * 4, hipMemcpyHostToDevice, c->stream));
  if (n == 0) return MRF_OK;
  HIPCHK(c, hipMemcpyAsync(c->d_tX, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, c->stream));
  launch_mlp_fwd(c->d_mlp, c->d_tX, (uint32_t)n, c->d_tT, c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(y, c->d_tT, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRF_OK;
}

int mrf_mlp_grad(mrf_ctx* c, const float* params, const float* X, const float* t, int64_t n, float* grad,
                 double* loss) {
  if (!c || !params || !X || !t || !grad || n <= 0) return MRF_E_INVALID;
  if (n > 0xFFFFFFFFll) return fail(c, MRF_E_LIMIT, "mlp_grad: more than 2^32 samples");
  int rc = ensure_train(c, (size_t)n, (uint32_t)n);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_mlp, params, kMlpParams * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tX, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tT, t, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  launch_zero_words(c->d_tLoss, 8, c->stream);
  launch_mlp_grad(c->d_mlp, c->d_tX, c->d_tT, nullptr, 0, (uint32_t)n, 1.0f / (float)n, c->d_tPart, c->d_tLpart,
                  c->stream);
  launch_mlp_reduce(c->d_tPart, c->d_tLpart, mlp_grad_waves((uint32_t)n), c->d_tG, c->d_tLoss, 1.0 / (double)n,
                    c->stream);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(grad, c->d_tG, kMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
  double l = 0;
  HIPCHK(c, hipMemcpyAsync(&l, c->d_tLoss, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (loss) *loss = l;
  return MRF_OK;
}

int mrf_mlp_grad_noalloc(mrf_ctx* c, const float* params, const float* X, const float* t, int64_t n, float* grad,
                         double* loss, void* stream) {
  if (!c || !params || !X || !t || !grad || n <= 0) return MRF_E_INVALID;
  if (n > (int64_t)kNoAllocMaxSamples)
    return fail(c, MRF_E_LIMIT, "mlp_grad_noalloc: more than " + std::to_string(kNoAllocMaxSamples) +
                                    " samples (size of the fixed workspace)");
  const uint32_t nw = mlp_grad_waves((uint32_t)n);
  if (nw > kNoAllocWaves) return fail(c, MRF_E_LIMIT, "mlp_grad_noalloc: workspace too small");
  if (!c->scratch) return fail(c, MRF_E_STATE, "mlp_grad_noalloc: no workspace on this device");
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  // every wave's partials are written by k_mlp_grad before k_mlp_reduce reads
  // them, so nothing of the workspace needs clearing; the loss output is
  // cleared by a kernel (no memset node under graph capture)
  float* part = (float*)c->scratch;
  double* lpart = (double*)((unsigned char*)c->scratch + kNoAllocPartBytes);
  if (loss) launch_zero_words(loss, sizeof(double), s);
  launch_mlp_grad(params, X, t, nullptr, 0, (uint32_t)n, 1.0f / (float)n, part, lpart, s);
  launch_mlp_reduce(part, lpart, nw, grad, loss, 1.0 / (double)n, s);
  HIPCHK(c, hipGetLastError());
  return MRF_OK;
}

void mrf_train_opts_default(mrf_train_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->epochs = 20;
  o->batch = 8192;
  o->lr = 1e-3f;
  o->beta1 = 0.9f;
  o->beta2 = 0.999f;
  o->eps = 1e-8f;
  o->lr_decay = 1.0f;
  o->seed = 1;
}

int mrf_mlp_train(mrf_ctx* c, float* params, const float* X, const float* t, int64_t n, const mrf_train_opts* o,
                  double* epoch_loss) {
  if (!c || !params || !X || !t || !o || n <= 0 || o->epochs < 0 || o->batch <= 0) return MRF_E_INVALID;
  if (n > 0xFFFFFFFFll) return fail(c, MRF_E_LIMIT, "mlp_train: more than 2^32 samples");
  const uint32_t batch = (uint32_t)std::min<int64_t>(o->batch, n);
  int rc = ensure_train(c, (size_t)n, batch);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->d_mlp, params, kMlpParams * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tX, X, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->d_tT, t, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  launch_zero_words(c->d_tM, kMlpParams * 4, c->stream);
  launch_zero_words(c->d_tV, kMlpParams * 4, c->stream);
  std::vector<uint32_t> sel((size_t)n);
  std::mt19937_64 rng(o->seed);
  float lr = o->lr;
  double b1t = 1.0, b2t = 1.0;
  for (int e = 0; e < o->epochs; ++e) {
    for (size_t i = 0; i < sel.size(); ++i) sel[i] = (uint32_t)i;
    for (size_t i = sel.size() - 1; i > 0; --i) {  // Fisher-Yates
      const size_t j = (size_t)(rng() % (uint64_t)(i + 1));
      std::swap(sel[i], sel[j]);
    }
    // the previous epoch's kernels may still read d_tSel
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_tSel, sel.data(), sel.size() * 4, hipMemcpyHostToDevice, c->stream));
    launch_zero_words(c->d_tLoss, 8, c->stream);
    for (uint64_t off = 0; off < (uint64_t)n; off += batch) {
      const uint32_t cnt = (uint32_t)std::min<uint64_t>(batch, (uint64_t)n - off);
      launch_mlp_grad(c->d_mlp, c->d_tX, c->d_tT, c->d_tSel, (uint32_t)off, cnt, 1.0f / (float)cnt, c->d_tPart,
                      c->d_tLpart, c->stream);
      launch_mlp_reduce(c->d_tPart, c->d_tLpart, mlp_grad_waves(cnt), c->d_tG, c->d_tLoss, 1.0 / (double)n,
                        c->stream);
      b1t *= o->beta1;
      b2t *= o->beta2;
      launch_adam(c->d_mlp, c->d_tG, c->d_tM, c->d_tV, lr, o->beta1, o->beta2, o->eps, (float)(1.0 - b1t),
                  (float)(1.0 - b2t), c->stream);
    }
    HIPCHK(c, hipGetLastError());
    double l = 0;
    HIPCHK(c, hipMemcpyAsync(&l, c->d_tLoss, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (epoch_loss) epoch_loss[e] = l;
    lr *= o->lr_decay;
  }
  HIPCHK(c, hipMemcpyAsync(params, c->d_mlp, kMlpParams * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRF_OK;
}

int mrf_num_bricks(mrf_ctx* c, int64_t* n) {
  if (!c || !n) return MRF_E_INVALID;
  int rc = prepare(c);
  if (rc) return rc;
  *n = (int64_t)c->slot_lin.size();
  return MRF_OK;
}

int mrf_get_occupied_bricks(mrf_ctx* c, float* centers, int64_t n) {
  if (!c || (!centers && n)) return MRF_E_INVALID;
  int rc = prepare(c);
  if (rc) return rc;
  if (n != (int64_t)c->slot_lin.size()) return fail(c, MRF_E_SHAPE, "get_occupied_bricks: wrong row count");
  for (size_t s = 0; s < c->slot_lin.size(); ++s) {
    const int lin = c->slot_lin[s];
    const int bz = lin % c->g.bz, by = (lin / c->g.bz) % c->g.by, bx = lin / (c->g.bz * c->g.by);
    centers[3 * s + 0] = (float)(bx * 8 + 4);
    centers[3 * s + 1] = (float)(by * 8 + 4);
    centers[3 * s + 2] = (float)(bz * 8 + 4);
  }
  return MRF_OK;
}

static int read_pool(mrf_ctx* c, float* out, int64_t rows, int64_t cols, bool alphas) {
  if (!c || (!out && rows)) return MRF_E_INVALID;
  int rc = prepare(c);
  if (rc) return rc;
  if (rows != (int64_t)c->slot_lin.size() || cols != kBrickVox)
    return fail(c, MRF_E_SHAPE, "Dimensions for the alphas array are incorrect!");
  const size_t n = (size_t)rows * kBrickVox;
  if (n == 0) return MRF_OK;
  if (alphas) {
    launch_alphas(c->d_B, c->d_upd, c->d_out, n, c->stream);
    HIPCHK(c, hipMemcpyAsync(out, c->d_out, n * 4, hipMemcpyDeviceToHost, c->stream));
  } else {
    HIPCHK(c, hipMemcpyAsync(out, c->d_B, n * 4, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipGetLastError());
  return MRF_OK;
}

int mrf_get_alphas(mrf_ctx* c, float* out, int64_t rows, int64_t cols) { return read_pool(c, out, rows, cols, true); }
int mrf_get_beliefs(mrf_ctx* c, float* out, int64_t rows, int64_t cols) {
  return read_pool(c, out, rows, cols, false);
}

int mrf_get_cum_lengths(mrf_ctx* c, uint32_t id, float* out, int64_t rows, int64_t cols) {
  if (!c || (!out && rows)) return MRF_E_INVALID;
  if (id >= c->cams.size()) return fail(c, MRF_E_RANGE, "get_cum_lengths: camera id out of range");
  int rc = prepare(c);
  if (rc) return rc;
  if (rows != (int64_t)c->slot_lin.size() || cols != kBrickVox)
    return fail(c, MRF_E_SHAPE, "get_cum_lengths: the array must be num_bricks x 512");
  const size_t n = (size_t)rows * kBrickVox;
  if (n == 0) return MRF_OK;
  const Camera& cam = c->cams[id];
  launch_fill_f32(c->d_out, 0.0f, n, c->stream);
  if (cam.npairs) launch_scatter_pairs(cam.d_pair_vox, cam.d_pair_c, cam.npairs, c->d_out, c->stream);
  HIPCHK(c, hipMemcpyAsync(out, c->d_out, n * 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipGetLastError());
  return MRF_OK;
}

int mrf_get_traversal(mrf_ctx* c, uint32_t id, int64_t* n_rays, int64_t* n_records, uint32_t* rays,
                      uint32_t* records) {
  if (!c) return MRF_E_INVALID;
  if (id >= c->cams.size()) return fail(c, MRF_E_RANGE, "get_traversal: camera id out of range");
  int rc = prepare(c);
  if (rc) return rc;
  const Camera& cam = c->cams[id];
  if (n_rays) *n_rays = cam.nrays;
  if (n_records) *n_records = (int64_t)cam.nlist;
  if (!rays && !records) return MRF_OK;
  // compact the sliced lists back to per-ray lists in tile order
  const size_t npix = (size_t)cam.nslices * kSlice;
  std::vector<FrontRec> front(npix);
  std::vector<RayRec> lanes(npix);
  std::vector<uint2> sell(cam.nrec);
  HIPCHK(c, hipMemcpy(front.data(), cam.d_front, npix * sizeof(FrontRec), hipMemcpyDeviceToHost));
  HIPCHK(c, hipMemcpy(lanes.data(), cam.d_lanes, npix * sizeof(RayRec), hipMemcpyDeviceToHost));
  if (cam.nrec) {
    uint2* tmp = nullptr;
    if ((rc = dalloc(c, &tmp, cam.nrec))) return rc;
    launch_export_vox(cam.d_inc, (uint32_t)cam.nrec, cam.d_pair_vox, tmp, c->stream);
    hipError_t e = hipMemcpyAsync(sell.data(), tmp, cam.nrec * sizeof(uint2), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(&tmp);
    if (e != hipSuccess) return fail(c, MRF_E_HIP, std::string("get_traversal: ") + hipGetErrorString(e));
  }
  uint64_t off = 0, r = 0;
  for (size_t pix = 0; pix < npix; ++pix) {
    const FrontRec& F = front[pix];
    const RayRec& L = lanes[pix];
    const uint32_t n = F.nf + L.nw;
    if (n == 0) continue;
    if (rays) {
      uint32_t* o = rays + 6 * r;
      o[0] = (uint32_t)off;
      o[1] = n;
      std::memcpy(&o[2], &L.z0, 4);
      std::memcpy(&o[3], &L.inv_norm, 4);
      std::memcpy(&o[4], &L.depth, 4);
      o[5] = L.cell_flags & ~kInfTerm;  // lookup offset | kRanOut, as the oracle lists them
    }
    if (records)
      for (uint32_t k = 0; k < n; ++k) {
        const uint2 e = k < F.nf ? sell[F.off + (size_t)k * kSlice] : sell[L.off_w + (size_t)(k - F.nf) * kSlice];
        records[2 * (off + k)] = e.x;
        records[2 * (off + k) + 1] = e.y;
      }
    off += n;
    ++r;
  }
  return MRF_OK;
}

int mrf_evaluate_camera(mrf_ctx* c, uint32_t id, float* ll, float* ed, float* acc, float* sacc, mrf_eval* out) {
  if (!c) return MRF_E_INVALID;
  if (id >= c->cams.size()) return fail(c, MRF_E_RANGE, "evaluate_camera: camera id out of range");
  const CamGeom cg = make_geom(c, c->cams[id]);
  return evaluate(c, cg, nullptr, c->cams[id].d_depth, ll, ed, acc, sacc, out);
}

int mrf_evaluate_at_pose(mrf_ctx* c, const float pose[16], const float K[9], const float* depth, int rows, int cols,
                         int layout, float* ll, float* ed, float* acc, float* sacc, mrf_eval* out) {
  if (!c || !pose || rows <= 0 || cols <= 0 || (layout != MRF_ROW_MAJOR && layout != MRF_COL_MAJOR))
    return fail(c, MRF_E_INVALID, "evaluate_at_pose: invalid argument");
  const float* k = K ? K : c->p.K;
  if (k[0] <= 0 || k[4] <= 0) return fail(c, MRF_E_INVALID, "evaluate_at_pose: focal lengths must be positive");
  int rc = ensure_tables(c, rows, cols);
  if (rc) return rc;
  Camera cam;
  if (layout == MRF_COL_MAJOR) {
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) cam.pose[i * 4 + j] = pose[j * 4 + i];
  } else {
    std::memcpy(cam.pose, pose, sizeof(cam.pose));
  }
  std::memcpy(cam.K, k, sizeof(cam.K));
  cam.rows = rows;
  cam.cols = cols;
  std::vector<float> d;
  if (depth) to_row_major(depth, rows, cols, layout, d);
  const CamGeom cg = make_geom(c, cam);
  return evaluate(c, cg, depth ? d.data() : nullptr, nullptr, ll, ed, acc, sacc, out);
}

int mrf_load_map(mrf_ctx* c, const float* points, const float* occupancies, int64_t n) {
  if (!c || n < 0 || (n > 0 && (!points || !occupancies))) return fail(c, MRF_E_INVALID, "load_map: invalid argument");
  int rc;
  if ((rc = prepare(c))) return rc;
  const GridGeom& g = c->g;
  const int M[3] = {g.mx, g.my, g.mz}, N[3] = {g.nx, g.ny, g.nz};
  // 1. bricks within one voxel of every point in the map box
  std::vector<uint8_t> flags((size_t)c->nbricks, 0);
  std::vector<int64_t> cell((size_t)n, -1);  // voxel of each point as x, y, z packed, -1 outside
  for (int64_t i = 0; i < n; ++i) {
    int lo[3], hi[3], v[3];
    bool ok = true;
    for (int a = 0; a < 3 && ok; ++a) {
      const double P = (double)points[3 * i + a];
      if (!(P >= 0.0 && P < (double)M[a])) {
        ok = false;
        break;
      }
      v[a] = (int)std::floor(P);
      lo[a] = std::max(0, (int)std::floor(P - 1.0)) >> kBrickLog;
      hi[a] = std::min(N[a] - 1, (int)std::floor(P + 1.0)) >> kBrickLog;
    }
    if (!ok) continue;
    for (int x = lo[0]; x <= hi[0]; ++x)
      for (int y = lo[1]; y <= hi[1]; ++y)
        for (int z = lo[2]; z <= hi[2]; ++z) flags[((size_t)x * g.by + y) * g.bz + z] = 1;
    cell[i] = ((int64_t)v[0] * N[1] + v[1]) * N[2] + v[2];
  }
  bool grew = false;
  for (int lin = 0; lin < c->nbricks; ++lin)
    if (flags[lin] && c->h_slot[lin] < 0) {
      c->h_slot[lin] = (int32_t)c->slot_lin.size();
      c->slot_lin.push_back(lin);
      grew = true;
    }
  if ((rc = ensure_pool(c, c->slot_lin.size()))) return rc;
  if (grew) {
    HIPCHK(c, hipMemcpyAsync(c->d_slot, c->h_slot.data(), c->h_slot.size() * 4, hipMemcpyHostToDevice, c->stream));
    c->geom_dirty = true;  // the cameras' rays now cross more active bricks
  }
  // 2. beliefs: logit of the occupancy (clamped like every marginal); the
  //    last point in a voxel wins (set_channel_kernel, gvdbkernels.cu:1040-1064)
  std::vector<std::pair<uint32_t, float>> vals;
  vals.reserve((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    if (cell[i] < 0) continue;
    const int64_t cz = cell[i] % N[2], cy = (cell[i] / N[2]) % N[1], cx = cell[i] / ((int64_t)N[1] * N[2]);
    const int lin = (int)(((cx >> kBrickLog) * g.by + (cy >> kBrickLog)) * g.bz + (cz >> kBrickLog));
    const uint32_t local = (uint32_t)((((cx & 7) << 3) | (cy & 7)) << 3 | (cz & 7));
    const float o = std::min(std::max(occupancies[i], kFEps), 1.0f - kFEps);
    vals.emplace_back((uint32_t)c->h_slot[lin] * kBrickVox + local, logf(o) - logf(1.0f - o));
  }
  std::stable_sort(vals.begin(), vals.end(),
                   [](const std::pair<uint32_t, float>& a, const std::pair<uint32_t, float>& b) { return a.first < b.first; });
  std::vector<uint32_t> vox;
  std::vector<float> lo;
  for (size_t i = 0; i < vals.size(); ++i)
    if (i + 1 == vals.size() || vals[i + 1].first != vals[i].first) {
      vox.push_back(vals[i].first);
      lo.push_back(vals[i].second);
    }
  if (!vox.empty()) {
    uint32_t* d_vox = nullptr;
    float* d_lo = nullptr;
    if ((rc = dalloc(c, &d_vox, vox.size())) || (rc = dalloc(c, &d_lo, lo.size()))) {
      dfree(&d_vox);
      return rc;
    }
    hipError_t e = hipMemcpyAsync(d_vox, vox.data(), vox.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(d_lo, lo.data(), lo.size() * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) launch_set_voxels(d_vox, d_lo, (uint32_t)vox.size(), c->d_B, c->d_upd, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipGetLastError();
    dfree(&d_vox);
    dfree(&d_lo);
    if (e != hipSuccess) return fail(c, MRF_E_HIP, std::string("load_map: ") + hipGetErrorString(e));
  }
  return MRF_OK;
}

int mrf_set_launch_mode(mrf_ctx* c, int mode) {
  if (!c) return MRF_E_INVALID;
  if (mode != MRF_LAUNCH_GRAPH && mode != MRF_LAUNCH_TIMED) return fail(c, MRF_E_INVALID, "unknown launch mode");
  c->launch_mode = mode;
  return MRF_OK;
}

int mrf_set_eval_thresholds(mrf_ctx* c, float occ_thresh, float acc_thresh) {
  if (!c) return MRF_E_INVALID;
  c->p.occ_thresh = occ_thresh;
  c->p.acc_thresh = acc_thresh;
  return MRF_OK;
}

int mrf_release_device_cache(void) {
  BlockCache& bc = block_cache();
  std::lock_guard<std::mutex> lock(bc.mu);
  release_idle(bc, -1);
  return MRF_OK;
}

const char* mrf_last_error(mrf_ctx* c) { return c ? c->err.c_str() : "null context"; }

}  // extern "C"
