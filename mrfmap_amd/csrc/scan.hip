// Device-wide exclusive prefix sums of uint32 values, deterministic by
// construction.
//  * launch_exclusive_scan: three launches (block totals, one-block scan of
//    the totals, block-local scan + offset) with a 64-bit total, for counts
//    whose sum may exceed 32 bits (list offsets; the callers check limits).
//  * launch_flag_scan: 0/1 flags, whose sum is at most n, so a 32-bit scan is
//    exact: hipCUB's single-pass scan and a one-thread total; used for the
//    brick-pool-sized flag arrays, where the three-launch scan's one-block
//    middle step dominates.
//  * launch_u64_scan: the same single-pass scan over uint64 values (the
//    per-brick counts of two flags, packed).
#include <hipcub/hipcub.hpp>

#include "mrf_kernels.h"

namespace mrf {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 8;
constexpr int kPerBlock = kThreads * kItems;

__device__ uint64_t block_excl_scan(uint64_t v, uint64_t* sh, uint64_t* total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = 1; off < kThreads; off <<= 1) {
    uint64_t add = t >= off ? sh[t - off] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  uint64_t incl = sh[t];
  *total = sh[kThreads - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kThreads) void k_block_sums(const uint32_t* __restrict__ in, size_t n,
                                                          uint64_t* __restrict__ sums) {
  __shared__ uint64_t sh[kThreads];
  const size_t base = (size_t)blockIdx.x * kPerBlock;
  uint64_t acc = 0;
  for (int k = 0; k < kItems; ++k) {
    size_t i = base + (size_t)k * kThreads + threadIdx.x;
    if (i < n) acc += in[i];
  }
  uint64_t tot;
  block_excl_scan(acc, sh, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kThreads) void k_scan_sums(uint64_t* __restrict__ sums, size_t nb,
                                                         uint64_t* __restrict__ d_total) {
  __shared__ uint64_t sh[kThreads];
  uint64_t carry = 0;
  for (size_t base = 0; base < nb; base += kThreads) {
    size_t i = base + threadIdx.x;
    uint64_t v = i < nb ? sums[i] : 0;
    uint64_t tot;
    uint64_t ex = block_excl_scan(v, sh, &tot);
    if (i < nb) sums[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) *d_total = carry;
}

__global__ __launch_bounds__(kThreads) void k_apply(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                     size_t n, const uint64_t* __restrict__ offs) {
  __shared__ uint64_t sh[kThreads];
  // each thread owns kItems consecutive items
  const size_t base = (size_t)blockIdx.x * kPerBlock + (size_t)threadIdx.x * kItems;
  uint32_t v[kItems];
  uint64_t acc = 0;
  for (int k = 0; k < kItems; ++k) {
    size_t i = base + k;
    v[k] = i < n ? in[i] : 0u;
    acc += v[k];
  }
  uint64_t tot;
  uint64_t run = offs[blockIdx.x] + block_excl_scan(acc, sh, &tot);
  for (int k = 0; k < kItems; ++k) {
    size_t i = base + k;
    if (i < n) out[i] = (uint32_t)run;
    run += v[k];
  }
}

__global__ void k_flag_total(const uint32_t* __restrict__ in, const uint32_t* __restrict__ out, size_t n,
                             uint64_t* __restrict__ d_total) {
  if (threadIdx.x == 0) *d_total = (uint64_t)out[n - 1] + in[n - 1];
}

__global__ void k_u64_total(const uint64_t* __restrict__ in, const uint64_t* __restrict__ out, size_t n,
                            uint64_t* __restrict__ d_total) {
  if (threadIdx.x == 0) *d_total = out[n - 1] + in[n - 1];
}

}  // namespace

size_t flag_scan_temp_bytes(size_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n);
  return bytes;
}

hipError_t launch_flag_scan(const uint32_t* in, uint32_t* out, size_t n, void* temp, size_t temp_bytes,
                            uint64_t* d_total, hipStream_t s) {
  if (n == 0) {
    launch_zero_words(d_total, sizeof(uint64_t), s);
    return hipGetLastError();
  }
  size_t tb = temp_bytes;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(temp, tb, in, out, (int)n, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_flag_total, dim3(1), dim3(64), 0, s, in, out, n, d_total);
  return hipGetLastError();
}

size_t u64_scan_temp_bytes(size_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int)n);
  return bytes;
}

hipError_t launch_u64_scan(const uint64_t* in, uint64_t* out, size_t n, void* temp, size_t temp_bytes,
                           uint64_t* d_total, hipStream_t s) {
  if (n == 0) {
    launch_zero_words(d_total, sizeof(uint64_t), s);
    return hipGetLastError();
  }
  size_t tb = temp_bytes;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(temp, tb, in, out, (int)n, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_u64_total, dim3(1), dim3(64), 0, s, in, out, n, d_total);
  return hipGetLastError();
}

size_t scan_tmp_words(size_t n) { return (n + kPerBlock - 1) / kPerBlock + 1; }

void launch_exclusive_scan(const uint32_t* in, uint32_t* out, size_t n, uint64_t* tmp, uint64_t* d_total,
                           hipStream_t s) {
  const size_t nb = (n + kPerBlock - 1) / kPerBlock;
  if (nb == 0) {
    (void)hipMemsetAsync(d_total, 0, sizeof(uint64_t), s);
    return;
  }
  hipLaunchKernelGGL(k_block_sums, dim3((unsigned)nb), dim3(kThreads), 0, s, in, n, tmp);
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(kThreads), 0, s, tmp, nb, d_total);
  hipLaunchKernelGGL(k_apply, dim3((unsigned)nb), dim3(kThreads), 0, s, in, out, n, tmp);
}

}  // namespace mrf
