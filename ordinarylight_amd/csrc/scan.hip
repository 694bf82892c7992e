// Exclusive prefix sum of uint32 arrays (three-phase: tile scan, recursive
// scan of tile totals, add).  Used by the radix sort, the BVH collapse and
// the order-preserving compaction of the reference-estimator mode.

#include "internal.h"

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr uint32_t kScanTile = kScanBlock * kScanItems;

__global__ void __launch_bounds__(kScanBlock) k_scan_tile(uint32_t* data, uint32_t n, uint32_t* tile_sums) {
  __shared__ uint32_t wave_tot[kScanBlock / 64];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint32_t i = base + k;
    v[k] = i < n ? data[i] : 0u;
    sum += v[k];
  }
  // Wave-inclusive scan of per-thread sums.
  uint32_t incl = sum;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t o = __shfl_up(incl, (unsigned)off, 64);
    if ((int)lane_id() >= off) incl += o;
  }
  const uint32_t wave = threadIdx.x / 64;
  if (lane_id() == 63) wave_tot[wave] = incl;
  __syncthreads();
  uint32_t wave_base = 0, block_total = 0;
  for (uint32_t w = 0; w < kScanBlock / 64; ++w) {
    if (w < wave) wave_base += wave_tot[w];
    block_total += wave_tot[w];
  }
  uint32_t run = wave_base + incl - sum;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    uint32_t i = base + k;
    if (i < n) data[i] = run;
    run += v[k];
  }
  if (threadIdx.x == 0 && tile_sums) tile_sums[blockIdx.x] = block_total;
}

__global__ void __launch_bounds__(kScanBlock) k_scan_add(uint32_t* data, uint32_t n, const uint32_t* tile_offsets) {
  const uint32_t add = tile_offsets[blockIdx.x];
  const uint32_t base = blockIdx.x * kScanTile;
  for (uint32_t k = threadIdx.x; k < kScanTile; k += kScanBlock) {
    uint32_t i = base + k;
    if (i < n) data[i] += add;
  }
}

__global__ void k_scan_total(const uint32_t* data, const uint32_t* last_in, uint32_t n, uint32_t* total) {
  // total = exclusive[n-1] + original[n-1]
  *total = n ? data[n - 1] + *last_in : 0u;
}

__global__ void __launch_bounds__(256) k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) dst[i] = src[i];
}

}  // namespace

void stream_copy(const float4* src, float4* dst, uint64_t n16, int cu_count, hipStream_t s) {
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)(cu_count * 8)), dim3(256), 0, s, src, dst, n16);
}

void scan_exclusive_u32(uint32_t* data, uint32_t n, uint32_t* total, hipStream_t s) {
  if (n == 0) {
    if (total) OL_HIP_CHECK(hipMemsetAsync(total, 0, 4, s));
    return;
  }
  DevBuf last;
  if (total) {
    last.reserve(4);
    OL_HIP_CHECK(hipMemcpyAsync(last.ptr, data + (n - 1), 4, hipMemcpyDeviceToDevice, s));
  }
  const uint32_t tiles = (n + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_scan_tile, dim3(1), dim3(kScanBlock), 0, s, data, n, (uint32_t*)nullptr);
  } else {
    DevBuf sums;
    sums.reserve(4ull * tiles);
    hipLaunchKernelGGL(k_scan_tile, dim3(tiles), dim3(kScanBlock), 0, s, data, n, sums.as<uint32_t>());
    scan_exclusive_u32(sums.as<uint32_t>(), tiles, nullptr, s);
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kScanBlock), 0, s, data, n, sums.as<uint32_t>());
    OL_HIP_CHECK(hipStreamSynchronize(s));  // sums is freed on return
  }
  if (total) {
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(1), 0, s, data, last.as<uint32_t>(), n, total);
    OL_HIP_CHECK(hipStreamSynchronize(s));  // last is freed on return
  }
}
