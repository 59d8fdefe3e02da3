// radix_sort.hip -- see radix_sort.h.
#include <hip/hip_runtime.h>

#include "gz_device.h"
#include "gz_scan.h"
#include "radix_sort.h"

namespace gz {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;                          // per thread
constexpr size_t kSortTile = kSortThreads * kSortItems;  // 4096
constexpr int kSubChunks = kSortTile / kWave;            // 64 runs of 64 items
constexpr int kDigits = 256;

__global__ void __launch_bounds__(kSortThreads)
    DigitHistogramKernel(const uint32_t *keys, size_t n, int shift, uint32_t tiles,
                         uint32_t *hist) {
  __shared__ uint32_t count[kDigits];
  count[threadIdx.x] = 0;
  __syncthreads();
  const size_t base = (size_t)blockIdx.x * kSortTile;
#pragma unroll 4
  for (int j = 0; j < kSortItems; ++j) {
    const size_t i = base + (size_t)j * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&count[(keys[i] >> shift) & 0xffu], 1u);
  }
  __syncthreads();
  hist[(size_t)threadIdx.x * tiles + blockIdx.x] = count[threadIdx.x];
}

// Item s * 64 + lane of the tile is handled by wave s % 4 as its item s / 4.
__global__ void __launch_bounds__(kSortThreads)
    DigitScatterKernel(const uint32_t *keys, const uint32_t *values, size_t n, int shift,
                       uint32_t tiles, const uint32_t *offsets, uint32_t *keys_out,
                       uint32_t *values_out) {
  __shared__ uint16_t count[kSubChunks][kDigits];  // 32 KB
  for (int i = threadIdx.x; i < kSubChunks * kDigits; i += kSortThreads)
    (&count[0][0])[i] = 0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const size_t base = (size_t)blockIdx.x * kSortTile;
  uint32_t key[kSortItems], val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int s = j * 4 + wave;
    const size_t i = base + (size_t)s * kWave + lane;
    const bool valid = i < n;
    key[j] = valid ? keys[i] : 0u;
    val[j] = valid ? values[i] : 0u;
    const uint32_t d = (key[j] >> shift) & 0xffu;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    rank[j] = (uint32_t)__popcll(peers & lt);
    if (valid && (peers & lt) == 0) count[s][d] = (uint16_t)__popcll(peers);
  }
  __syncthreads();
  {
    // exclusive prefix over the runs, per digit
    const int d = threadIdx.x;
    uint32_t run = 0;
    for (int s = 0; s < kSubChunks; ++s) {
      const uint32_t c = count[s][d];
      count[s][d] = (uint16_t)run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int s = j * 4 + wave;
    const size_t i = base + (size_t)s * kWave + lane;
    if (i < n) {
      const uint32_t d = (key[j] >> shift) & 0xffu;
      const size_t o = (size_t)offsets[(size_t)d * tiles + blockIdx.x] + count[s][d] + rank[j];
      keys_out[o] = key[j];
      values_out[o] = val[j];
    }
  }
}

}  // namespace

size_t RadixSortTempBytes(size_t n) {
  const size_t tiles = (n + kSortTile - 1) / kSortTile;
  const size_t h = AlignUp(tiles * kDigits * sizeof(uint32_t), 256);
  return 2 * h + DeviceScanTempBytes<uint32_t>(tiles * kDigits);
}

hipError_t RadixSortPairs(uint32_t *keys, uint32_t *values, uint32_t *keys_alt,
                          uint32_t *values_alt, size_t n, int bits, void *temp,
                          hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (bits <= 0 || bits > 32 || bits % 16 != 0 || n > 0xffffffffull) return hipErrorInvalidValue;
  const size_t tiles = (n + kSortTile - 1) / kSortTile;
  const size_t h = AlignUp(tiles * kDigits * sizeof(uint32_t), 256);
  uint32_t *hist = (uint32_t *)temp;
  uint32_t *offs = (uint32_t *)((char *)temp + h);
  void *scan_tmp = (char *)temp + 2 * h;
  uint32_t *src_k = keys, *src_v = values, *dst_k = keys_alt, *dst_v = values_alt;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(DigitHistogramKernel, dim3((unsigned)tiles), dim3(kSortThreads), 0, s,
                       src_k, n, shift, (uint32_t)tiles, hist);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = DeviceScan<uint32_t, false>(hist, offs, tiles * kDigits, false, scan_tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(DigitScatterKernel, dim3((unsigned)tiles), dim3(kSortThreads), 0, s,
                       src_k, src_v, n, shift, (uint32_t)tiles, offs, dst_k, dst_v);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    uint32_t *t = src_k;
    src_k = dst_k;
    dst_k = t;
    t = src_v;
    src_v = dst_v;
    dst_v = t;
  }
  return hipSuccess;
}

}  // namespace gz
