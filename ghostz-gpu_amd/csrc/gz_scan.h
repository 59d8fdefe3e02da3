// gz_scan.h -- device-wide scans (the project's replacement for the vendored
// cub::DeviceScan::InclusiveScan(Max) of src/ungapped_extender_gpu.cu:838-841
// and src/device_seeds_memory.cu:28-34).
//
// Reduce-then-scan over tiles of 1024 items (256 threads x 4), recursing on
// the tile aggregates.  Unsigned element types only (identity 0 for both
// sum and max).
#pragma once

#include "gz_device.h"

namespace gz {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr size_t kScanTile = kScanThreads * kScanItems;

template <typename T, bool kMax>
__global__ void __launch_bounds__(kScanThreads)
    ScanTileReduceKernel(const T *in, size_t n, T *aggregates) {
  __shared__ T lds[kScanThreads / kWave];
  const size_t base = (size_t)blockIdx.x * kScanTile;
  T acc = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const size_t i = base + threadIdx.x + (size_t)j * kScanThreads;
    if (i < n) {
      const T v = in[i];
      acc = kMax ? (v > acc ? v : acc) : acc + v;
    }
  }
  T inc, exc, tot;
  BlockScan<T, kMax>(acc, lds, &inc, &exc, &tot);
  if (threadIdx.x == 0) aggregates[blockIdx.x] = tot;
}

// Each thread owns kScanItems consecutive items.  prefix (optional) holds
// the exclusive scan of the tile aggregates.  Safe in place (in == out).
template <typename T, bool kMax, bool kInclusive>
__global__ void __launch_bounds__(kScanThreads)
    ScanTileKernel(const T *in, T *out, size_t n, const T *prefix) {
  __shared__ T lds[kScanThreads / kWave];
  const size_t base =
      (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
  T v[kScanItems];
  T run = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = (base + j < n) ? in[base + j] : T(0);
    run = kMax ? (v[j] > run ? v[j] : run) : run + v[j];
  }
  T inc, exc;
  BlockScan<T, kMax>(run, lds, &inc, &exc, nullptr);
  T acc = exc;
  if (prefix) {
    const T p = prefix[blockIdx.x];
    acc = kMax ? (p > acc ? p : acc) : acc + p;
  }
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    const T before = acc;
    acc = kMax ? (v[j] > acc ? v[j] : acc) : acc + v[j];
    if (base + j < n) out[base + j] = kInclusive ? acc : before;
  }
}

inline size_t AlignUp(size_t x, size_t a) { return (x + a - 1) / a * a; }

template <typename T>
inline size_t DeviceScanTempBytes(size_t n) {
  size_t bytes = 0;
  size_t tiles = (n + kScanTile - 1) / kScanTile;
  while (tiles > 1) {
    bytes += AlignUp(tiles * sizeof(T), 256);
    tiles = (tiles + kScanTile - 1) / kScanTile;
  }
  return bytes + 256;
}

template <typename T, bool kMax>
inline hipError_t DeviceScan(const T *in, T *out, size_t n, bool inclusive,
                             void *temp, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const size_t tiles = (n + kScanTile - 1) / kScanTile;
  T *agg = nullptr;
  if (tiles > 1) {
    agg = reinterpret_cast<T *>(temp);
    void *next = reinterpret_cast<char *>(temp) +
                 AlignUp(tiles * sizeof(T), 256);
    hipLaunchKernelGGL((ScanTileReduceKernel<T, kMax>), dim3(tiles),
                       dim3(kScanThreads), 0, stream, in, n, agg);
    hipError_t e = DeviceScan<T, kMax>(agg, agg, tiles, false, next, stream);
    if (e != hipSuccess) return e;
  }
  if (inclusive) {
    hipLaunchKernelGGL((ScanTileKernel<T, kMax, true>), dim3(tiles),
                       dim3(kScanThreads), 0, stream, in, out, n, agg);
  } else {
    hipLaunchKernelGGL((ScanTileKernel<T, kMax, false>), dim3(tiles),
                       dim3(kScanThreads), 0, stream, in, out, n, agg);
  }
  return hipGetLastError();
}

}  // namespace gz
