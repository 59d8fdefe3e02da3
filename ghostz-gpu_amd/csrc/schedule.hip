// schedule.hip -- see schedule.h.
#include <hip/hip_runtime.h>

#include "gz_device.h"
#include "gz_scan.h"
#include "schedule.h"

namespace gz {

namespace {

constexpr int kOrderThreads = 256;
constexpr uint32_t kOrderChunk = 32768;  // jobs per block (few global bin atomics per job)

__global__ void __launch_bounds__(kOrderThreads)
    BinHistogramKernel(const uint16_t *keys, uint32_t n, uint32_t *counts) {
  __shared__ uint32_t hist[kScheduleBins];
  for (uint32_t b = threadIdx.x; b < kScheduleBins; b += kOrderThreads) hist[b] = 0;
  __syncthreads();
  const uint32_t begin = blockIdx.x * kOrderChunk;
  const uint32_t end = min(n, begin + kOrderChunk);
  for (uint32_t i = begin + threadIdx.x; i < end; i += kOrderThreads)
    atomicAdd(&hist[keys[i]], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < kScheduleBins; b += kOrderThreads)
    if (hist[b]) atomicAdd(&counts[b], hist[b]);
}

// cursor[b] = number of jobs in bins above b (descending layout); one block
__global__ void __launch_bounds__(kOrderThreads)
    BinScanKernel(const uint32_t *counts, uint32_t *cursor) {
  __shared__ uint32_t lds[kOrderThreads / kWave];
  constexpr uint32_t kPer = kScheduleBins / kOrderThreads;
  // thread t owns bins in descending order: kScheduleBins-1-(t*kPer+j)
  uint32_t v[kPer];
  uint32_t sum = 0;
  for (uint32_t j = 0; j < kPer; ++j) {
    v[j] = counts[kScheduleBins - 1 - (threadIdx.x * kPer + j)];
    sum += v[j];
  }
  uint32_t inc, exc;
  BlockScan<uint32_t, false>(sum, lds, &inc, &exc, nullptr);
  uint32_t run = exc;
  for (uint32_t j = 0; j < kPer; ++j) {
    cursor[kScheduleBins - 1 - (threadIdx.x * kPer + j)] = run;
    run += v[j];
  }
}

__global__ void __launch_bounds__(kOrderThreads)
    BinScatterKernel(const uint16_t *keys, uint32_t n, uint32_t *cursor, uint32_t *perm) {
  __shared__ uint32_t hist[kScheduleBins];
  for (uint32_t b = threadIdx.x; b < kScheduleBins; b += kOrderThreads) hist[b] = 0;
  __syncthreads();
  const uint32_t begin = blockIdx.x * kOrderChunk;
  const uint32_t end = min(n, begin + kOrderChunk);
  for (uint32_t i = begin + threadIdx.x; i < end; i += kOrderThreads)
    atomicAdd(&hist[keys[i]], 1u);
  __syncthreads();
  // reserve this block's range of every bin it uses
  for (uint32_t b = threadIdx.x; b < kScheduleBins; b += kOrderThreads)
    if (hist[b]) hist[b] = atomicAdd(&cursor[b], hist[b]);
  __syncthreads();
  for (uint32_t i = begin + threadIdx.x; i < end; i += kOrderThreads) {
    const uint32_t pos = atomicAdd(&hist[keys[i]], 1u);
    perm[pos] = i;
  }
}

__global__ void ZeroKernel(uint32_t *p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

__global__ void CostToKeyKernel(const float *cost, uint32_t n, float scale, uint16_t *keys) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = cost[i] * scale;
  int k = (x == x) ? (int)lrintf(fminf(fmaxf(x, 0.f), (float)(kScheduleBins - 1))) : 0;
  keys[i] = (uint16_t)k;
}

__global__ void ClearCountKernel(uint32_t *count) {
  if (threadIdx.x == 0) *count = 0;
}

// Length bins and the gate of one side in one pass: every block takes
// kGateChunk consecutive jobs (kGateItems per thread, strided so that the
// loads coalesce), writes their length-bin keys, and appends the gated ones
// with one atomic on the side's count per block (a block scan ranks them).
constexpr uint32_t kGateItems = 16;
constexpr uint32_t kGateChunk = 256 * kGateItems;

__global__ void __launch_bounds__(256)
    BinGateKernel(const uint16_t *length, uint32_t n, uint32_t gate, uint16_t *keys,
                  uint32_t *index, uint32_t *count) {
  __shared__ uint32_t lds[256 / kWave];
  __shared__ uint32_t base_lds;
  const uint32_t begin = blockIdx.x * kGateChunk;
  uint32_t take = 0;  // bit r: item r of this thread is gated
#pragma unroll
  for (uint32_t r = 0; r < kGateItems; ++r) {
    const uint32_t i = begin + r * 256 + threadIdx.x;
    if (i < n) {
      const uint32_t len = length[i];
      keys[i] = (uint16_t)(min(len, kLengthBins - 1) << 6);
      if (len >= gate) take |= 1u << r;
    }
  }
  uint32_t inc, exc, total;
  BlockScan<uint32_t, false>((uint32_t)__popc(take), lds, &inc, &exc, &total);
  if (total == 0) return;  // block-uniform
  if (threadIdx.x == 0) base_lds = atomicAdd(count, total);
  __syncthreads();
  uint32_t at = base_lds + exc;
#pragma unroll
  for (uint32_t r = 0; r < kGateItems; ++r)
    if ((take >> r) & 1u) index[at++] = begin + r * 256 + threadIdx.x;
}

__global__ void PredictionBinKernel(const float *pred, const uint32_t *index, uint32_t m,
                                    const uint32_t *count, uint16_t *keys) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (count) m = min(m, *count);
  if (k >= m) return;
  const float x = pred[k] * 4.f;
  const int b = (x == x) ? (int)fminf(fmaxf(floorf(x), 0.f), (float)(kLengthBins - 1)) : 0;
  const uint32_t j = index[k];
  keys[j] = (uint16_t)(keys[j] | (uint32_t)b);
}

inline unsigned Blocks(uint32_t n, unsigned t = 256) { return (n + t - 1) / t; }

}  // namespace

size_t ScheduleTempBytes(size_t) { return 2 * kScheduleBins * sizeof(uint32_t); }

hipError_t DeviceOrderByKey(const uint16_t *keys, uint32_t n, uint32_t *perm, void *temp,
                            hipStream_t s) {
  if (n == 0) return hipSuccess;
  uint32_t *counts = (uint32_t *)temp;
  uint32_t *cursor = counts + kScheduleBins;
  const unsigned blocks = (n + kOrderChunk - 1) / kOrderChunk;
  hipLaunchKernelGGL(ZeroKernel, dim3(Blocks(kScheduleBins)), dim3(256), 0, s, counts,
                     kScheduleBins);
  hipLaunchKernelGGL(BinHistogramKernel, dim3(blocks), dim3(kOrderThreads), 0, s, keys, n,
                     counts);
  hipLaunchKernelGGL(BinScanKernel, dim3(1), dim3(kOrderThreads), 0, s, counts, cursor);
  hipLaunchKernelGGL(BinScatterKernel, dim3(blocks), dim3(kOrderThreads), 0, s, keys, n, cursor,
                     perm);
  return hipGetLastError();
}

hipError_t CostToKey(const float *cost, uint32_t n, float scale, uint16_t *keys, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(CostToKeyKernel, dim3(Blocks(n)), dim3(256), 0, s, cost, n, scale, keys);
  return hipGetLastError();
}

hipError_t BinAndGateJobs(const uint16_t *length, uint32_t n, uint32_t gate, uint16_t *keys,
                          uint32_t *index, uint32_t *d_count, hipStream_t s) {
  hipLaunchKernelGGL(ClearCountKernel, dim3(1), dim3(64), 0, s, d_count);
  if (n > 0)
    hipLaunchKernelGGL(BinGateKernel, dim3((n + kGateChunk - 1) / kGateChunk), dim3(256), 0, s,
                       length, n, gate, keys, index, d_count);
  return hipGetLastError();
}

hipError_t AddPredictionBins(const float *pred, const uint32_t *index, uint32_t m,
                             uint16_t *keys, hipStream_t s, const uint32_t *d_count) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(PredictionBinKernel, dim3(Blocks(m)), dim3(256), 0, s, pred, index, m,
                     d_count, keys);
  return hipGetLastError();
}

}  // namespace gz
