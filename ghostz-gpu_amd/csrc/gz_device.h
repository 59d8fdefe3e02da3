// gz_device.h -- device-side building blocks for gfx950 (wave64).
//
//  * packed-alphabet cursors that stream 5-bit codes out of 64-bit words in
//    either direction (the reference's InitPackedAlphabetCode /
//    GetPackedAlphabetCode<Dir> pair, src/packed_alphabet_code.h:22-142,
//    restated as a small register-resident cursor);
//  * wave64 and block scans built on DPP/shuffle, used for seed expansion
//    (inclusive max-scan) and ballot/popcount compaction (exclusive sum).
#pragma once

#include <hip/hip_runtime.h>

#include "gz_common.h"

namespace gz {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t PackedCodeAt(const uint64_t *packed,
                                                 uint32_t position) {
  const uint32_t p = position + kPackPadCodes;
  const uint32_t w = p / kCodesPerWord;
  const uint32_t k = p - w * kCodesPerWord;
  return (uint32_t)(packed[w] >> (kCodeBits * k)) & 31u;
}

// Streams codes at position, position+1, ...
struct ForwardCursor {
  const uint64_t *base;
  uint64_t cur;
  uint32_t w;
  uint32_t k;
  __device__ __forceinline__ void Init(const uint64_t *b, uint32_t position) {
    base = b;
    const uint32_t p = position + kPackPadCodes;
    w = p / kCodesPerWord;
    k = p - w * kCodesPerWord;
    cur = b[w] >> (kCodeBits * k);
  }
  __device__ __forceinline__ uint32_t Next() {
    const uint32_t c = (uint32_t)cur & 31u;
    if (++k == kCodesPerWord) {
      k = 0;
      cur = base[++w];
    } else {
      cur >>= kCodeBits;
    }
    return c;
  }
};

// Streams codes at position, position-1, ...
struct ReverseCursor {
  const uint64_t *base;
  uint64_t cur;
  uint32_t w;
  int k;
  __device__ __forceinline__ void Init(const uint64_t *b, uint32_t position) {
    base = b;
    const uint32_t p = position + kPackPadCodes;
    w = p / kCodesPerWord;
    k = (int)(p - w * kCodesPerWord);
    cur = b[w];
  }
  __device__ __forceinline__ uint32_t Next() {
    const uint32_t c = (uint32_t)(cur >> (kCodeBits * k)) & 31u;
    if (k == 0) {
      k = kCodesPerWord - 1;
      cur = base[--w];
    } else {
      --k;
    }
    return c;
  }
};

// Copies an nl x nl int matrix into a 32 x 32 int8 LDS table (zero padded,
// so delimiter rows/columns score 0 as in src/aligner_gpu_data.cu:108-122).
__device__ __forceinline__ void StageScoreTable(int8_t *lds, const int *matrix,
                                                uint32_t nl) {
  for (uint32_t i = threadIdx.x; i < kTableDim * kTableDim; i += blockDim.x) {
    const uint32_t r = i / kTableDim, c = i % kTableDim;
    lds[i] = (r < nl && c < nl) ? (int8_t)matrix[r * nl + c] : (int8_t)0;
  }
}

// ---- wave64 scans (lane order = index order) -------------------------------
template <typename T>
__device__ __forceinline__ T WaveInclusiveSum(T v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T o = __shfl_up(v, d, kWave);
    if (lane >= d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T WaveInclusiveMax(T v) {
  const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T o = __shfl_up(v, d, kWave);
    if (lane >= d) v = (o > v) ? o : v;
  }
  return v;
}

// Inclusive max-scan of an int over the wave in DPP moves (no LDS):
// Hillis-Steele inside each row of 16 lanes (row_shr 1, 2, 4, 8), then each
// row's last lane carried into the rows after it (row_bcast 15 into rows 1
// and 3, row_bcast 31 into rows 2 and 3).  A lane whose source lies outside
// its row keeps `identity`, the smallest value the scan sees.
__device__ __forceinline__ int WaveInclusiveMaxDpp(int v, int identity) {
  int t = __builtin_amdgcn_update_dpp(identity, v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v = v < t ? t : v;
  t = __builtin_amdgcn_update_dpp(identity, v, 0x112, 0xf, 0xf, false);      // row_shr:2
  v = v < t ? t : v;
  t = __builtin_amdgcn_update_dpp(identity, v, 0x114, 0xf, 0xf, false);      // row_shr:4
  v = v < t ? t : v;
  t = __builtin_amdgcn_update_dpp(identity, v, 0x118, 0xf, 0xf, false);      // row_shr:8
  v = v < t ? t : v;
  t = __builtin_amdgcn_update_dpp(identity, v, 0x142, 0xa, 0xf, false);      // row_bcast:15
  v = v < t ? t : v;
  t = __builtin_amdgcn_update_dpp(identity, v, 0x143, 0xc, 0xf, false);      // row_bcast:31
  v = v < t ? t : v;
  return v;
}

// Lane i - 1's value (lane 0: fill), one DPP move (wave_shr:1).
__device__ __forceinline__ int WaveShiftUp1(int v, int fill) {
  return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xf, 0xf, false);
}

// Block-wide scan of one value per thread (unsigned types: identity 0).
// lds must hold blockDim.x / 64 elements; blockDim.x a multiple of 64.
template <typename T, bool kMax>
__device__ __forceinline__ void BlockScan(T v, T *lds, T *inclusive,
                                          T *exclusive, T *total) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const int nwaves = blockDim.x / kWave;
  T wi = kMax ? WaveInclusiveMax(v) : WaveInclusiveSum(v);
  T we = __shfl_up(wi, 1, kWave);
  if (lane == 0) we = T(0);
  if (lane == kWave - 1) lds[wave] = wi;
  __syncthreads();
  if (wave == 0) {
    T w = (lane < nwaves) ? lds[lane] : T(0);
    w = kMax ? WaveInclusiveMax(w) : WaveInclusiveSum(w);
    if (lane < nwaves) lds[lane] = w;
  }
  __syncthreads();
  const T prefix = (wave > 0) ? lds[wave - 1] : T(0);
  if (kMax) {
    *inclusive = (prefix > wi) ? prefix : wi;
    *exclusive = (prefix > we) ? prefix : we;
  } else {
    *inclusive = prefix + wi;
    *exclusive = prefix + we;
  }
  if (total) *total = lds[nwaves - 1];
  __syncthreads();
}

}  // namespace gz
