// mlp.hip -- see mlp.h.
//
// Every contraction runs through one bf16 MFMA GEMM kernel
// (v_mfma_f32_32x32x8bf16_1k, fp32 accumulation): a 256-thread block owns a
// 64 x 64 tile of C = A * B^T, each wave a 32 x 32 sub-tile; A and B are
// addressed through (row, k) strides so the transposed products of the
// backward pass reuse the same kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "cost_model.h"
#include "cost_model_gpu.h"
#include "gz_runtime.h"
#include "gz_common.h"
#include "mlp.h"

namespace gz {

namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float16_t __attribute__((ext_vector_type(16)));

constexpr size_t kOffEmbed = 0;
constexpr size_t kOffW1 = kOffEmbed + (size_t)kCostAlphabet * kCostEmbed;
constexpr size_t kOffB1 = kOffW1 + (size_t)kCostH1 * kCostIn;
constexpr size_t kOffW2 = kOffB1 + kCostH1;
constexpr size_t kOffB2 = kOffW2 + (size_t)kCostH2 * kCostH1;
constexpr size_t kOffW3 = kOffB2 + kCostH2;
constexpr size_t kOffB3 = kOffW3 + kCostH2;
constexpr size_t kFlat = kOffB3 + 1;

__device__ __forceinline__ uint16_t F2Bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float Bf2F(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float Gelu(float z) {
  return 0.5f * z * (1.f + erff(z * 0.70710678118654752f));
}
__device__ __forceinline__ float GeluGrad(float z) {
  const float cdf = 0.5f * (1.f + erff(z * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
  return cdf + z * pdf;
}

enum { kEpiStore = 0, kEpiBiasGelu = 1, kEpiGeluGrad = 2 };

// C[m][n] = sum_k A(m, k) * B(n, k), A(m,k) = A[m*sam + k*sak], B likewise;
// every operand has one unit stride (k for the forward and dX products, the
// row for the weight gradients).  Blocks of four waves own 64 x 64 tiles of
// C (32 x 32 per wave, v_mfma_f32_32x32x16_bf16); K advances 64 at a time
// through LDS, where both operand tiles are stored k-contiguous, each filled
// with 16-byte loads along the operand's unit stride; the next K slice is
// read into registers while the current one is multiplied.  The weight
// gradients have K = batch and few tiles: they split K over gridDim.z and
// add their partial tiles into C (which the caller has zeroed).
constexpr int kGemmTile = 64, kGemmK = 64, kGemmPad = 8;
typedef __bf16 gemm_bf16x8 __attribute__((ext_vector_type(8)));

struct GemmSlice {
  uint4 v[2];  // this thread's 16 values of one operand's 64 x 64 slice
};

// Thread t's part of the slice of rows [r0, r0 + 64) x k [k0, k0 + 64):
// with unit k stride, row r0 + t / 4 and k's (t % 4) * 16 .. + 16; with unit
// row stride, k0 + t / 4 and rows (t % 4) * 16 .. + 16.  Out-of-range values
// read 0.
template <bool kUnitK>
__device__ __forceinline__ void GemmLoadSlice(GemmSlice &g, const uint16_t *X, long sr, long sk,
                                              int rows, int K, int r0, int k0) {
  const int t = threadIdx.x;
  const int a = t >> 2, b = (t & 3) * 16;
  const int gr = kUnitK ? r0 + a : r0 + b, gk = kUnitK ? k0 + b : k0 + a;
  const uint16_t *p = X + (long)gr * sr + (long)gk * sk;
  const bool fast = kUnitK ? (gr < rows && gk + 16 <= K) : (gk < K && gr + 16 <= rows);
  if (fast && ((reinterpret_cast<uintptr_t>(p) & 15) == 0)) {
    g.v[0] = *reinterpret_cast<const uint4 *>(p);
    g.v[1] = *reinterpret_cast<const uint4 *>(p + 8);
  } else {
    uint16_t h[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int rr = kUnitK ? gr : gr + j, kk = kUnitK ? gk + j : gk;
      h[j] = (rr < rows && kk < K) ? X[(long)rr * sr + (long)kk * sk] : (uint16_t)0;
    }
    memcpy(&g.v[0], h, 32);
  }
}

template <bool kUnitK>
__device__ __forceinline__ void GemmStoreSlice(uint16_t (*T)[kGemmK + kGemmPad],
                                               const GemmSlice &g) {
  const int t = threadIdx.x;
  const int a = t >> 2, b = (t & 3) * 16;
  if (kUnitK) {
    *reinterpret_cast<uint4 *>(&T[a][b]) = g.v[0];
    *reinterpret_cast<uint4 *>(&T[a][b + 8]) = g.v[1];
  } else {
    uint16_t h[16];
    memcpy(h, &g.v[0], 32);
#pragma unroll
    for (int j = 0; j < 16; ++j) T[b + j][a] = h[j];
  }
}

template <int EPI, bool kUnitKA, bool kUnitKB>
__global__ void __launch_bounds__(256)
    GemmKernel(int M, int N, int K, const uint16_t *A, long sam, long sak,
               const uint16_t *B, long sbn, long sbk, float *C, int ldc,
               const float *bias, uint16_t *Ch, const float *zsaved) {
  __shared__ __align__(16) uint16_t As[kGemmTile][kGemmK + kGemmPad];
  __shared__ __align__(16) uint16_t Bs[kGemmTile][kGemmK + kGemmPad];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int mb = blockIdx.y * kGemmTile, nb = blockIdx.x * kGemmTile;
  const int m0 = mb + (wave >> 1) * 32, n0 = nb + (wave & 1) * 32;
  const int am = (wave >> 1) * 32 + (lane & 31), bn = (wave & 1) * 32 + (lane & 31);
  const int kq = (lane >> 5) * 8;
  // this block's K range (split K: gridDim.z slices of whole 64-steps)
  const int steps = (K + kGemmK - 1) / kGemmK;
  const int per = (steps + (int)gridDim.z - 1) / (int)gridDim.z;
  const int kb = (int)blockIdx.z * per * kGemmK, ke = min(K, kb + per * kGemmK);
  float16_t acc;
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  GemmSlice ga, gb;
  if (kb < ke) {
    GemmLoadSlice<kUnitKA>(ga, A, sam, sak, M, ke, mb, kb);
    GemmLoadSlice<kUnitKB>(gb, B, sbn, sbk, N, ke, nb, kb);
  }
  for (int k0 = kb; k0 < ke; k0 += kGemmK) {
    GemmStoreSlice<kUnitKA>(As, ga);
    GemmStoreSlice<kUnitKB>(Bs, gb);
    __syncthreads();
    if (k0 + kGemmK < ke) {  // the next slice's loads overlap this slice's MFMAs
      GemmLoadSlice<kUnitKA>(ga, A, sam, sak, M, ke, mb, k0 + kGemmK);
      GemmLoadSlice<kUnitKB>(gb, B, sbn, sbk, N, ke, nb, k0 + kGemmK);
    }
#pragma unroll
    for (int ks = 0; ks < kGemmK; ks += 16) {
      const gemm_bf16x8 a = *reinterpret_cast<const gemm_bf16x8 *>(&As[am][ks + kq]);
      const gemm_bf16x8 b = *reinterpret_cast<const gemm_bf16x8 *>(&Bs[bn][ks + kq]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int row = m0 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    const int col = n0 + (lane & 31);
    if (row < M && col < N) {
      const long idx = (long)row * ldc + col;
      const float v = acc[r];
      if (EPI == kEpiStore) {
        if (gridDim.z > 1) atomicAdd(&C[idx], v);
        else C[idx] = v;
      } else if (EPI == kEpiBiasGelu) {
        const float z = v + bias[col];
        C[idx] = z;
        Ch[idx] = F2Bf(Gelu(z));
      } else {
        const float dz = v * GeluGrad(zsaved[idx]);
        C[idx] = dz;
        Ch[idx] = F2Bf(dz);
      }
    }
  }
}

__global__ void EmbedKernel(int B, const uint8_t *codes, const uint16_t *embh, uint16_t *x) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)B * kCostIn) return;
  const long b = i / kCostIn;
  const int c = (int)(i % kCostIn);
  const int r = c / kCostEmbed, e = c % kCostEmbed;
  const int a = min((int)codes[b * kCostResidues + r], kCostAlphabet - 1);
  x[i] = embh[a * kCostEmbed + e];
}

__global__ void OutputKernel(int B, const uint16_t *h2, const float *w3, const float *b3,
                             float *y) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float s = 0.f;
  for (int k = 0; k < kCostH2; ++k) s += Bf2F(h2[b * kCostH2 + k]) * Bf2F(F2Bf(w3[k]));
  y[b] = s + b3[0];
}

// Batch rows from a device-resident training set: row r < n of step *done
// is row order_base[*done * n + r] of (codes, targets); padding rows read
// code 20 and target 0.  One thread per 16 bytes of the batch's codes.
__global__ void GatherRowsKernel(const uint8_t *codes, const float *targets,
                                 const uint32_t *order_base, const uint32_t *done, int n, int B,
                                 uint8_t *out_codes, float *out_t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  constexpr int kPieces = kCostResidues / 16;
  if (i >= B * kPieces) return;
  // the batch of step *done
  const uint32_t *order = order_base + (size_t)*done * (size_t)n;
  const int r = i / kPieces, k = i % kPieces;
  uint4 v = make_uint4(0x14141414u, 0x14141414u, 0x14141414u, 0x14141414u);
  if (r < n) v = reinterpret_cast<const uint4 *>(codes + (size_t)order[r] * kCostResidues)[k];
  reinterpret_cast<uint4 *>(out_codes + (size_t)r * kCostResidues)[k] = v;
  if (k == 0) out_t[r] = r < n ? targets[order[r]] : 0.f;
}

// End of a step: its loss into the sum, the step count up by one.
__global__ void AdvanceKernel(const float *loss, float *sum, uint32_t *done) {
  if (threadIdx.x == 0) {
    *sum += *loss;
    *done += 1;
  }
}

// MSE loss over the first n rows; dY = 2 (y - t) / n, zero on padding.
__global__ void LossKernel(int B, int n, const float *y, const float *t, float *dy,
                           float *loss) {
  __shared__ float part[256];
  float acc = 0.f;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    float g = 0.f;
    if (b < n) {
      const float d = y[b] - t[b];
      acc += d * d;
      g = 2.f * d / n;
    }
    dy[b] = g;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = part[0] / n;
}

// dZ2 = dY w3 * gelu'(Z2), elementwise over B x 64
__global__ void OutputBackwardKernel(int B, const float *dy, const float *z2, const float *w3,
                                     float *dz2, uint16_t *dz2h) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * kCostH2) return;
  const int b = i / kCostH2, k = i % kCostH2;
  const float d = dy[b] * Bf2F(F2Bf(w3[k])) * GeluGrad(z2[i]);
  dz2[i] = d;
  dz2h[i] = F2Bf(d);
}

// out[c] += sum_b wt[b] * M[b][c] (wt optional); M is fp32 or bf16.  Blocks
// own 64 columns and a slice of rows; partial sums are added atomically.
template <typename T>
__global__ void __launch_bounds__(256)
    ColumnSumKernel(int B, int N, const T *m, const float *wt, float *out) {
  __shared__ float part[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int lane_row = threadIdx.x >> 6;
  const int rows = (B + gridDim.y - 1) / gridDim.y;
  const int r0 = blockIdx.y * rows, r1 = min(B, r0 + rows);
  float s = 0.f;
  if (c < N)
    for (int b = r0 + lane_row; b < r1; b += 4) {
      float v;
      if constexpr (sizeof(T) == 2) v = Bf2F((uint16_t)m[(long)b * N + c]);
      else v = (float)m[(long)b * N + c];
      s += wt ? wt[b] * v : v;
    }
  part[lane_row][threadIdx.x & 63] = s;
  __syncthreads();
  if (threadIdx.x < 64 && c < N)
    atomicAdd(&out[c], part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                           part[3][threadIdx.x]);
}

// d_embed[a][e] = sum over (b, r) with code a of dX[b][r*8 + e]
__global__ void EmbedBackwardKernel(int B, const uint8_t *codes, const float *dx, float *gemb) {
  __shared__ float acc[kCostAlphabet * kCostEmbed];
  for (int i = threadIdx.x; i < kCostAlphabet * kCostEmbed; i += blockDim.x) acc[i] = 0.f;
  __syncthreads();
  const long total = (long)B * kCostIn;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    const long b = i / kCostIn;
    const int c = (int)(i % kCostIn);
    const int a = min((int)codes[b * kCostResidues + c / kCostEmbed], kCostAlphabet - 1);
    atomicAdd(&acc[a * kCostEmbed + c % kCostEmbed], dx[i]);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kCostAlphabet * kCostEmbed; i += blockDim.x)
    atomicAdd(&gemb[i], acc[i]);
}

__global__ void AdamKernel(size_t n, float *w, const float *g, float *m, float *v, float lr,
                           float b1, float b2, float c1, float c2, float eps) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// Adam with the step number on the device: t = base + *done + 1 (a step
// sequence replayed from a graph keeps its own count).
__global__ void AdamCountedKernel(size_t n, float *w, const float *g, float *m, float *v,
                                  float lr, float b1, float b2, float eps, long base,
                                  const uint32_t *done) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float t = (float)(base + (long)*done + 1);
  const float c1 = 1.f - powf(b1, t), c2 = 1.f - powf(b2, t);
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void ZeroFloatKernel(size_t n, float *p) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

__global__ void ToBf16Kernel(size_t n, const float *src, uint16_t *dst) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = F2Bf(src[i]);
}

inline unsigned Blocks(size_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

// ---------------------------------------------------------------------------
// Inference: one persistent kernel on v_mfma_f32_32x32x16_bf16 runs all three
// layers per tile of seeds, so no activation leaves the CU.
//
// Operand layout of the 32x32x16 MFMA: lane l supplies A[row l % 32][k] and
// B[k][column l % 32] for k = 8 * (l / 32) + j, j < 8; accumulator register r
// of lane l holds C[8 * (r / 4) + 4 * (l / 32) + (r % 4)][l % 32].
//
// Layers are computed transposed (hidden unit = row, seed = column); each
// wave owns 32 seeds:
//   layer 1: C1[h][seed] = W1[h][:] . x[seed][:] in kHB row blocks of 32
//            units; A = W1 from LDS ([k/8][unit][8]), a 32nd of the inputs
//            (16 KB) per stage in two alternating buffers, the next
//            stage read into registers during this one's MFMAs and stored
//            into the other buffer; B = the residue embeddings (one 16-byte
//            LDS read);
//   layer 2: C2[n][seed] = W2[n][:] . h1[:][seed], whose B operand is the
//            lane's own layer-1 accumulators (a lane holds one seed's column):
//            K step t takes registers 8 (t % 2) .. 8 (t % 2) + 7 of row block
//            t / 2, i.e. hidden units 16 t + 8 (j / 4) + 4 (l / 32) + j % 4,
//            and W2's columns are staged in LDS in that order;
//   layer 3: each lane sums its 32 rows of C2 and adds its partner lane's.
// Rounding points are those of the training forward (bf16 h1, h2, w3).
// ---------------------------------------------------------------------------
static_assert(kCostH1 % 32 == 0 && kCostH2 == 64, "inference tiles: H1 in 32s, H2 = 64");
constexpr int kHB = kCostH1 / 32;                  // layer-1 row blocks
// Blocks of four waves and 66 KB of LDS, two per CU: their stage barriers
// are independent, so one block's GELU epilogue overlaps the other's MFMAs.
constexpr int kInferWaves = 4;
constexpr int kInferTile = kInferWaves * 32;       // seeds per block pass
constexpr int kStageIn = 32;                       // W1 inputs per stage (4 residues)
constexpr int kStages = kCostIn / kStageIn;        // 32
static_assert(kStageIn == 32, "a stage's residues are one 32-bit word of codes");
constexpr int kStageVec = kCostH1 * kStageIn / 8;  // 16-byte pieces per stage
constexpr int kPrefetch = kStageVec / (kInferWaves * 64);
static_assert(kStageVec % (kInferWaves * 64) == 0, "a stage splits evenly over the block");
static_assert(kStages % 2 == 0, "stages alternate between two buffers across tiles");
static_assert((kStages & (kStages - 1)) == 0 && kStages * 4 == kCostResidues,
              "a row is kStages words; rotations wrap with a mask");
constexpr int kL2Steps = kCostH1 / 16;             // layer-2 K steps

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ bf16x8 Load8(const uint16_t *p) {
  return *reinterpret_cast<const bf16x8 *>(p);
}

__device__ __forceinline__ __bf16 AsBf16(uint16_t h) {
  return __builtin_bit_cast(__bf16, h);
}

// Piece i of stage s: unit h = i / 4, inputs s * 32 + (i % 4) * 8 .. + 8.
__device__ __forceinline__ uint4 StagePiece(const uint16_t *w1h, int s, int i) {
  const int h = i / (kStageIn / 8), k8 = i % (kStageIn / 8);
  return *reinterpret_cast<const uint4 *>(w1h + (size_t)h * kCostIn + s * kStageIn + k8 * 8);
}
__device__ __forceinline__ void PutPiece(uint16_t *w1s, int i, uint4 v) {
  const int h = i / (kStageIn / 8), k8 = i % (kStageIn / 8);
  *reinterpret_cast<uint4 *>(&w1s[(k8 * kCostH1 + h) * 8]) = v;
}

__global__ void __launch_bounds__(kInferWaves * 64) __attribute__((amdgpu_waves_per_eu(2)))
    InferKernel(const uint8_t *codes, uint32_t n, const uint32_t *count, uint32_t count_offset,
                const uint16_t *w1h, const uint16_t *embh, const float *b1, const uint16_t *w2h,
                const float *b2, const float *w3, const float *b3, uint32_t rotate, float *y) {
  // with `count`, the rows are the first *count - count_offset (at most n)
  if (count) {
    const uint32_t c = *count;
    n = c > count_offset ? min(n, c - count_offset) : 0u;
  }
  __shared__ __align__(16) uint16_t w1s[2][kStageIn * kCostH1];     // two stages, 32 KB
  __shared__ __align__(16) uint16_t w2[kL2Steps * 2 * 64 * 8];      // [t][row block][lane][8]
  __shared__ __align__(16) uint16_t emb[kCostAlphabet * kCostEmbed];
  __shared__ float b1s[kCostH1], b2s[kCostH2], w3s[kCostH2];
  for (int idx = threadIdx.x; idx < kL2Steps * 2 * 64 * 8; idx += blockDim.x) {
    const int j = idx & 7, l = (idx >> 3) & 63, nb = (idx >> 9) & 1, t = idx >> 10;
    const int row = 32 * nb + (l & 31);
    const int h = 16 * t + 8 * (j >> 2) + 4 * (l >> 5) + (j & 3);
    w2[idx] = w2h[row * kCostH1 + h];
  }
  for (int i = threadIdx.x; i < kCostAlphabet * kCostEmbed; i += blockDim.x) emb[i] = embh[i];
  for (int i = threadIdx.x; i < kCostH1; i += blockDim.x) b1s[i] = b1[i];
  if (threadIdx.x < kCostH2) {
    b2s[threadIdx.x] = b2[threadIdx.x];
    w3s[threadIdx.x] = Bf2F(F2Bf(w3[threadIdx.x]));
  }
  // stage 0 of W1; every later stage is read one stage ahead
#pragma unroll
  for (int p = 0; p < kPrefetch; ++p) {
    const int i = (int)threadIdx.x + p * kInferWaves * 64;
    PutPiece(w1s[0], i, StagePiece(w1h, 0, i));
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int kq = lane >> 5, r32 = lane & 31;
  const uint32_t kshift = 8u * (uint32_t)kq;
  const float bias3 = b3[0];
  const uint32_t tiles = (n + kInferTile - 1) / kInferTile;
  for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
    // the codes of this lane's seed stream through registers 4 at a time;
    // seeds past the end read seed 0's row as code 20 throughout
    const uint32_t seed = tile * kInferTile + wave * 32 + r32;
    const uint4 *src =
        reinterpret_cast<const uint4 *>(codes + (size_t)(seed < n ? seed : 0) * kCostResidues);
    const uint32_t fill = seed < n ? 0u : 0x14141414u;
    float16_t acc[kHB];
#pragma unroll
    for (int hb = 0; hb < kHB; ++hb)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[hb][r] = 0.f;
#pragma unroll 1
    for (int s = 0; s < kStages; ++s) {
      // stage s is in buffer s & 1; the next one (the first again after the
      // last; kStages is even) goes to the other buffer, which every wave
      // finished reading before the barrier that ended stage s - 1
      uint4 pf[kPrefetch];
      const int ns = (s + 1) % kStages;
#pragma unroll
      for (int p = 0; p < kPrefetch; ++p)
        pf[p] = StagePiece(w1h, ns, (int)threadIdx.x + p * kInferWaves * 64);
      const uint16_t *cur = w1s[s & 1];
      // residues 4 s .. 4 s + 3 of this lane's seed, the row read from word
      // `rotate` on (16 swaps its two halves of windows)
      const uint32_t cw =
          reinterpret_cast<const uint32_t *>(src)[(s + (int)rotate) & (kStages - 1)] | fill;
      // a K step's operands are read as one batch, a step ahead of its
      // MFMAs (the scheduling barriers keep that order), so the LDS latency
      // overlaps the previous step's MFMAs
      bf16x8 a[2][kHB], bx[2];
      auto load_step = [&](int q, bf16x8 (&av)[kHB], bf16x8 &b) {
        // residue 4 s + 2 q + kq: byte 2 q + kq of the word
        const int k8 = 2 * q + kq;
        uint32_t code = (cw >> (16 * q + kshift)) & 0xffu;
        code = code < 20 ? code : 20;
        b = Load8(&emb[code * kCostEmbed]);
#pragma unroll
        for (int hb = 0; hb < kHB; ++hb) av[hb] = Load8(&cur[(k8 * kCostH1 + 32 * hb + r32) * 8]);
      };
      load_step(0, a[0], bx[0]);
#pragma unroll
      for (int q = 0; q < kStageIn / 16; ++q) {
        if (q + 1 < kStageIn / 16) load_step(q + 1, a[(q + 1) & 1], bx[(q + 1) & 1]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int hb = 0; hb < kHB; ++hb)
          acc[hb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[q & 1][hb], bx[q & 1], acc[hb],
                                                            0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int p = 0; p < kPrefetch; ++p)
        PutPiece(w1s[ns & 1], (int)threadIdx.x + p * kInferWaves * 64, pf[p]);
      __syncthreads();
    }
    float16_t acc2[2];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      acc2[0][r] = 0.f;
      acc2[1][r] = 0.f;
    }
#pragma unroll
    for (int t = 0; t < kL2Steps; ++t) {
      bf16x8 bh;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int h = 16 * t + 8 * (j >> 2) + 4 * kq + (j & 3);
        bh[j] = AsBf16(F2Bf(Gelu(acc[t >> 1][8 * (t & 1) + j] + b1s[h])));
      }
#pragma unroll
      for (int nb = 0; nb < 2; ++nb) {
        const bf16x8 a = Load8(&w2[((t * 2 + nb) * 64 + lane) * 8]);
        acc2[nb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, bh, acc2[nb], 0, 0, 0);
      }
    }
    float part = 0.f;
#pragma unroll
    for (int nb = 0; nb < 2; ++nb)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = 32 * nb + 8 * (r >> 2) + 4 * kq + (r & 3);
        part += Bf2F(F2Bf(Gelu(acc2[nb][r] + b2s[row]))) * w3s[row];
      }
    part += __shfl_xor(part, 32, 64);
    if (kq == 0 && seed < n) y[seed] = part + bias3;
  }
}

// K splits for a weight gradient of `tiles` output tiles and K = batch:
// about 512 blocks, each with at least four 64-steps of K
inline int SplitK(int K, int tiles) {
  const int steps = (K + kGemmK - 1) / kGemmK;
  return std::max(1, std::min(512 / tiles, steps / 4));
}

// split_k > 1 only for kEpiStore into a zeroed C
template <int EPI>
void Gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, long sam, long sak,
          const uint16_t *B, long sbn, long sbk, float *C, int ldc, const float *bias = nullptr,
          uint16_t *Ch = nullptr, const float *z = nullptr, int split_k = 1) {
  dim3 grid((N + 63) / 64, (M + 63) / 64, split_k);
  // every operand here has one unit stride: k or the row
  if (sak == 1 && sbk == 1)
    hipLaunchKernelGGL((GemmKernel<EPI, true, true>), grid, dim3(256), 0, s, M, N, K, A, sam,
                       sak, B, sbn, sbk, C, ldc, bias, Ch, z);
  else if (sak == 1)
    hipLaunchKernelGGL((GemmKernel<EPI, true, false>), grid, dim3(256), 0, s, M, N, K, A, sam,
                       sak, B, sbn, sbk, C, ldc, bias, Ch, z);
  else if (sbk == 1)
    hipLaunchKernelGGL((GemmKernel<EPI, false, true>), grid, dim3(256), 0, s, M, N, K, A, sam,
                       sak, B, sbn, sbk, C, ldc, bias, Ch, z);
  else
    hipLaunchKernelGGL((GemmKernel<EPI, false, false>), grid, dim3(256), 0, s, M, N, K, A, sam,
                       sak, B, sbn, sbk, C, ldc, bias, Ch, z);
}

}  // namespace

MlpDevice::MlpDevice() {
  count_ = kFlat;
  hipError_t e = hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking);
  if (e == hipSuccess) e = gz::Malloc((void **)&master_, kFlat * 4);
  if (e == hipSuccess) e = gz::Malloc((void **)&grad_, kFlat * 4);
  if (e == hipSuccess) e = gz::Malloc((void **)&m_, kFlat * 4);
  if (e == hipSuccess) e = gz::Malloc((void **)&v_, kFlat * 4);
  if (e == hipSuccess) e = hipMemset(m_, 0, kFlat * 4);
  if (e == hipSuccess) e = hipMemset(v_, 0, kFlat * 4);
  if (e == hipSuccess) e = gz::Malloc((void **)&w1h_, (size_t)kCostH1 * kCostIn * 2);
  if (e == hipSuccess) e = gz::Malloc((void **)&w2h_, (size_t)kCostH2 * kCostH1 * 2);
  if (e == hipSuccess) e = gz::Malloc((void **)&embh_, (size_t)kCostAlphabet * kCostEmbed * 2);
  init_status_ = e;
}

MlpDevice::~MlpDevice() {
  void *ps[] = {master_, grad_, m_,   v_,    w1h_,  w2h_,  embh_, codes_, x_,
                z1_,     h1_,   z2_,  h2_,   y_,    t_,    loss_, dz2_,   dz2h_,
                dz1_,    dz1h_, dh1_, dx_};
  for (void *p : ps)
    if (p) (void)gz::Free(p);
  if (stream_) (void)hipStreamDestroy(stream_);
}

void MlpDevice::SyncBf16() {
  hipLaunchKernelGGL(ToBf16Kernel, dim3(Blocks((size_t)kCostH1 * kCostIn)), dim3(256), 0,
                     stream_, (size_t)kCostH1 * kCostIn, master_ + kOffW1, w1h_);
  hipLaunchKernelGGL(ToBf16Kernel, dim3(Blocks((size_t)kCostH2 * kCostH1)), dim3(256), 0,
                     stream_, (size_t)kCostH2 * kCostH1, master_ + kOffW2, w2h_);
  hipLaunchKernelGGL(ToBf16Kernel, dim3(1), dim3(256), 0, stream_,
                     (size_t)kCostAlphabet * kCostEmbed, master_ + kOffEmbed, embh_);
}

hipError_t MlpDevice::SetWeights(const CostModelWeights &w) {
  if (init_status_ != hipSuccess) return init_status_;
  std::vector<float> flat(kFlat);
  float *f = flat.data();
  auto put = [&](const std::vector<float> &v, size_t off) { memcpy(f + off, v.data(), v.size() * 4); };
  put(w.embed, kOffEmbed);
  put(w.w1, kOffW1);
  put(w.b1, kOffB1);
  put(w.w2, kOffW2);
  put(w.b2, kOffB2);
  put(w.w3, kOffW3);
  put(w.b3, kOffB3);
  hipError_t e = hipMemcpyAsync(master_, flat.data(), kFlat * 4, hipMemcpyHostToDevice, stream_);
  if (e == hipSuccess) e = hipMemsetAsync(m_, 0, kFlat * 4, stream_);
  if (e == hipSuccess) e = hipMemsetAsync(v_, 0, kFlat * 4, stream_);
  step_ = 0;
  if (e == hipSuccess) {
    SyncBf16();
    e = hipGetLastError();
  }
  const hipError_t e2 = hipStreamSynchronize(stream_);
  return e != hipSuccess ? e : e2;
}

hipError_t MlpDevice::GetWeights(CostModelWeights *w) {
  if (init_status_ != hipSuccess) return init_status_;
  std::vector<float> flat(kFlat);
  const hipError_t e = hipMemcpy(flat.data(), master_, kFlat * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  auto get = [&](std::vector<float> &v, size_t off, size_t n) {
    v.assign(flat.begin() + off, flat.begin() + off + n);
  };
  get(w->embed, kOffEmbed, (size_t)kCostAlphabet * kCostEmbed);
  get(w->w1, kOffW1, (size_t)kCostH1 * kCostIn);
  get(w->b1, kOffB1, kCostH1);
  get(w->w2, kOffW2, (size_t)kCostH2 * kCostH1);
  get(w->b2, kOffB2, kCostH2);
  get(w->w3, kOffW3, kCostH2);
  get(w->b3, kOffB3, 1);
  return hipSuccess;
}

hipError_t MlpDevice::Reserve(size_t batch) {
  if (init_status_ != hipSuccess) return init_status_;
  if (batch <= cap_) return hipSuccess;
  void **ps[] = {(void **)&codes_, (void **)&x_,   (void **)&z1_,  (void **)&h1_,
                 (void **)&z2_,    (void **)&h2_,  (void **)&y_,   (void **)&t_,
                 (void **)&loss_,  (void **)&dz2_, (void **)&dz2h_, (void **)&dz1_,
                 (void **)&dz1h_,  (void **)&dh1_, (void **)&dx_};
  for (void **p : ps)
    if (*p) {
      (void)gz::Free(*p);
      *p = nullptr;
    }
  const size_t B = batch;
  hipError_t e = hipSuccess;
  auto al = [&](void **p, size_t bytes) {
    if (e == hipSuccess) e = gz::Malloc(p, std::max<size_t>(bytes, 64));
  };
  al((void **)&codes_, B * kCostResidues);
  al((void **)&x_, B * kCostIn * 2);
  al((void **)&z1_, B * kCostH1 * 4);
  al((void **)&h1_, B * kCostH1 * 2);
  al((void **)&z2_, B * kCostH2 * 4);
  al((void **)&h2_, B * kCostH2 * 2);
  al((void **)&y_, B * 4);
  al((void **)&t_, B * 4);
  al((void **)&loss_, 64);
  al((void **)&dz2_, B * kCostH2 * 4);
  al((void **)&dz2h_, B * kCostH2 * 2);
  al((void **)&dz1_, B * kCostH1 * 4);
  al((void **)&dz1h_, B * kCostH1 * 2);
  al((void **)&dh1_, B * kCostH1 * 4);
  al((void **)&dx_, B * kCostIn * 4);
  if (e == hipSuccess) cap_ = batch;
  return e;
}

hipError_t MlpDevice::ForwardDevice(size_t n, size_t B) {
  (void)n;
  hipStream_t s = stream_;
  hipLaunchKernelGGL(EmbedKernel, dim3(Blocks(B * kCostIn)), dim3(256), 0, s, (int)B, codes_,
                     embh_, x_);
  Gemm<kEpiBiasGelu>(s, (int)B, kCostH1, kCostIn, x_, kCostIn, 1, w1h_, kCostIn, 1, z1_, kCostH1,
                     master_ + kOffB1, h1_);
  Gemm<kEpiBiasGelu>(s, (int)B, kCostH2, kCostH1, h1_, kCostH1, 1, w2h_, kCostH1, 1, z2_,
                     kCostH2, master_ + kOffB2, h2_);
  hipLaunchKernelGGL(OutputKernel, dim3(Blocks(B)), dim3(256), 0, s, (int)B, h2_,
                     master_ + kOffW3, master_ + kOffB3, y_);
  return hipGetLastError();
}

hipError_t MlpDevice::InferDevice(const uint8_t *d_codes, size_t n, float *d_out,
                                  hipStream_t s, uint32_t rotate_words, const uint32_t *d_count,
                                  uint32_t count_offset) {
  if (init_status_ != hipSuccess) return init_status_;
  if (n == 0) return hipSuccess;
  if (n > 0xFFFFFFFFull - kInferTile) return hipErrorInvalidValue;
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint32_t tiles = (uint32_t)((n + kInferTile - 1) / kInferTile);
  hipLaunchKernelGGL(InferKernel, dim3(std::min<unsigned>(tiles, 2u * (unsigned)cus)),
                     dim3(kInferWaves * 64), 0, s, d_codes, (uint32_t)n, d_count, count_offset,
                     w1h_, embh_, master_ + kOffB1, w2h_, master_ + kOffB2, master_ + kOffW3,
                     master_ + kOffB3, rotate_words & (kStages - 1), d_out);
  return hipGetLastError();
}

hipError_t MlpDevice::Forward(const uint8_t *codes, size_t n, float *out, uint32_t rotate_words) {
  last_forward_ms = 0;
  if (n == 0) return hipSuccess;
  uint8_t *d_codes = nullptr;
  float *d_out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = gz::Malloc((void **)&d_codes, n * kCostResidues);
  if (e == hipSuccess) e = gz::Malloc((void **)&d_out, n * 4);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_codes, codes, n * kCostResidues, hipMemcpyHostToDevice, stream_);
  if (e == hipSuccess) e = hipEventRecord(e0, stream_);
  if (e == hipSuccess) e = InferDevice(d_codes, n, d_out, stream_, rotate_words);
  if (e == hipSuccess) e = hipEventRecord(e1, stream_);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, n * 4, hipMemcpyDeviceToHost, stream_);
  if (e == hipSuccess) e = hipStreamSynchronize(stream_);
  float ms = 0;
  if (e == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) last_forward_ms = ms;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (d_codes) (void)gz::Free(d_codes);
  if (d_out) (void)gz::Free(d_out);
  return e;
}

hipError_t MlpDevice::TrainStep(const uint8_t *codes, const float *targets, size_t n,
                                float lr, float *loss, float *flat_grad) {
  const size_t B = (n + 63) / 64 * 64;
  hipError_t e = Reserve(B);
  if (e != hipSuccess) return e;
  hipStream_t s = stream_;
  (void)hipMemsetAsync(codes_, 20, B * kCostResidues, s);
  (void)hipMemcpyAsync(codes_, codes, n * kCostResidues, hipMemcpyHostToDevice, s);
  (void)hipMemsetAsync(t_, 0, B * 4, s);
  (void)hipMemcpyAsync(t_, targets, n * 4, hipMemcpyHostToDevice, s);
  e = StepOnDevice(n, B, lr, 0, nullptr);
  if (e != hipSuccess) return e;
  if (loss) (void)hipMemcpyAsync(loss, loss_, 4, hipMemcpyDeviceToHost, s);
  if (flat_grad) (void)hipMemcpyAsync(flat_grad, grad_, kFlat * 4, hipMemcpyDeviceToHost, s);
  return hipStreamSynchronize(s);
}

hipError_t MlpDevice::TrainEpochs(const uint8_t *codes, const float *targets, size_t n,
                                  const uint32_t *order, size_t m, size_t batch, float lr,
                                  float *mean_loss, float *kernel_ms) {
  if (mean_loss) *mean_loss = 0.f;
  if (kernel_ms) *kernel_ms = 0.f;
  if (n == 0 || m == 0 || batch == 0) return hipSuccess;
  const size_t B = (batch + 63) / 64 * 64;
  hipError_t e = Reserve(B);
  if (e != hipSuccess) return e;
  hipStream_t s = stream_;
  uint8_t *d_codes = nullptr;
  float *d_t = nullptr, *d_sum = nullptr;
  uint32_t *d_order = nullptr, *d_done = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  e = gz::Malloc((void **)&d_codes, n * kCostResidues);
  if (e == hipSuccess) e = gz::Malloc((void **)&d_t, n * 4);
  if (e == hipSuccess) e = gz::Malloc((void **)&d_order, m * 4);
  if (e == hipSuccess) e = gz::Malloc((void **)&d_sum, 256);
  d_done = reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(d_sum) + 128);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_codes, codes, n * kCostResidues, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_t, targets, n * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_order, order, m * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemsetAsync(d_sum, 0, 256, s);
  if (e == hipSuccess) e = hipEventRecord(e0, s);
  // whole batches only (the remainder of the order is left out, as the
  // per-step loop of the callers does); every step reads its batch and its
  // Adam step number from d_done, so a captured run of kGraphSteps steps is
  // replayed as is
  const size_t steps = m / batch;
  const long base = step_;
  auto step = [&]() -> hipError_t {
    hipLaunchKernelGGL(GatherRowsKernel, dim3((unsigned)Blocks(B * kCostResidues / 16)),
                       dim3(256), 0, s, d_codes, d_t, d_order, d_done, (int)batch, (int)B,
                       codes_, t_);
    hipError_t r = StepOnDevice(batch, B, lr, base, d_done);
    if (r == hipSuccess) {
      hipLaunchKernelGGL(AdvanceKernel, dim3(1), dim3(64), 0, s, loss_, d_sum, d_done);
      r = hipGetLastError();
    }
    return r;
  };
  constexpr size_t kGraphSteps = 16;
  size_t done = 0;
  if (e == hipSuccess && steps >= 2 * kGraphSteps) {
    {
      std::lock_guard<std::recursive_mutex> lock(CaptureLock());
      e = hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed);
      hipError_t ec = hipSuccess;
      for (size_t k = 0; k < kGraphSteps && e == hipSuccess; ++k) e = step();
      ec = hipStreamEndCapture(s, &graph);
      if (e == hipSuccess) e = ec;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    // each replay waits for the previous one (one stream); the host
    // synchronizes between replays so that a replay never races the next
    for (; e == hipSuccess && done + kGraphSteps <= steps; done += kGraphSteps) {
      e = hipGraphLaunch(exec, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
  }
  for (; e == hipSuccess && done < steps; ++done) e = step();
  if (e == hipSuccess) e = hipEventRecord(e1, s);
  float sum = 0.f;
  if (e == hipSuccess) e = hipMemcpyAsync(&sum, d_sum, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e == hipSuccess) step_ += (long)steps;
  if (e == hipSuccess && mean_loss && steps) *mean_loss = sum / (float)steps;
  float ms = 0.f;
  if (e == hipSuccess && kernel_ms && hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
    *kernel_ms = ms;
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (void *p : {(void *)d_codes, (void *)d_t, (void *)d_order, (void *)d_sum})
    if (p) (void)gz::Free(p);
  return e;
}

// One optimisation step on the batch in codes_ / t_ (n rows, B padded):
// forward, loss, backward, Adam (lr != 0); nothing is read back.  With
// done, Adam's step number is base + *done + 1 on the device (the host's
// step count is the caller's to advance); without, the step is counted here.
hipError_t MlpDevice::StepOnDevice(size_t n, size_t B, float lr, long base,
                                   const uint32_t *done) {
  hipStream_t s = stream_;
  hipError_t e = ForwardDevice(n, B);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(LossKernel, dim3(1), dim3(256), 0, s, (int)B, (int)n, y_, t_, y_, loss_);
  // y_ now holds dY; the gradient is cleared by a kernel (graph-safe)
  hipLaunchKernelGGL(ZeroFloatKernel, dim3(Blocks(kFlat)), dim3(256), 0, s, kFlat, grad_);
  const unsigned row_split = std::max(1u, std::min(64u, (unsigned)(B / 64)));
  // dW3 = H2^T dY, dB3 = sum dY, then dZ2
  hipLaunchKernelGGL((ColumnSumKernel<uint16_t>), dim3(1, row_split), dim3(256), 0, s, (int)B,
                     kCostH2, h2_, y_, grad_ + kOffW3);
  hipLaunchKernelGGL((ColumnSumKernel<float>), dim3(1, row_split), dim3(256), 0, s, (int)B, 1, y_,
                     (const float *)nullptr, grad_ + kOffB3);
  hipLaunchKernelGGL(OutputBackwardKernel, dim3(Blocks(B * kCostH2)), dim3(256), 0, s, (int)B,
                     y_, z2_, master_ + kOffW3, dz2_, dz2h_);
  // dW2 = dZ2^T H1
  Gemm<kEpiStore>(s, kCostH2, kCostH1, (int)B, dz2h_, 1, kCostH2, h1_, 1, kCostH1,
                  grad_ + kOffW2, kCostH1, nullptr, nullptr, nullptr, SplitK((int)B, 4));
  hipLaunchKernelGGL((ColumnSumKernel<float>), dim3(1, row_split), dim3(256), 0, s, (int)B,
                     kCostH2, dz2_, (const float *)nullptr, grad_ + kOffB2);
  // dZ1 = (dZ2 W2) * gelu'(Z1)
  Gemm<kEpiGeluGrad>(s, (int)B, kCostH1, kCostH2, dz2h_, kCostH2, 1, w2h_, 1, kCostH1, dz1_,
                     kCostH1, nullptr, dz1h_, z1_);
  // dW1 = dZ1^T X
  Gemm<kEpiStore>(s, kCostH1, kCostIn, (int)B, dz1h_, 1, kCostH1, x_, 1, kCostIn,
                  grad_ + kOffW1, kCostIn, nullptr, nullptr, nullptr, SplitK((int)B, 64));
  hipLaunchKernelGGL((ColumnSumKernel<float>), dim3(kCostH1 / 64, row_split), dim3(256), 0, s,
                     (int)B, kCostH1, dz1_, (const float *)nullptr, grad_ + kOffB1);
  // dX = dZ1 W1, then the embedding gradient
  Gemm<kEpiStore>(s, (int)B, kCostIn, kCostH1, dz1h_, kCostH1, 1, w1h_, 1, kCostIn, dx_,
                  kCostIn);
  hipLaunchKernelGGL(EmbedBackwardKernel, dim3(std::min<size_t>(Blocks(B * kCostIn), 1024)),
                     dim3(256), 0, s, (int)n, codes_, dx_, grad_ + kOffEmbed);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (lr != 0.f) {
    const float b1 = 0.9f, b2 = 0.999f;
    if (done) {
      hipLaunchKernelGGL(AdamCountedKernel, dim3(Blocks(kFlat)), dim3(256), 0, s, kFlat,
                         master_, grad_, m_, v_, lr, b1, b2, 1e-8f, base, done);
    } else {
      ++step_;
      const float c1 = 1.f - std::pow(b1, (float)step_), c2 = 1.f - std::pow(b2, (float)step_);
      hipLaunchKernelGGL(AdamKernel, dim3(Blocks(kFlat)), dim3(256), 0, s, kFlat, master_,
                         grad_, m_, v_, lr, b1, b2, c1, c2, 1e-8f);
    }
    SyncBf16();
  }
  return hipGetLastError();
}

}  // namespace gz

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace gz;

struct gz_mlp_s {
  MlpDevice dev;
};

extern "C" {

size_t gz_mlp_weight_count(void) { return kFlat; }

gz_status gz_mlp_create(const float *flat, size_t count, gz_mlp_t *out) {
  GZ_REQUIRE(out && flat && count == kFlat, "gz_mlp_create: bad weights");
  CostModelWeights w;
  const float *f = flat;
  auto take = [&](std::vector<float> &v, size_t n) {
    v.assign(f, f + n);
    f += n;
  };
  take(w.embed, (size_t)kCostAlphabet * kCostEmbed);
  take(w.w1, (size_t)kCostH1 * kCostIn);
  take(w.b1, kCostH1);
  take(w.w2, (size_t)kCostH2 * kCostH1);
  take(w.b2, kCostH2);
  take(w.w3, kCostH2);
  take(w.b3, 1);
  std::unique_ptr<gz_mlp_s> h(new gz_mlp_s());
  GZ_HIP_CHECK(h->dev.SetWeights(w));
  *out = h.release();
  return GZ_OK;
}

gz_status gz_mlp_init_weights(uint64_t seed, float *flat) {
  CostModelWeights w;
  w.Init(seed);
  float *f = flat;
  for (auto *v : {&w.embed, &w.w1, &w.b1, &w.w2, &w.b2, &w.w3, &w.b3}) {
    memcpy(f, v->data(), v->size() * 4);
    f += v->size();
  }
  return GZ_OK;
}

gz_status gz_mlp_destroy(gz_mlp_t h) {
  delete h;
  return GZ_OK;
}

gz_status gz_mlp_predict(gz_mlp_t h, const uint8_t *codes, size_t n, float *out,
                         float *kernel_ms) {
  GZ_REQUIRE(h && (n == 0 || (codes && out)), "gz_mlp_predict: null argument");
  GZ_HIP_CHECK(h->dev.Forward(codes, n, out));
  if (kernel_ms) *kernel_ms = h->dev.last_forward_ms;
  return GZ_OK;
}

gz_status gz_mlp_predict_swapped(gz_mlp_t h, const uint8_t *codes, size_t n, float *out) {
  GZ_REQUIRE(h && (n == 0 || (codes && out)), "gz_mlp_predict_swapped: null argument");
  GZ_HIP_CHECK(h->dev.Forward(codes, n, out, kCostResidues / 8));
  return GZ_OK;
}

gz_status gz_mlp_train_step(gz_mlp_t h, const uint8_t *codes, const float *targets, size_t n,
                            float lr, float *loss, float *flat_grad) {
  GZ_REQUIRE(h && codes && targets && n > 0, "gz_mlp_train_step: bad argument");
  GZ_HIP_CHECK(h->dev.TrainStep(codes, targets, n, lr, loss, flat_grad));
  return GZ_OK;
}

gz_status gz_mlp_train_epochs(gz_mlp_t h, const uint8_t *codes, const float *targets, size_t n,
                              const uint32_t *order, size_t order_length, size_t batch, float lr,
                              float *mean_loss, float *kernel_ms) {
  GZ_REQUIRE(h && codes && targets && order && n > 0 && batch > 0,
             "gz_mlp_train_epochs: bad argument");
  for (size_t i = 0; i < order_length; ++i)
    GZ_REQUIRE(order[i] < n, "gz_mlp_train_epochs: order index out of range");
  GZ_HIP_CHECK(h->dev.TrainEpochs(codes, targets, n, order, order_length, batch, lr, mean_loss,
                                  kernel_ms));
  return GZ_OK;
}

gz_status gz_mlp_get_weights(gz_mlp_t h, float *flat) {
  GZ_REQUIRE(h && flat, "null argument");
  CostModelWeights w;
  GZ_HIP_CHECK(h->dev.GetWeights(&w));
  float *f = flat;
  for (auto *v : {&w.embed, &w.w1, &w.b1, &w.w2, &w.b2, &w.w3, &w.b3}) {
    memcpy(f, v->data(), v->size() * 4);
    f += v->size();
  }
  return GZ_OK;
}

gz_status gz_cost_features_packed_async(const uint64_t *pq, const uint64_t *pd,
                                        const uint32_t *qpos, const uint32_t *dpos, size_t n,
                                        int reverse, uint8_t *out, gz_stream_t stream) {
  GZ_REQUIRE(n < (1ull << 32), "gz_cost_features_packed_async: too many seeds");
  GZ_HIP_CHECK(CostFeaturesOnDevice(pq, pd, qpos, dpos, (uint32_t)n, reverse != 0, out,
                                    stream ? stream->s : nullptr));
  return GZ_OK;
}

void gz_cost_features(const uint8_t *queries, const uint8_t *database, const uint32_t *qpos,
                      const uint32_t *dpos, size_t n, int reverse, uint8_t *out) {
  for (size_t i = 0; i < n; ++i)
    CostFeatures(queries, database, qpos[i], dpos[i], reverse != 0, out + i * kCostResidues);
}

}  // extern "C"
