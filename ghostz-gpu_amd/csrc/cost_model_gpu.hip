// cost_model_gpu.hip -- device inference of the learned extension cost
// (see cost_model.h).  Feature codes are gathered on the device straight from
// the packed query / database sequences; the MLP runs through the fused bf16
// MFMA inference kernels of mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "cost_model_gpu.h"
#include "gz_common.h"
#include "gz_runtime.h"
#include "gz_device.h"
#include "mlp.h"
#include "ungapped.h"

namespace gz {

size_t FlatWeightCount() {
  return (size_t)kCostAlphabet * kCostEmbed + (size_t)kCostH1 * kCostIn + kCostH1 +
         (size_t)kCostH2 * kCostH1 + kCostH2 + kCostH2 + 1;
}

bool UnflattenWeights(const float *f, size_t count, CostModelWeights *w) {
  if (count != FlatWeightCount()) return false;
  auto take = [&](std::vector<float> &v, size_t n) {
    v.assign(f, f + n);
    f += n;
  };
  take(w->embed, (size_t)kCostAlphabet * kCostEmbed);
  take(w->w1, (size_t)kCostH1 * kCostIn);
  take(w->b1, kCostH1);
  take(w->w2, (size_t)kCostH2 * kCostH1);
  take(w->b2, kCostH2);
  take(w->w3, kCostH2);
  take(w->b3, 1);
  return true;
}

namespace {

// One feature window: 32 codes from position `start` stepping forward or
// backward, codes >= 20 as 20, the delimiter and everything after it as 20
// (CostFeatures, cost_model.cpp).  Twelve codes come per packed word pair and
// the three pairs are consecutive words at one offset, so the four words are
// read at once: for any start inside the sequences they lie between the
// 48-code front pad and the four tail words.
template <bool kBackward>
__device__ __forceinline__ void FeatureWindow(const uint64_t *packed, uint32_t start,
                                              uint32_t out[8]) {
  // part p covers codes start + 12 p + j (forward) or start - 12 p - j (backward)
  const uint32_t c0 = (kBackward ? start - 11u : start) + kPackPadCodes;
  const uint32_t q = c0 / kCodesPerWord, off = c0 - q * kCodesPerWord;
  const uint32_t wb = kBackward ? q - 2u : q;
  uint64_t wd[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) wd[i] = packed[wb + i] & kCodeBitsMask;
  bool stopped = false;
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = 0x14141414u;
#pragma unroll
  for (int part = 0; part < 3; ++part) {
    const uint64_t win = kBackward ? Window12(wd[2 - part], wd[3 - part], off)
                                   : Window12(wd[part], wd[part + 1], off);
#pragma unroll
    for (int j = 0; j < 12; ++j) {
      const int k = 12 * part + j;  // index in the window
      if (k >= kCostWindow) break;
      const uint32_t code = (uint32_t)(win >> (kCodeBits * (kBackward ? 11 - j : j))) & 31u;
      stopped = stopped || code == kDelimiter;
      const uint32_t x = stopped ? 20u : (code < 20u ? code : 20u);
      out[k >> 2] = (out[k >> 2] & ~(0xffu << (8 * (k & 3)))) | (x << (8 * (k & 3)));
    }
  }
}

// CostFeatures (cost_model.cpp) on packed codes, one thread per (seed,
// window): windows along the extension (query, subject) then away from it
// (query, subject).  The grid is bounded and strides over the rows, so a
// launch sized for more rows than the device-side count holds costs little.
__global__ void CostFeatureKernel(const uint64_t *pq, const uint64_t *pd, const uint32_t *qpos,
                                  const uint32_t *dpos, const uint32_t *index, uint32_t n,
                                  const uint32_t *count, uint32_t count_offset, int reverse,
                                  uint8_t *out) {
  if (count) {  // rows beyond *count - count_offset are not computed
    const uint32_t c = *count;
    n = c > count_offset ? min(n, c - count_offset) : 0u;
  }
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 4ull * n;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = (uint32_t)(t >> 2), wi = (uint32_t)t & 3u;
    const uint32_t job = index ? index[i] : i;  // row i holds job index[i]
    const uint64_t *seq = (wi & 1u) ? pd : pq;
    const uint32_t pos = (wi & 1u) ? dpos[job] : qpos[job];
    // along: from pos in the extension's direction; away: from the next
    // position the other way
    const bool along = wi < 2u;
    const bool backward = along == (reverse != 0);
    const uint32_t start = along ? pos : (reverse ? pos + 1 : pos - 1);
    uint32_t codes[8];
    if (backward) FeatureWindow<true>(seq, start, codes);
    else FeatureWindow<false>(seq, start, codes);
    uint4 *dst = reinterpret_cast<uint4 *>(out + (size_t)i * kCostResidues + wi * kCostWindow);
    dst[0] = make_uint4(codes[0], codes[1], codes[2], codes[3]);
    dst[1] = make_uint4(codes[4], codes[5], codes[6], codes[7]);
  }
}

}  // namespace

hipError_t CostFeaturesOnDevice(const uint64_t *pq, const uint64_t *pd, const uint32_t *qpos,
                                const uint32_t *dpos, uint32_t n, bool reverse, uint8_t *out,
                                hipStream_t s, const uint32_t *index, const uint32_t *d_count,
                                uint32_t count_offset) {
  if (n == 0) return hipSuccess;
  const size_t blocks = std::min<size_t>((4 * (size_t)n + 255) / 256, 4096);
  hipLaunchKernelGGL(CostFeatureKernel, dim3((unsigned)blocks), dim3(256), 0, s, pq, pd, qpos,
                     dpos, index, n, d_count, count_offset, reverse ? 1 : 0, out);
  return hipGetLastError();
}

struct GpuCostModel::Impl {
  MlpDevice mlp;
  uint8_t *codes = nullptr;
  size_t codes_cap = 0;
  ~Impl() {
    if (codes) (void)gz::Free(codes);
  }
};

GpuCostModel::GpuCostModel(const CostModelWeights &w) : impl_(new Impl()) {
  (void)impl_->mlp.SetWeights(w);
}

hipError_t GpuCostModel::status() const { return impl_->mlp.status(); }

GpuCostModel::~GpuCostModel() { delete impl_; }

void GpuCostModel::Predict(const uint8_t *queries, const uint8_t *database,
                           const uint32_t *qpos, const uint32_t *dpos, size_t n,
                           bool reverse, float *out) const {
  if (n == 0) return;
  std::vector<uint8_t> codes(n * kCostResidues);
  for (size_t i = 0; i < n; ++i)
    CostFeatures(queries, database, qpos[i], dpos[i], reverse, &codes[i * kCostResidues]);
  // a failed prediction only loses the ordering: equal costs keep the
  // plain order (the schedule never changes results)
  if (impl_->mlp.Forward(codes.data(), n, out) != hipSuccess)
    std::fill(out, out + n, 0.f);
}

hipError_t GpuCostModel::PredictOnDevice(const uint64_t *pq, const uint64_t *pd,
                                         const uint32_t *d_qpos, const uint32_t *d_dpos,
                                         uint32_t n, bool reverse, float *d_out,
                                         hipStream_t s) const {
  return Predict2(pq, pd, d_qpos, d_dpos, n, reverse, d_out, nullptr, s);
}

hipError_t GpuCostModel::PredictBothSidesOnDevice(const uint64_t *pq, const uint64_t *pd,
                                                  const uint32_t *d_qpos,
                                                  const uint32_t *d_dpos, uint32_t n,
                                                  float *d_out_reverse, float *d_out_forward,
                                                  hipStream_t s) const {
  return Predict2(pq, pd, d_qpos, d_dpos, n, true, d_out_reverse, d_out_forward, s);
}

hipError_t GpuCostModel::PredictIndexedOnDevice(const uint64_t *pq, const uint64_t *pd,
                                                const uint32_t *d_qpos, const uint32_t *d_dpos,
                                                const uint32_t *d_index, uint32_t m,
                                                bool reverse, float *d_out, hipStream_t s,
                                                const uint32_t *d_count) const {
  if (m == 0) return hipSuccess;
  const size_t chunk = (size_t)1 << 22;
  const size_t want = std::min<size_t>(m, chunk) * kCostResidues;
  if (impl_->codes_cap < want) {
    if (impl_->codes) (void)gz::Free(impl_->codes);
    impl_->codes = nullptr;
    impl_->codes_cap = 0;
    const hipError_t e = gz::Malloc((void **)&impl_->codes, want);
    if (e != hipSuccess) return e;
    impl_->codes_cap = want;
  }
  for (size_t b = 0; b < m; b += chunk) {
    const uint32_t k = (uint32_t)std::min<size_t>(chunk, m - b);
    hipError_t e = CostFeaturesOnDevice(pq, pd, d_qpos, d_dpos, k, reverse, impl_->codes, s,
                                        d_index + b, d_count, (uint32_t)b);
    if (e == hipSuccess)
      e = impl_->mlp.InferDevice(impl_->codes, k, d_out + b, s, 0, d_count, (uint32_t)b);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// features of n jobs on side `reverse` into d_out; with d_out_other, also the
// other side's predictions from the same rows, halves swapped
hipError_t GpuCostModel::Predict2(const uint64_t *pq, const uint64_t *pd,
                                  const uint32_t *d_qpos, const uint32_t *d_dpos, uint32_t n,
                                  bool reverse, float *d_out, float *d_out_other,
                                  hipStream_t s) const {
  if (n == 0) return hipSuccess;
  const size_t chunk = (size_t)1 << 22;
  const size_t want = std::min<size_t>(n, chunk) * kCostResidues;
  if (impl_->codes_cap < want) {
    if (impl_->codes) (void)gz::Free(impl_->codes);
    impl_->codes = nullptr;
    impl_->codes_cap = 0;
    const hipError_t e = gz::Malloc((void **)&impl_->codes, want);
    if (e != hipSuccess) return e;
    impl_->codes_cap = want;
  }
  for (size_t b = 0; b < n; b += chunk) {
    const uint32_t m = (uint32_t)std::min<size_t>(chunk, n - b);
    hipError_t e = CostFeaturesOnDevice(pq, pd, d_qpos + b, d_dpos + b, m, reverse,
                                        impl_->codes, s);
    if (e == hipSuccess) e = impl_->mlp.InferDevice(impl_->codes, m, d_out + b, s);
    if (e == hipSuccess && d_out_other)
      e = impl_->mlp.InferDevice(impl_->codes, m, d_out_other + b, s, kCostResidues / 8);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace gz
