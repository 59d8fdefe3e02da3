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
#include "gz_device.h"
#include "mlp.h"

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

template <class Cursor>
__device__ __forceinline__ void Window(Cursor &c, uint8_t *out) {
  bool stopped = false;
  for (int k = 0; k < kCostWindow; ++k) {
    uint32_t x = 20;
    if (!stopped) {
      const uint32_t v = c.Next();
      if (v == kDelimiter) stopped = true;
      else x = v < 20 ? v : 20;
    }
    out[k] = (uint8_t)x;
  }
}

// CostFeatures (cost_model.cpp) on packed codes: windows along the extension
// (query, subject) then away from it (query, subject).
__global__ void CostFeatureKernel(const uint64_t *pq, const uint64_t *pd, const uint32_t *qpos,
                                  const uint32_t *dpos, uint32_t n, int reverse, uint8_t *out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint8_t w[kCostResidues];
  const uint32_t q = qpos[i], d = dpos[i];
  if (reverse) {
    ReverseCursor a, b;
    a.Init(pq, q);
    b.Init(pd, d);
    Window(a, w);
    Window(b, w + kCostWindow);
    ForwardCursor c, e;
    c.Init(pq, q + 1);
    e.Init(pd, d + 1);
    Window(c, w + 2 * kCostWindow);
    Window(e, w + 3 * kCostWindow);
  } else {
    ForwardCursor a, b;
    a.Init(pq, q);
    b.Init(pd, d);
    Window(a, w);
    Window(b, w + kCostWindow);
    ReverseCursor c, e;
    c.Init(pq, q - 1);
    e.Init(pd, d - 1);
    Window(c, w + 2 * kCostWindow);
    Window(e, w + 3 * kCostWindow);
  }
  uint4 *dst = reinterpret_cast<uint4 *>(out + (size_t)i * kCostResidues);
  const uint4 *src = reinterpret_cast<const uint4 *>(w);
#pragma unroll
  for (int k = 0; k < kCostResidues / 16; ++k) dst[k] = src[k];
}

}  // namespace

hipError_t CostFeaturesOnDevice(const uint64_t *pq, const uint64_t *pd, const uint32_t *qpos,
                                const uint32_t *dpos, uint32_t n, bool reverse, uint8_t *out,
                                hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(CostFeatureKernel, dim3((n + 255) / 256), dim3(256), 0, s, pq, pd, qpos,
                     dpos, n, reverse ? 1 : 0, out);
  return hipGetLastError();
}

struct GpuCostModel::Impl {
  MlpDevice mlp;
  uint8_t *codes = nullptr;
  size_t codes_cap = 0;
  ~Impl() {
    if (codes) (void)hipFree(codes);
  }
};

GpuCostModel::GpuCostModel(const CostModelWeights &w) : impl_(new Impl()) {
  impl_->mlp.SetWeights(w);
}

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
  if (n == 0) return hipSuccess;
  const size_t chunk = (size_t)1 << 22;
  const size_t want = std::min<size_t>(n, chunk) * kCostResidues;
  if (impl_->codes_cap < want) {
    if (impl_->codes) (void)hipFree(impl_->codes);
    impl_->codes = nullptr;
    impl_->codes_cap = 0;
    const hipError_t e = hipMalloc((void **)&impl_->codes, want);
    if (e != hipSuccess) return e;
    impl_->codes_cap = want;
  }
  for (size_t b = 0; b < n; b += chunk) {
    const uint32_t m = (uint32_t)std::min<size_t>(chunk, n - b);
    hipError_t e = CostFeaturesOnDevice(pq, pd, d_qpos + b, d_dpos + b, m, reverse,
                                        impl_->codes, s);
    if (e == hipSuccess) e = impl_->mlp.InferDevice(impl_->codes, m, d_out + b, s);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace gz
