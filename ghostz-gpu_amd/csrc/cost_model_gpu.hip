// cost_model_gpu.hip -- device inference of the learned extension cost
// (see cost_model.h).  Feature codes are gathered on the host from the
// unpacked sequences; the MLP runs on the device through the bf16 MFMA
// kernels of mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "cost_model_gpu.h"
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

struct GpuCostModel::Impl {
  MlpDevice mlp;
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

}  // namespace gz
