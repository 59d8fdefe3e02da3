// cost_model_gpu.h -- device inference of the learned extension cost.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include "cost_model.h"

namespace gz {

bool UnflattenWeights(const float *flat, size_t count, CostModelWeights *w);
size_t FlatWeightCount();

class GpuCostModel : public CostModel {
 public:
  explicit GpuCostModel(const CostModelWeights &w);
  ~GpuCostModel() override;
  void Predict(const uint8_t *queries, const uint8_t *database, const uint32_t *qpos,
               const uint32_t *dpos, size_t n, bool reverse, float *out) const override;

 private:
  struct Impl;
  Impl *impl_;
};

}  // namespace gz
