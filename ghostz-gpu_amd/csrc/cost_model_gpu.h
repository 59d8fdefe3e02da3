// cost_model_gpu.h -- device inference of the learned extension cost.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

#include "cost_model.h"

namespace gz {

bool UnflattenWeights(const float *flat, size_t count, CostModelWeights *w);
size_t FlatWeightCount();

// Feature codes of n one-sided jobs straight from the packed sequences (the
// same codes CostFeatures produces from unpacked ones): out is n x 128.
// With `index`, row i holds job index[i] of the position arrays; with
// d_count (device), only rows below *d_count - count_offset are written.
hipError_t CostFeaturesOnDevice(const uint64_t *packed_queries, const uint64_t *packed_database,
                                const uint32_t *qpos, const uint32_t *dpos, uint32_t n,
                                bool reverse, uint8_t *out, hipStream_t stream,
                                const uint32_t *index = nullptr,
                                const uint32_t *d_count = nullptr, uint32_t count_offset = 0);

// The device predictions share one feature scratch buffer: a model serves one
// stream at a time (each aligner owns its own model and calls it from its
// schedule stream only, on one host thread).  Concurrent calls on two streams would overwrite
// each other's features, which changes only the predicted order, never a
// search result.
class GpuCostModel : public CostModel {
 public:
  explicit GpuCostModel(const CostModelWeights &w);
  ~GpuCostModel() override;
  // hipSuccess, or the error of the device allocations made at construction
  hipError_t status() const;
  void Predict(const uint8_t *queries, const uint8_t *database, const uint32_t *qpos,
               const uint32_t *dpos, size_t n, bool reverse, float *out) const override;
  // Predicted log2 cells of n jobs whose positions are on the device; all
  // work is queued on `stream` (features, fused MLP); out is a device array.
  hipError_t PredictOnDevice(const uint64_t *packed_queries, const uint64_t *packed_database,
                             const uint32_t *d_qpos, const uint32_t *d_dpos, uint32_t n,
                             bool reverse, float *d_out, hipStream_t stream) const;
  // Predictions for the jobs d_index[0..m) of one side (out[k] for job
  // d_index[k]); with d_count (device), for the first *d_count of them only
  // (m then bounds the count and sizes the launches).
  hipError_t PredictIndexedOnDevice(const uint64_t *packed_queries,
                                    const uint64_t *packed_database, const uint32_t *d_qpos,
                                    const uint32_t *d_dpos, const uint32_t *d_index, uint32_t m,
                                    bool reverse, float *d_out, hipStream_t stream,
                                    const uint32_t *d_count = nullptr) const;
  // Both sides of n hits: job k extends backwards from (d_qpos[k], d_dpos[k])
  // and its forward side starts one position later.  The forward side's
  // windows are the backward side's in the other order (along the
  // extension, then away), so one feature pass serves both: the forward
  // predictions read each row with its halves swapped.
  hipError_t PredictBothSidesOnDevice(const uint64_t *packed_queries,
                                      const uint64_t *packed_database, const uint32_t *d_qpos,
                                      const uint32_t *d_dpos, uint32_t n, float *d_out_reverse,
                                      float *d_out_forward, hipStream_t stream) const;

 private:
  hipError_t Predict2(const uint64_t *packed_queries, const uint64_t *packed_database,
                      const uint32_t *d_qpos, const uint32_t *d_dpos, uint32_t n, bool reverse,
                      float *d_out, float *d_out_other, hipStream_t stream) const;
  struct Impl;
  Impl *impl_;
};

}  // namespace gz
