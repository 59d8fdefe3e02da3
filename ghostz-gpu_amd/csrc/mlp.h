// mlp.h -- the cost-model MLP on the device: embedding 21x8 per residue,
// 1024 -> 256 -> 64 -> 1 with GELU, bf16 operands, fp32 accumulation (MFMA),
// fp32 master weights and Adam for training.
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime.h>

#include <vector>

#include "cost_model.h"

namespace gz {

class MlpDevice {
 public:
  MlpDevice();
  ~MlpDevice();
  // hipSuccess, or the first error of the constructor's stream creation and
  // allocations (every other call then returns it without doing anything)
  hipError_t status() const { return init_status_; }
  hipError_t SetWeights(const CostModelWeights &w);
  hipError_t GetWeights(CostModelWeights *w);
  // Predicted log2 cells for n seeds given their 128 feature codes each.
  hipError_t Forward(const uint8_t *codes, size_t n, float *out, uint32_t rotate_words = 0);
  // One training step on a batch: mean squared error against targets;
  // computes gradients, applies Adam (lr) unless lr == 0, and optionally
  // returns the flat gradient (same layout as the flat weights).
  hipError_t TrainStep(const uint8_t *codes, const float *targets, size_t n, float lr,
                       float *loss, float *flat_grad);
  // Training on a device-resident set: codes (n x 128) and targets are
  // uploaded once, then every whole batch of `order` (m row indices, e.g.
  // the epochs' permutations one after the other) is one TrainStep without
  // a host round trip; *mean_loss is the mean of the steps' losses and
  // *kernel_ms the device time of the steps (HIP events).
  hipError_t TrainEpochs(const uint8_t *codes, const float *targets, size_t n,
                         const uint32_t *order, size_t m, size_t batch, float lr,
                         float *mean_loss, float *kernel_ms);
  // Forward on device-resident codes (n x 128) into device outputs, on the
  // caller's stream (one persistent kernel for all three layers).
  // rotate_words: each 128-code row is read from its 32-bit word
  // rotate_words on, wrapping (16 swaps the row's two halves)
  // With d_count (device), only the first min(n, *d_count - count_offset)
  // rows are inferred: n sizes the launch, the device decides the rows.
  hipError_t InferDevice(const uint8_t *d_codes, size_t n, float *d_out, hipStream_t stream,
                         uint32_t rotate_words = 0, const uint32_t *d_count = nullptr,
                         uint32_t count_offset = 0);
  // Kernel-only forward time of the last Forward call (HIP events).
  float last_forward_ms = 0;

 private:
  hipError_t Reserve(size_t batch);
  hipError_t ForwardDevice(size_t n, size_t padded);
  hipError_t StepOnDevice(size_t n, size_t padded, float lr, long base, const uint32_t *done);
  void SyncBf16();

  hipError_t init_status_ = hipSuccess;
  hipStream_t stream_ = nullptr;
  size_t cap_ = 0;
  // fp32 master weights + gradients + Adam moments, flat
  float *master_ = nullptr, *grad_ = nullptr, *m_ = nullptr, *v_ = nullptr;
  uint16_t *w1h_ = nullptr, *w2h_ = nullptr, *embh_ = nullptr;  // bf16 copies
  size_t count_ = 0;
  long step_ = 0;
  // activations
  uint8_t *codes_ = nullptr;
  uint16_t *x_ = nullptr;       // [B][1024] bf16
  float *z1_ = nullptr;         // [B][256]
  uint16_t *h1_ = nullptr;      // [B][256] bf16
  float *z2_ = nullptr;         // [B][64]
  uint16_t *h2_ = nullptr;      // [B][64] bf16
  float *y_ = nullptr, *t_ = nullptr, *loss_ = nullptr;
  float *dz2_ = nullptr;        // [B][64]
  uint16_t *dz2h_ = nullptr;
  float *dz1_ = nullptr;        // [B][256]
  uint16_t *dz1h_ = nullptr;
  float *dh1_ = nullptr;        // [B][256]
  float *dx_ = nullptr;         // [B][1024]
};

}  // namespace gz
