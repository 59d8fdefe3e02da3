// cost_model.cpp -- feature extraction and initialisation of the learned
// gapped-extension cost model (see cost_model.h).
#include "cost_model.h"

#include <cmath>
#include <random>

#include "gz_common.h"

namespace gz {

void CostModelWeights::Init(uint64_t seed) {
  std::mt19937_64 rng(seed);
  auto fill = [&](std::vector<float> &v, size_t n, float scale) {
    std::normal_distribution<float> d(0.f, scale);
    v.resize(n);
    for (size_t i = 0; i < n; ++i) v[i] = d(rng);
  };
  fill(embed, kCostAlphabet * kCostEmbed, 0.5f);
  fill(w1, (size_t)kCostH1 * kCostIn, std::sqrt(2.0f / kCostIn));
  b1.assign(kCostH1, 0.f);
  fill(w2, (size_t)kCostH2 * kCostH1, std::sqrt(2.0f / kCostH1));
  b2.assign(kCostH2, 0.f);
  fill(w3, kCostH2, std::sqrt(1.0f / kCostH2));
  b3.assign(1, 0.f);
}

size_t CostModelWeights::Count() const {
  return embed.size() + w1.size() + b1.size() + w2.size() + b2.size() + w3.size() + b3.size();
}

void CostFeatures(const uint8_t *queries, const uint8_t *database, uint32_t qpos,
                  uint32_t dpos, bool reverse, uint8_t *out) {
  const int inc = reverse ? -1 : 1;
  // window k: (sequence, start, step); the first two look along the
  // extension, the last two away from it
  const uint8_t *seqs[4] = {queries, database, queries, database};
  const int64_t starts[4] = {(int64_t)qpos, (int64_t)dpos, (int64_t)qpos - inc,
                             (int64_t)dpos - inc};
  const int steps[4] = {inc, inc, -inc, -inc};
  for (int w = 0; w < 4; ++w) {
    int64_t p = starts[w];
    bool stopped = false;
    for (int k = 0; k < kCostWindow; ++k) {
      uint8_t c = 20;
      if (!stopped && p >= 0) {
        const uint8_t x = seqs[w][p];
        if (x == kDelimiter) stopped = true;
        else c = x < 20 ? x : 20;
      }
      out[w * kCostWindow + k] = c;
      p += steps[w];
    }
  }
}

}  // namespace gz
