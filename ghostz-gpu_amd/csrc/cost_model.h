// cost_model.h -- learned cost of a gapped X-drop extension.
//
// A small MLP (embedding 21 x 8 per residue, 128 residues = the 32-residue
// windows of query and subject on either side of the seed, 1024 -> 256 ->
// 64 -> 1 with GELU) predicts log2 of the DP cells the one-sided extension
// will evaluate.  The prediction only orders and bins extensions for load
// balance; it never changes which extensions run or what they return.
#pragma once

#include <stdint.h>
#include <stddef.h>

#include <vector>

namespace gz {

constexpr int kCostWindow = 32;                    // residues per side
constexpr int kCostResidues = 4 * kCostWindow;     // query/subject x left/right
constexpr int kCostEmbed = 8;
constexpr int kCostAlphabet = 21;                  // 20 amino acids + other
constexpr int kCostIn = kCostResidues * kCostEmbed;  // 1024
constexpr int kCostH1 = 256;
constexpr int kCostH2 = 64;

// Parameters in fp32 master form; the device kernels use bf16 copies.
struct CostModelWeights {
  std::vector<float> embed;  // [21][8]
  std::vector<float> w1, b1; // [256][1024], [256]
  std::vector<float> w2, b2; // [64][256], [64]
  std::vector<float> w3, b3; // [64], [1]
  void Init(uint64_t seed);
  size_t Count() const;
};

class CostModel {
 public:
  virtual ~CostModel() {}
  // Predicted log2 cells for seeds (qpos, dpos) in the concatenated query /
  // chunk sequences (unpacked codes); reverse selects the side.
  virtual void Predict(const uint8_t *queries, const uint8_t *database,
                       const uint32_t *qpos, const uint32_t *dpos, size_t n,
                       bool reverse, float *out) const = 0;
};

// Feature codes for one seed: 128 residue indexes in [0, 21) in four windows
// of 32: query along the extension, subject along the extension, query away
// from it, subject away from it (positions at or beyond a delimiter map to
// 20).  A hit's forward side (seed one position later) sees the reverse
// side's windows with the two halves of the row swapped.
void CostFeatures(const uint8_t *queries, const uint8_t *database, uint32_t qpos,
                  uint32_t dpos, bool reverse, uint8_t *out128);

}  // namespace gz
