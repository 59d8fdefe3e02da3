// statistics.h -- Karlin-Altschul statistics of the search path.
//
// Per-query ungapped lambda/K feed the ungapped X-drop cutoff and the gapped
// trigger (src/aligner_gpu.cpp:213-242); the hard-coded gapped parameters of
// BLOSUM62 11/1 feed the gapped cutoffs and the E-values of the output
// (src/statistics.cpp:28-269).  The arithmetic keeps the reference's float /
// double mix so that the integer cutoffs and the printed values agree.
#pragma once

#include <stdint.h>
#include <vector>

namespace gz {

struct KarlinParameters {
  float lambda;
  float K;
  float H;
};

// Karlin-Altschul block calculation over a score distribution p[min..max]
// (lambda by Newton-Raphson with a bisection fallback, H, K by the
// Karlin-Altschul series; ext/karlin/src/karlin.c:17-31).
void KarlinBlockCalc(const double *p_at_min, int min, int max, float *lambda,
                     float *K, float *H);

class Statistics {
 public:
  Statistics();  // protein, Robinson background (src/statistics.cpp:136-176)
  // src/statistics.cpp:193-229
  bool IdealUngapped(const int *matrix, uint32_t nl, KarlinParameters *p) const;
  bool QueryUngapped(const uint8_t *seq, uint32_t length, const int *matrix,
                     uint32_t nl, KarlinParameters *p) const;
  static KarlinParameters GappedBlosum62();  // 11/1: lambda .267 K .041 H .14
  static uint64_t SearchSpace(uint64_t query_length, uint64_t db_length,
                              uint64_t db_sequences, const KarlinParameters &p,
                              double alpha, double beta);

  static int Normalized2Nominal(float normalized, const KarlinParameters &p);
  static int NormalizedCutoff2NominalCutoff(float normalized,
                                            const KarlinParameters &p);
  static float Nominal2Normalized(int score, const KarlinParameters &p);
  static double Nominal2EValue(int score, uint64_t search_space,
                               const KarlinParameters &p);

 private:
  bool ScoreProbabilities(const int *matrix, uint32_t nl, const double *f1,
                          const double *f2, int lowest, int highest,
                          double *p_at_lowest) const;
  std::vector<double> letter_prob_;
};

}  // namespace gz
