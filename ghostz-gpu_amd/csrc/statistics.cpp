// statistics.cpp -- see statistics.h.
//
// The Karlin-Altschul routines follow the published NCBI algorithm as used
// by the reference (ext/karlin/src/karlin.c): the operation order is kept
// because lambda and K are rounded to float and then floored into integer
// cutoffs, so any reassociation could move a cutoff by one.
#include "statistics.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "gz_common.h"

namespace gz {

namespace {

constexpr int kScoreMin = -10000;
constexpr int kScoreMax = 1000;
constexpr int kScoreRangeMax = kScoreMax - kScoreMin;
constexpr int kKIterMax = 100;
constexpr double kKSumLimit = 0.01;
constexpr double kLambda0 = 0.5;
constexpr double kLambdaAccuracy = 1.e-5;
constexpr int kLambdaBisectionIterations = 17;

bool BadScoreRange(int lo, int hi) {
  if (lo >= 0 || hi <= 0 || lo < kScoreMin || hi > kScoreMax) return true;
  return hi - lo > kScoreRangeMax;
}

long Gcd(long a, long b) {
  b = b >= 0 ? b : -b;
  if (b > a) std::swap(a, b);
  while (b != 0) {
    long c = a % b;
    a = b;
    b = c;
  }
  return a;
}

// sum_i p[i] * e^{lambda i}, evaluated by repeated multiplication when
// e^{lambda (low-1)} is representable.
double ExpSum(const double *p, int low, int high, double lambda) {
  double x0 = std::exp(lambda);
  double x1 = std::pow(x0, low - 1);
  double sum = 0.;
  if (x1 > 0.) {
    for (int i = low; i <= high; ++i) sum += p[i] * (x1 *= x0);
  } else {
    for (int i = low; i <= high; ++i) sum += p[i] * std::exp(lambda * i);
  }
  return sum;
}

double LambdaBisection(const double *p, int low, int high) {
  if (BadScoreRange(low, high)) return -1.;
  double up = kLambda0, lambda = 0.;
  for (;;) {
    up *= 2;
    if (ExpSum(p, low, high, up) >= 1.0) break;
    lambda = up;
  }
  for (int j = 0; j < kLambdaBisectionIterations; ++j) {
    const double mid = (lambda + up) / 2.;
    if (ExpSum(p, low, high, mid) > 1.0)
      up = mid;
    else
      lambda = mid;
  }
  return (lambda + up) / 2.;
}

double LambdaNewton(const double *p, int low, int high) {
  if (BadScoreRange(low, high)) return -1.;
  double lambda = kLambda0;
  for (int j = 0; j < 20; ++j) {
    double sum = -1.0, slope = 0.0;
    if (lambda < 0.01) break;
    const double x0 = std::exp(lambda);
    double x1 = std::pow(x0, low - 1);
    if (x1 == 0.) break;
    for (int i = low; i <= high; ++i) {
      const double t = p[i] * (x1 *= x0);
      sum += t;
      slope += t * i;
    }
    const double step = sum / slope;
    lambda -= step;
    const double rel = step / lambda;
    if ((rel >= 0 ? rel : -rel) < kLambdaAccuracy) {
      if (lambda > kLambdaAccuracy) return lambda;
      break;
    }
  }
  return LambdaBisection(p, low, high);
}

double Entropy(const double *p, int low, int high, double lambda) {
  if (lambda < 0.) return -1.;
  if (BadScoreRange(low, high)) return -1.;
  const double etolam = std::exp(lambda);
  double etolami = std::pow(etolam, low - 1);
  double av = 0.0;
  if (etolami > 0.) {
    for (int s = low; s <= high; ++s) av += p[s] * s * (etolami *= etolam);
  } else {
    for (int s = low; s <= high; ++s) av += p[s] * s * std::exp(lambda * s);
  }
  return lambda * av;
}

// K from lambda and H: the series of Karlin & Altschul (PNAS 87, 1990)
// evaluated by repeated convolution of the score distribution, plus a
// geometric tail correction and the lattice gcd.
double KFromLambdaH(const double *p_scores, int low, int high, double lambda,
                    double H) {
  if (lambda <= 0. || H <= 0.) return -1.;
  const int range = high - low;
  const double av = H / lambda;
  const double etolam = std::exp(lambda);
  if (low == -1 || high == 1) return av * (1.0 - 1. / etolam);
  double sumlimit = kKSumLimit;
  const int iter = kKIterMax;
  if ((long)iter * range + 1 > (long)kKIterMax * kScoreRangeMax + 1) return -1.;
  std::vector<double> P(iter * range + 1);
  const double *pr = p_scores + low;  // pr[k] = probability of score low + k
  double Sum = 0., sum = 1., oldsum = 1., oldsum2 = 1.;
  int lo = 0, hi = 0, j;
  P[0] = 1.;
  for (j = 0; j < iter && sum > sumlimit;) {
    int first = range, last = range;
    lo += low;
    hi += high;
    // P <- P * p (convolution), written from the top index down
    for (long top = hi - lo; top >= 0; --top) {
      long a = top - first;
      const long ae = top - last;
      int b = first;
      double s = 0.;
      for (; a >= ae;) s += P[a--] * pr[b++];
      if (first) --first;
      if (top <= range) --last;
      P[top] = s;
      sum = s;
    }
    double etolami = std::pow(etolam, lo - 1);
    long idx = -1;
    sum = 0.;
    int i;
    for (i = lo; i != 0; ++i) {
      etolami *= etolam;
      sum += P[++idx] * etolami;
    }
    for (; i <= hi; ++i) sum += P[++idx];
    oldsum2 = oldsum;
    oldsum = sum;
    Sum += sum /= ++j;
  }
  const double ratio = oldsum / oldsum2;
  if (ratio >= (1.0 - sumlimit * 0.001)) return -1.;
  sumlimit *= 0.01;
  while (sum > sumlimit) {
    oldsum *= ratio;
    Sum += sum = oldsum / ++j;
  }
  int i;
  long g;
  for (i = 1, g = -low; i <= range && g > 1; ++i)
    if (pr[i]) g = Gcd(g, i);
  double K;
  if (g * etolam > 0.05) {
    const double etolami = std::pow(etolam, -g);
    K = g * std::exp(-2.0 * Sum) / (av * (1.0 - etolami));
  } else {
    K = -g * std::exp(-2.0 * Sum) / (av * (std::exp(-g * lambda) - 1));
  }
  return K;
}

}  // namespace

void KarlinBlockCalc(const double *p_at_min, int min, int max, float *lambda,
                     float *K, float *H) {
  const double *p = p_at_min - min;  // indexed by score
  const float l = (float)LambdaNewton(p, min, max);
  const float h = (float)Entropy(p, min, max, l);
  const float k = (float)KFromLambdaH(p, min, max, l, h);
  *lambda = l;
  *H = h;
  *K = k;
}

Statistics::Statistics() : letter_prob_(kNumberLetters, 0.0) {
  // Robinson & Robinson frequencies in alphabet-code order
  // (A R N D C Q E G H I L K M F P S T W Y V)
  static const double kRobinson[20] = {78.05, 51.29, 44.87, 53.64, 19.25,
                                       42.64, 62.95, 73.77, 21.99, 51.42,
                                       90.19, 57.44, 22.43, 38.56, 52.03,
                                       71.20, 58.41, 13.30, 32.16, 64.41};
  for (int c = 0; c < 20; ++c) letter_prob_[c] = kRobinson[c];
  double sum = 0.0;
  for (uint32_t c = 0; c < kNumberLetters; ++c) sum += letter_prob_[c];
  for (uint32_t c = 0; c < kNumberLetters; ++c) {
    letter_prob_[c] /= sum;
    letter_prob_[c] *= 1.0;
  }
}

static void MatrixRange(const int *m, uint32_t nl, int *lowest, int *highest) {
  int hi = 0, lo = 0;
  for (uint32_t i = 0; i < nl * nl; ++i) {
    hi = std::max(hi, m[i]);
    lo = std::min(lo, m[i]);
  }
  *lowest = lo;
  *highest = hi;
}

bool Statistics::ScoreProbabilities(const int *m, uint32_t nl, const double *f1,
                                    const double *f2, int lowest, int highest,
                                    double *pl) const {
  double *p = pl - lowest;
  for (int i = lowest; i <= highest; ++i) p[i] = 0.0;
  for (uint32_t c1 = 0; c1 < kNumberLetters; ++c1) {
    for (uint32_t c2 = 0; c2 < kNumberLetters; ++c2) {
      p[m[c1 * nl + c2]] += f1[c1] * f2[c2];
    }
  }
  double sum = 0.0;
  for (int i = lowest; i <= highest; ++i)
    if (p[i] > 0.0) sum += p[i];
  if (sum <= 0.0) return false;
  for (int i = lowest; i <= highest; ++i)
    if (p[i] > 0.0) p[i] /= sum;
  return true;
}

bool Statistics::IdealUngapped(const int *m, uint32_t nl,
                               KarlinParameters *out) const {
  int lo, hi;
  MatrixRange(m, nl, &lo, &hi);
  std::vector<double> p(hi - lo + 1);
  if (!ScoreProbabilities(m, nl, letter_prob_.data(), letter_prob_.data(), lo,
                          hi, p.data()))
    return false;
  KarlinBlockCalc(p.data(), lo, hi, &out->lambda, &out->K, &out->H);
  return true;
}

bool Statistics::QueryUngapped(const uint8_t *seq, uint32_t length,
                               const int *m, uint32_t nl,
                               KarlinParameters *out) const {
  // composition of the regular letters, every code at least 1
  // (src/statistics.cpp:332-352), normalised to 1 (:313-330)
  std::vector<double> f(kNumberLetters, 0.0);
  for (uint32_t i = 0; i < length; ++i)
    if (seq[i] < 20) ++f[seq[i]];
  for (uint32_t c = 0; c < kNumberLetters; ++c)
    if (f[c] < 1.0) ++f[c];
  double sum = 0.0;
  for (uint32_t c = 0; c < kNumberLetters; ++c) sum += f[c];
  if (sum <= 0.0) return false;
  for (uint32_t c = 0; c < kNumberLetters; ++c) {
    f[c] /= sum;
    f[c] *= 1.0;
  }
  int lo, hi;
  MatrixRange(m, nl, &lo, &hi);
  std::vector<double> p(hi - lo + 1);
  if (!ScoreProbabilities(m, nl, f.data(), letter_prob_.data(), lo, hi,
                          p.data()))
    return false;
  KarlinBlockCalc(p.data(), lo, hi, &out->lambda, &out->K, &out->H);
  return true;
}

KarlinParameters Statistics::GappedBlosum62() {
  KarlinParameters p;
  p.lambda = 0.267f;
  p.K = 0.041f;
  p.H = 0.14f;
  return p;
}

int Statistics::Normalized2Nominal(float normalized, const KarlinParameters &p) {
  float s = (normalized * static_cast<float>(std::log(2.0)) + std::log(p.K)) /
            p.lambda;
  return static_cast<int>(std::floor(s));
}

int Statistics::NormalizedCutoff2NominalCutoff(float normalized,
                                               const KarlinParameters &p) {
  float s = (normalized * static_cast<float>(std::log(2.0))) / p.lambda;
  return static_cast<int>(std::floor(s));
}

float Statistics::Nominal2Normalized(int score, const KarlinParameters &p) {
  return ((static_cast<float>(score) * p.lambda) - std::log(p.K)) /
         static_cast<float>(std::log(2.0));
}

double Statistics::Nominal2EValue(int score, uint64_t space,
                                  const KarlinParameters &p) {
  return space * p.K * std::exp(static_cast<double>(-1.0 * score * p.lambda));
}

uint64_t Statistics::SearchSpace(uint64_t m_len, uint64_t n_len, uint64_t n_seq,
                                 const KarlinParameters &p, double alpha,
                                 double beta) {
  // length adjustment: fixed point of
  //   l = alpha/lambda (log K + log((m - l)(n - N l))) + beta
  // (src/statistics.cpp:56-125)
  uint64_t adj = 0;
  const double m = (double)m_len, n = (double)n_len, N = (double)n_seq;
  const double log_k = std::log(p.K);
  const double adl = (double)(alpha / p.lambda);
  const double a = N, mb = m * N + n, c = n * m - std::max(m, n) / p.K;
  if (c < 0) {
    adj = 0;
  } else {
    double e_min = 0, e_max = 2 * c / (mb + std::sqrt((mb * mb - 4 * a * c)));
    bool converged = false;
    double e = 0, e_next = 0.0;
    for (int i = 0; i < 20; ++i) {
      e = e_next;
      const double ss = (m - e) * (n - N * e);
      const double e_bar = adl * (log_k + std::log(ss)) + beta;
      if (e_bar >= e) {
        e_min = e;
        if (e_bar - e_min <= 1.0) {
          converged = true;
          break;
        }
        if (e_min >= e_max) break;
      } else {
        e_max = e;
      }
      if (e_min <= e_bar && e_bar <= e_max)
        e_next = e_bar;
      else
        e_next = (i == 0) ? e_max : (e_min + e_max) / 2;
    }
    if (converged) {
      adj = static_cast<int>(e_min);
      e = std::ceil(e_min);
      if (e <= e_max) {
        const double ss = (m - e) * (n - N * e);
        if (adl * (log_k + std::log(ss)) + beta > e) adj = static_cast<int>(e);
      }
    } else {
      adj = static_cast<int>(e_min);
    }
  }
  const uint64_t eff_db = n_len - (n_seq * adj);
  return eff_db * (m_len - adj);
}

}  // namespace gz
