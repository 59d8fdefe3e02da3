// ungapped.h -- the seed tests of the presearch on packed codes, twelve
// residues (one 64-bit word) per step:
//
//  * UngappedSideFrom / UngappedWithTrigger: SeedSearcher::UngappedExtend
//    (src/seed_searcher.h:133-158).  The reference walks one residue at a
//    time, committing the running sum whenever it rises above the best
//    score and stopping at a delimiter or when the sum falls cutoff below
//    the best.  That is the same as: best = max(0, max prefix sum) over the
//    residues before the first stop, where a stop is a delimiter or a
//    prefix sum at or below best - cutoff.  Here a delimiter scores
//    kStopScore, which is a stop for every cutoff below -kStopScore, so a
//    step is branch-free: add, max, compare per residue.
//  * DistanceWindow / DistanceFrom: DistanceCalculator::CalculateDistance
//    (src/distance_calculator.cpp:35-54), the reduced-alphabet mismatches of
//    a window of up to 12 residues, through a 32 x 32 "differs" table.
//
// The functions compile for the host too (tests check them against the
// CPU oracle).
#pragma once

#include <stdint.h>

#include "gz_common.h"

#if defined(__HIP__)
#define GZ_HD __host__ __device__ __forceinline__
#else
#define GZ_HD inline
#endif

namespace gz {

constexpr int kStopScore = -16384;          // delimiter score in ungapped tables
constexpr int kMaxUngappedCutoff = 16384;   // cutoffs must stay below this
constexpr uint64_t kCodeBitsMask = (1ull << (kCodeBits * kCodesPerWord)) - 1;

// Ungapped table entry (row = database code, column = query code): the
// matrix score, kStopScore for codes >= nl (the delimiter among them).
GZ_HD int16_t UngappedTableEntry(const int *matrix, uint32_t nl, uint32_t i) {
  const uint32_t r = i / kTableDim, c = i % kTableDim;
  return (r < nl && c < nl) ? (int16_t)matrix[r * nl + c] : (int16_t)kStopScore;
}

// Twelve codes starting at packed code index k of word lo (k in [0, 12)):
// codes k..11 of lo, then 0..k-1 of hi; code j at bits [5j, 5j+5).
GZ_HD uint64_t Window12(uint64_t lo, uint64_t hi, uint32_t k) {
  return (lo >> (kCodeBits * k)) | (hi << (kCodeBits * (kCodesPerWord - k)));
}

// The 36 codes around a seed position `pos`: words w0..w2 hold codes
// [pos - 12, pos + 12) at offset k, so the reverse side's first twelve codes
// (pos - 12 .. pos - 1) are Window12(w0, w1, k) and the forward side's
// (pos .. pos + 11) Window12(w1, w2, k).  Loading both sides' words up front
// lets a kernel issue them together.
struct SeedWindow {
  uint64_t w0, w1, w2;
  uint32_t wi;  // word index of w0
  uint32_t k;
};

GZ_HD SeedWindow LoadSeedWindow(const uint64_t *packed, uint32_t pos) {
  SeedWindow s;
  const uint32_t c = pos - kCodesPerWord + kPackPadCodes;
  s.wi = c / kCodesPerWord;
  s.k = c - s.wi * kCodesPerWord;
  s.w0 = packed[s.wi] & kCodeBitsMask;
  s.w1 = packed[s.wi + 1] & kCodeBitsMask;
  s.w2 = packed[s.wi + 2] & kCodeBitsMask;
  return s;
}

// One side of an ungapped X-drop extension walked a window (twelve
// residues) at a time: the reverse side from pos - 1 backwards, the forward
// side from pos onwards.  The word after the current pair is read before the
// pair is scored, so its latency overlaps the scoring; it lies inside the
// padded array because the pair holds no delimiter before the side's last
// window (a side stops at the first one), and the front pad and the tail hold
// four words each.
struct SideWalk {
  uint64_t q0, q1, d0, d1;  // the current word pair of each sequence
  uint32_t qi, di;          // word index of q0 / d0
  uint32_t qk, dk;          // code offset of the window in the pair
  int c, best;
};

template <bool kReverse>
GZ_HD SideWalk SideStart(const SeedWindow &sq, const SeedWindow &sd) {
  SideWalk w;
  w.q0 = kReverse ? sq.w0 : sq.w1;
  w.q1 = kReverse ? sq.w1 : sq.w2;
  w.d0 = kReverse ? sd.w0 : sd.w1;
  w.d1 = kReverse ? sd.w1 : sd.w2;
  w.qi = kReverse ? sq.wi : sq.wi + 1;
  w.di = kReverse ? sd.wi : sd.wi + 1;
  w.qk = sq.k;
  w.dk = sd.k;
  w.c = 0;
  w.best = 0;
  return w;
}

// Scores one window; false once the side has stopped (w.best is then its
// score).
template <bool kReverse>
GZ_HD bool SideStep(const uint64_t *qw, const uint64_t *dw, const int16_t *tab, int cutoff,
                    SideWalk &w) {
  const uint64_t qn = kReverse ? qw[w.qi - 1] : qw[w.qi + 2];
  const uint64_t dn = kReverse ? dw[w.di - 1] : dw[w.di + 2];
  const uint64_t Q = Window12(w.q0, w.q1, w.qk), D = Window12(w.d0, w.d1, w.dk);
  int c = w.c, best = w.best;
  bool alive = true;
#pragma unroll
  for (int j = 0; j < (int)kCodesPerWord; ++j) {
    const int sh = (int)kCodeBits * (kReverse ? (int)kCodesPerWord - 1 - j : j);
    const uint32_t a = (uint32_t)(Q >> sh) & 31u, b = (uint32_t)(D >> sh) & 31u;
    c += tab[(b << 5) | a];  // c is not used once the side has stopped
    best = alive && c > best ? c : best;
    alive = alive && c - best > -cutoff;
  }
  w.c = c;
  w.best = best;
  // the pair moves on before the caller's exit test, so the read-ahead is
  // used on every path and stays at the top of the window
  if (kReverse) {
    w.q1 = w.q0;
    w.d1 = w.d0;
    w.q0 = qn & kCodeBitsMask;
    w.d0 = dn & kCodeBitsMask;
    --w.qi;
    --w.di;
  } else {
    w.q0 = w.q1;
    w.d0 = w.d1;
    w.q1 = qn & kCodeBitsMask;
    w.d1 = dn & kCodeBitsMask;
    ++w.qi;
    ++w.di;
  }
  return alive;
}

template <bool kReverse>
GZ_HD int UngappedSideFrom(const uint64_t *qw, const SeedWindow &sq, const uint64_t *dw,
                           const SeedWindow &sd, const int16_t *tab, int cutoff) {
  SideWalk w = SideStart<kReverse>(sq, sd);
  while (SideStep<kReverse>(qw, dw, tab, cutoff, w)) {
  }
  return w.best;
}

// SeedSearcher::UngappedExtend: reverse from (qpos-1, dpos-1), forward from
// (qpos, dpos) only while the sum is <= trigger.
GZ_HD int UngappedWithTrigger(const uint64_t *qw, const SeedWindow &sq, const uint64_t *dw,
                              const SeedWindow &sd, const int16_t *tab, int cutoff,
                              int trigger) {
  int sum = UngappedSideFrom<true>(qw, sq, dw, sd, tab, cutoff);
  if (sum <= trigger) sum += UngappedSideFrom<false>(qw, sq, dw, sd, tab, cutoff);
  return sum;
}

GZ_HD int UngappedWithTrigger(const uint64_t *qw, uint32_t qpos, const uint64_t *dw,
                              uint32_t dpos, const int16_t *tab, int cutoff, int trigger) {
  const SeedWindow sq = LoadSeedWindow(qw, qpos), sd = LoadSeedWindow(dw, dpos);
  return UngappedWithTrigger(qw, sq, dw, sd, tab, cutoff, trigger);
}

// Mismatches of the reduced codes over [qc - length/2, qc - length/2 +
// length) against the same window around dc; differs[(b << 5) | a] is 1
// when the reduced codes of a and b differ.
GZ_HD int DistanceWindow(const uint64_t *qw, uint32_t qc, const uint64_t *dw, uint32_t dc,
                         uint32_t length, const uint8_t *differs) {
  uint32_t qs = qc - length / 2 + kPackPadCodes, ds = dc - length / 2 + kPackPadCodes;
  int count = 0;
  for (uint32_t done = 0; done < length; done += kCodesPerWord) {
    const uint32_t qi = qs / kCodesPerWord, di = ds / kCodesPerWord;
    const uint64_t Q = Window12(qw[qi] & kCodeBitsMask, qw[qi + 1] & kCodeBitsMask,
                                qs - qi * kCodesPerWord);
    const uint64_t D = Window12(dw[di] & kCodeBitsMask, dw[di + 1] & kCodeBitsMask,
                                ds - di * kCodesPerWord);
#pragma unroll
    for (int j = 0; j < (int)kCodesPerWord; ++j) {
      const uint32_t a = (uint32_t)(Q >> (kCodeBits * j)) & 31u;
      const uint32_t b = (uint32_t)(D >> (kCodeBits * j)) & 31u;
      if (done + (uint32_t)j < length) count += differs[(b << 5) | a];
    }
    qs += kCodesPerWord;
    ds += kCodesPerWord;
  }
  return count;
}

// DistanceWindow from preloaded words (length <= 12: the window lies inside
// the 36 preloaded codes).
GZ_HD int DistanceFrom(const SeedWindow &sq, const SeedWindow &sd, uint32_t length,
                       const uint8_t *differs) {
  const uint32_t oq = sq.k + kCodesPerWord - length / 2, od = sd.k + kCodesPerWord - length / 2;
  const uint64_t Q = oq >= kCodesPerWord ? Window12(sq.w1, sq.w2, oq - kCodesPerWord)
                                         : Window12(sq.w0, sq.w1, oq);
  const uint64_t D = od >= kCodesPerWord ? Window12(sd.w1, sd.w2, od - kCodesPerWord)
                                         : Window12(sd.w0, sd.w1, od);
  int count = 0;
#pragma unroll
  for (int j = 0; j < (int)kCodesPerWord; ++j) {
    const uint32_t a = (uint32_t)(Q >> (kCodeBits * j)) & 31u;
    const uint32_t b = (uint32_t)(D >> (kCodeBits * j)) & 31u;
    if ((uint32_t)j < length) count += differs[(b << 5) | a];
  }
  return count;
}

// differs table entry i = (b << 5) | a for the 25-entry reduced map; codes
// from 25 up are each their own class.
GZ_HD uint8_t DiffersEntry(const uint8_t *reduced_map, uint32_t i) {
  const uint32_t b = i >> 5, a = i & 31u;
  const uint32_t ra = a < kNumberLetters ? reduced_map[a] : 100u + a;
  const uint32_t rb = b < kNumberLetters ? reduced_map[b] : 100u + b;
  return ra != rb ? 1 : 0;
}

}  // namespace gz
