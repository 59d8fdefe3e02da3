// ungapped.h -- the seed tests of the presearch on packed codes, twelve
// residues (one 64-bit word) per step:
//
//  * UngappedSide / UngappedWithTrigger: SeedSearcher::UngappedExtend
//    (src/seed_searcher.h:133-158).  The reference walks one residue at a
//    time, committing the running sum whenever it rises above the best
//    score and stopping at a delimiter or when the sum falls cutoff below
//    the best.  That is the same as: best = max(0, max prefix sum) over the
//    residues before the first stop, where a stop is a delimiter or a
//    prefix sum at or below best - cutoff.  Here a delimiter scores
//    kStopScore, which is a stop for every cutoff below -kStopScore, so a
//    step is branch-free: add, max, compare per residue.
//  * DistanceWindow: DistanceCalculator::CalculateDistance
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

// One side of the ungapped extension from (qpos, dpos): codes at qpos,
// qpos+1, ... (kReverse: qpos, qpos-1, ...) of the padded packed arrays.
// tab: 32 x 32 ungapped table.  Returns the best prefix sum (>= 0).
template <bool kReverse>
GZ_HD int UngappedSide(const uint64_t *qw, uint32_t qpos, const uint64_t *dw, uint32_t dpos,
                       const int16_t *tab, int cutoff) {
  // first window: [pos, pos + 12), or [pos - 11, pos] read backwards
  const uint32_t qs = (kReverse ? qpos - (kCodesPerWord - 1) : qpos) + kPackPadCodes;
  const uint32_t ds = (kReverse ? dpos - (kCodesPerWord - 1) : dpos) + kPackPadCodes;
  uint32_t qi = qs / kCodesPerWord, di = ds / kCodesPerWord;
  const uint32_t qk = qs - qi * kCodesPerWord, dk = ds - di * kCodesPerWord;
  uint64_t q0 = qw[qi] & kCodeBitsMask, q1 = qw[qi + 1] & kCodeBitsMask;
  uint64_t d0 = dw[di] & kCodeBitsMask, d1 = dw[di + 1] & kCodeBitsMask;
  int c = 0, best = 0;
  for (;;) {
    const uint64_t Q = Window12(q0, q1, qk), D = Window12(d0, d1, dk);
    bool alive = true;
#pragma unroll
    for (int j = 0; j < (int)kCodesPerWord; ++j) {
      const int sh = (int)kCodeBits * (kReverse ? (int)kCodesPerWord - 1 - j : j);
      const uint32_t a = (uint32_t)(Q >> sh) & 31u, b = (uint32_t)(D >> sh) & 31u;
      c += tab[(b << 5) | a];  // c is not used once the side has stopped
      best = alive && c > best ? c : best;
      alive = alive && c - best > -cutoff;
    }
    if (!alive) break;
    if (kReverse) {
      q1 = q0;
      d1 = d0;
      q0 = qw[--qi] & kCodeBitsMask;
      d0 = dw[--di] & kCodeBitsMask;
    } else {
      q0 = q1;
      d0 = d1;
      ++qi;
      ++di;
      q1 = qw[qi + 1] & kCodeBitsMask;
      d1 = dw[di + 1] & kCodeBitsMask;
    }
  }
  return best;
}

// SeedSearcher::UngappedExtend: reverse from (qpos-1, dpos-1), forward from
// (qpos, dpos) only while the sum is <= trigger.
GZ_HD int UngappedWithTrigger(const uint64_t *qw, uint32_t qpos, const uint64_t *dw,
                              uint32_t dpos, const int16_t *tab, int cutoff, int trigger) {
  int sum = UngappedSide<true>(qw, qpos - 1, dw, dpos - 1, tab, cutoff);
  if (sum <= trigger) sum += UngappedSide<false>(qw, qpos, dw, dpos, tab, cutoff);
  return sum;
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

// differs table entry i = (b << 5) | a for the 25-entry reduced map; codes
// from 25 up are each their own class.
GZ_HD uint8_t DiffersEntry(const uint8_t *reduced_map, uint32_t i) {
  const uint32_t b = i >> 5, a = i & 31u;
  const uint32_t ra = a < kNumberLetters ? reduced_map[a] : 100u + a;
  const uint32_t rb = b < kNumberLetters ? reduced_map[b] : 100u + b;
  return ra != rb ? 1 : 0;
}

}  // namespace gz
