// ungapped.h -- the seed tests of the presearch on packed codes, twelve
// residues (one 64-bit word) per step:
//
//  * PairSide / UngappedWithTrigger: SeedSearcher::UngappedExtend
//    (src/seed_searcher.h:133-158) as a walk over whole query words (the
//    pair walk below).
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

constexpr uint64_t kCodeBitsMask = (1ull << (kCodeBits * kCodesPerWord)) - 1;

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

// ---------------------------------------------------------------------------
// The pair walk: one side of an ungapped X-drop extension, twelve residue
// pairs per step.
//
// A step takes the twelve query codes and the twelve database codes of the
// next window (a funnel shift of each sequence's current word pair, the
// first window starting at the seed) and merges them into ten-bit table
// indices with two bit-field inserts:
//   E = (D << 5) on the fields [5i+5, 5i+10) of even i, Q elsewhere: bits
//       [5i, 5i+10) of E hold (b_i << 5) | a_i for even i,
//   O = Q on the fields [5i, 5i+5) of odd i, (D >> 5) elsewhere: bits
//       [5i-5, 5i+5) of O hold (a_i << 5) | b_i for odd i,
// (a = query code, b = database code), so residue pair i costs one field
// extract and one table read: even pairs read the table ba[(b << 5) | a],
// odd pairs its transpose ab[(a << 5) | b].  With byte tables the field is
// the LDS byte address itself.
//
// Scoring (SeedSearcher::UngappedExtend, src/seed_searcher.h:133-158, over
// UngappedExtender::ExtendOneSideScoreOnly, src/ungapped_extender.h:34-80):
// the reference walks one residue at a time, committing the running sum
// whenever it rises above the best score and stopping at a delimiter or
// when the sum falls to cutoff below the best.  For cutoffs >= 1 that is:
// best = max(0, max prefix sum) over the residues before the first stop,
// a stop being a delimiter or a prefix sum at or below best - cutoff.  (A
// cutoff below 1 stops the reference only after a non-positive residue;
// callers reject such cutoffs.)  A delimiter (and any code >= nl) scores the
// table's stop score, which is a stop for every cutoff up to -stop: c <=
// best, so c + stop <= best - cutoff.  Byte tables (stop -128) serve cutoffs
// in [1, 128] and matrices within [-127, 127]; 16-bit tables (stop -16384)
// the rest.
//
// Reads: the step reads each sequence's word after the current pair before
// scoring, so their latency overlaps the scoring.  A side stops in the
// window holding the first delimiter of either sequence, and both packed
// arrays carry four delimiter words in front and four words behind, so the
// read-ahead stays inside them (words 2 up to last + 2).
// ---------------------------------------------------------------------------
constexpr int kStopScore = -16384;          // 16-bit table stop score
constexpr int kByteStopScore = -128;        // byte table stop score
constexpr int kMaxUngappedCutoff = 16384;   // 16-bit tables: cutoffs up to this
constexpr int kMaxByteCutoff = 128;         // byte tables: cutoffs up to this
// fields [5i+5, 5i+10) of the even pairs i = 0, 2, .., 10
constexpr uint64_t kPairFieldMask = 0x0f83e0f83e0f83e0ull;

template <typename T>
struct PairTables {
  T ba[kTableDim * kTableDim];  // [(b << 5) | a]
  T ab[kTableDim * kTableDim];  // [(a << 5) | b]
};

template <typename T>
GZ_HD constexpr int PairStopScore() {
  return sizeof(T) == 1 ? kByteStopScore : kStopScore;
}

// Score of query code a against database code b (matrix row = database
// code, as in the reference's sm[d * nl + q]).
template <typename T>
GZ_HD T PairEntry(const int *matrix, uint32_t nl, uint32_t a, uint32_t b) {
  return (a < nl && b < nl) ? (T)matrix[b * nl + a] : (T)PairStopScore<T>();
}

// Entry i of both tables (i < 1024).
template <typename T>
GZ_HD void PairTablesFill(PairTables<T> &t, const int *matrix, uint32_t nl, uint32_t i) {
  t.ba[i] = PairEntry<T>(matrix, nl, i & 31u, i >> 5);
  t.ab[i] = PairEntry<T>(matrix, nl, i >> 5, i & 31u);
}

// (x & mask) | (y & ~mask), one v_bfi_b32 per half on the device (the
// compiler splits a constant-mask select into and / and / or)
GZ_HD uint32_t Bfi32(uint32_t mask, uint32_t x, uint32_t y) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "s"(mask), "v"(x), "v"(y));
  return r;
#else
  return (x & mask) | (y & ~mask);
#endif
}

GZ_HD uint64_t Bfi64(uint64_t mask, uint64_t x, uint64_t y) {
  return (uint64_t)Bfi32((uint32_t)mask, (uint32_t)x, (uint32_t)y) |
         ((uint64_t)Bfi32((uint32_t)(mask >> 32), (uint32_t)(x >> 32), (uint32_t)(y >> 32))
          << 32);
}

struct PairWalk {
  uint64_t q0, q1, d0, d1;  // the current word pair of each sequence
  uint32_t qi, di;          // word indices of q0 and d0
  uint32_t qr, ql, dr, dl;  // each pair's funnel shifts: 5 k and 60 - 5 k
  int c, best;
};

template <bool kReverse>
GZ_HD PairWalk PairStart(const SeedWindow &sq, const SeedWindow &sd) {
  PairWalk w;
  w.q0 = kReverse ? sq.w0 : sq.w1;
  w.q1 = kReverse ? sq.w1 : sq.w2;
  w.d0 = kReverse ? sd.w0 : sd.w1;
  w.d1 = kReverse ? sd.w1 : sd.w2;
  w.qi = kReverse ? sq.wi : sq.wi + 1;
  w.di = kReverse ? sd.wi : sd.wi + 1;
  w.qr = kCodeBits * sq.k;
  w.ql = kCodeBits * (kCodesPerWord - sq.k);
  w.dr = kCodeBits * sd.k;
  w.dl = kCodeBits * (kCodesPerWord - sd.k);
  w.c = 0;
  w.best = 0;
  return w;
}

// Scores one window; false once the side has stopped (w.best is then its
// score).
template <bool kReverse, typename T>
GZ_HD bool PairStep(const uint64_t *qw, const uint64_t *dw, const PairTables<T> *tab,
                    int cutoff, PairWalk &w) {
  const uint64_t qn = (kReverse ? qw[w.qi - 1] : qw[w.qi + 2]) & kCodeBitsMask;
  const uint64_t dn = (kReverse ? dw[w.di - 1] : dw[w.di + 2]) & kCodeBitsMask;
  // Window12 of each pair; with k = 0 bits 60..63 hold stray bits, which no
  // field below reads
  const uint64_t Q = (w.q0 >> w.qr) | (w.q1 << w.ql);
  const uint64_t D = (w.d0 >> w.dr) | (w.d1 << w.dl);
  const uint64_t E = Bfi64(kPairFieldMask, D << kCodeBits, Q);
  const uint64_t O = Bfi64(kPairFieldMask, Q, D >> kCodeBits);
  int t[kCodesPerWord];
#pragma unroll
  for (int p = 0; p < (int)kCodesPerWord / 2; ++p) {
    t[2 * p] = tab->ba[(uint32_t)(E >> (10 * p)) & 1023u];
    t[2 * p + 1] = tab->ab[(uint32_t)(O >> (10 * p)) & 1023u];
  }
  int c = w.c, best = w.best;
  bool alive = true;
#pragma unroll
  for (int j = 0; j < (int)kCodesPerWord; ++j) {
    c += t[kReverse ? (int)kCodesPerWord - 1 - j : j];  // unused once stopped
    best = alive && c > best ? c : best;
    alive = alive && c - best > -cutoff;
  }
  w.c = c;
  w.best = best;
  // the pairs move on before the caller's exit test, so the read-ahead is
  // used on every path and stays at the top of the step
  if (kReverse) {
    w.q1 = w.q0;
    w.q0 = qn;
    w.d1 = w.d0;
    w.d0 = dn;
    --w.qi;
    --w.di;
  } else {
    w.q0 = w.q1;
    w.q1 = qn;
    w.d0 = w.d1;
    w.d1 = dn;
    ++w.qi;
    ++w.di;
  }
  return alive;
}

template <bool kReverse, typename T>
GZ_HD int PairSide(const uint64_t *qw, const SeedWindow &sq, const uint64_t *dw,
                   const SeedWindow &sd, const PairTables<T> *tab, int cutoff) {
  PairWalk w = PairStart<kReverse>(sq, sd);
  while (PairStep<kReverse>(qw, dw, tab, cutoff, w)) {
  }
  return w.best;
}

// SeedSearcher::UngappedExtend: reverse from (qpos-1, dpos-1), forward from
// (qpos, dpos) only while the sum is <= trigger.  Cutoffs must lie in
// [1, -stop score of T].
template <typename T>
GZ_HD int UngappedWithTrigger(const uint64_t *qw, const SeedWindow &sq, const uint64_t *dw,
                              const SeedWindow &sd, const PairTables<T> *tab, int cutoff,
                              int trigger) {
  int sum = PairSide<true>(qw, sq, dw, sd, tab, cutoff);
  if (sum <= trigger) sum += PairSide<false>(qw, sq, dw, sd, tab, cutoff);
  return sum;
}

template <typename T>
GZ_HD int UngappedWithTrigger(const uint64_t *qw, uint32_t qpos, const uint64_t *dw,
                              uint32_t dpos, const PairTables<T> *tab, int cutoff, int trigger) {
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
