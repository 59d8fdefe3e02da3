// extenders.hip -- the search-path kernels and the extender classes of the
// C-ABI (DistanceCalculatorGpu, UngappedExtenderGpu, GappedExtenderGpu).
//
// Kernels are integer-only and written for gfx950 (wave64):
//  * PackKernel           -- PackAlphabetCodeSequence on the device
//  * DistanceKernel       -- reduced-alphabet mismatches over a window,
//                            one pair per lane, reduced map in LDS
//  * UngappedTriggerKernel-- reverse X-drop, forward continuation while the
//                            sum is <= trigger; one seed per lane, the score
//                            table in LDS, codes streamed from packed words
//  * SeedScatter/Gather   -- SetQuerySeeds expansion via own max-scan
//  * GappedKernel<W>      -- the CPU's row-major affine X-drop
//                            (src/gapped_extender.cpp:306-471), one seed per
//                            lane, DP row kept in a W-column LDS ring; seeds
//                            whose band outgrows W go to a wide-ring pass.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gz_common.h"
#include "gz_device.h"
#include "gz_runtime.h"
#include "gz_scan.h"

namespace gz {

// ---------------------------------------------------------------------------
// packing
// ---------------------------------------------------------------------------
__global__ void PackKernel(const uint8_t *codes, size_t n, uint64_t *packed,
                           size_t words) {
  const size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words) return;
  uint64_t word = 0;
#pragma unroll
  for (uint32_t j = 0; j < kCodesPerWord; ++j) {
    const size_t ci = w * kCodesPerWord + j;
    uint64_t c = kDelimiter;
    if (ci >= kPackPadCodes && ci - kPackPadCodes < n) c = codes[ci - kPackPadCodes];
    word |= c << (kCodeBits * j);
  }
  packed[w] = word;
}

// ---------------------------------------------------------------------------
// distance (src/distance_calculator_gpu.cu:187-286 semantics)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    DistanceKernel(size_t size, const uint32_t *qcenters,
                   const uint32_t *dcenters, const uint64_t *queries,
                   const uint64_t *database, const uint8_t *reduced_map,
                   uint32_t length, char *out) {
  __shared__ uint8_t map[kTableDim];
  if (threadIdx.x < kTableDim) {
    map[threadIdx.x] = (threadIdx.x < kNumberLetters)
                           ? reduced_map[threadIdx.x]
                           : (uint8_t)(100 + threadIdx.x);
  }
  __syncthreads();
  const uint32_t half = length / 2;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < size;
       i += (size_t)gridDim.x * blockDim.x) {
    ForwardCursor a, b;
    a.Init(queries, qcenters[i] - half);
    b.Init(database, dcenters[i] - half);
    int count = 0;
    for (uint32_t k = 0; k < length; ++k) {
      count += (map[a.Next()] != map[b.Next()]) ? 1 : 0;
    }
    out[i] = (char)count;
  }
}

// ---------------------------------------------------------------------------
// ungapped X-drop with trigger (src/seed_searcher.h:133-158,
// src/ungapped_extender.h:34-82)
// ---------------------------------------------------------------------------
template <class Cursor>
__device__ __forceinline__ int UngappedSide(Cursor &q, Cursor &d,
                                            const int8_t *sm, uint32_t delim,
                                            int cutoff) {
  int best = 0;
  uint32_t qc = q.Next(), dc = d.Next();
  if (dc != delim && qc != delim) {
    int score = 0;
    do {
      if (dc == delim || qc == delim) break;
      score += sm[dc * kTableDim + qc];
      qc = q.Next();
      dc = d.Next();
      if (score > 0) {
        do {
          best += score;
          if (dc == delim || qc == delim) break;
          score = sm[dc * kTableDim + qc];
          qc = q.Next();
          dc = d.Next();
        } while (score > 0);
      }
    } while (score > -cutoff);
  }
  return best;
}

__device__ __forceinline__ int UngappedWithTrigger(
    const uint64_t *queries, uint32_t qpos, const uint64_t *database,
    uint32_t dpos, const int8_t *sm, uint32_t delim, int cutoff, int trigger) {
  ReverseCursor rq, rd;
  rq.Init(queries, qpos - 1);
  rd.Init(database, dpos - 1);
  int sum = UngappedSide(rq, rd, sm, delim, cutoff);
  if (sum <= trigger) {
    ForwardCursor fq, fd;
    fq.Init(queries, qpos);
    fd.Init(database, dpos);
    sum += UngappedSide(fq, fd, sm, delim, cutoff);
  }
  return sum;
}

__global__ void __launch_bounds__(256)
    UngappedTriggerKernel(size_t size, const uint32_t *qids,
                          const uint32_t *qpos, const uint32_t *dpos,
                          const uint64_t *queries, const uint64_t *database,
                          const int *matrix, uint32_t nl, uint32_t delim,
                          const int *cutoffs, const int *triggers,
                          char *flags) {
  __shared__ int8_t sm[kTableDim * kTableDim];
  StageScoreTable(sm, matrix, nl);
  __syncthreads();
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < size;
       i += (size_t)gridDim.x * blockDim.x) {
    const uint32_t q = qids[i];
    const int trigger = triggers[q];
    const int s = UngappedWithTrigger(queries, qpos[i], database, dpos[i], sm,
                                      delim, cutoffs[q], trigger);
    flags[i] = (s > trigger) ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------
// SetQuerySeeds expansion (src/ungapped_extender_gpu.cu:332-408, 818-852)
// ---------------------------------------------------------------------------
__global__ void SeedScatterKernel(size_t runs, size_t size,
                                  const uint32_t *starts, uint32_t *temp) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < runs && starts[i] < size) temp[starts[i]] = (uint32_t)i;
}

__global__ void SeedGatherKernel(size_t size, const uint32_t *run_index,
                                 const uint32_t *seed_ids,
                                 const uint32_t *seed_query_ids,
                                 const uint32_t *seed_query_positions,
                                 uint32_t *qids, uint32_t *qpos) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= size) return;
  const uint32_t s = seed_ids[run_index[i]];
  qids[i] = seed_query_ids[s];
  qpos[i] = seed_query_positions[s];
}

// ---------------------------------------------------------------------------
// gapped X-drop, one seed per lane
// ---------------------------------------------------------------------------
struct LaneRingLds {
  static constexpr bool kShort = false;
  int2 *cells;     // [W][64]
  uint8_t *codes;  // [W][64]
  uint32_t mask;
  int lane;
  __device__ __forceinline__ int2 Get(int i) const { return cells[(i & mask) * kWave + lane]; }
  __device__ __forceinline__ void Set(int i, int2 v) { cells[(i & mask) * kWave + lane] = v; }
  __device__ __forceinline__ void SetInRange(int i, int2 v) { Set(i, v); }
  __device__ __forceinline__ uint8_t &Code(int i) {
    return codes[(i & mask) * kWave + lane];
  }
};

// 16-bit cells (half the LDS of LaneRingLds, so more waves per CU).  Exact
// under two conditions.  The host uses it only when the cutoff is at most
// 8192 and gap open + extension at most 1024; then every cell value that is
// compared is above -16384 + the gap extension, and values stored below
// -16384 (the band's first gap column keeps decreasing row after row) can
// be clamped there without changing a comparison.  And the extension leaves
// the ring (as when its band outgrows it) once its best score passes 16384,
// so stored values stay below 16384 + 127, the largest int8 matrix entry.
// A row's cells are stored unclamped (SetInRange): a pruned cell stores
// kInitScore, and a kept cell's score and gap value are at least
// -(cutoff + gap open + extension) > -16384.
struct LaneRingLds16 {
  static constexpr bool kShort = true;
  static constexpr int kLow = -16384;
  short2 *cells;   // [W][64]
  uint8_t *codes;  // [W][64]
  uint32_t mask;
  int lane;
  __device__ __forceinline__ int2 Get(int i) const {
    const short2 v = cells[(i & mask) * kWave + lane];
    return make_int2(v.x, v.y);
  }
  __device__ __forceinline__ void Set(int i, int2 v) {
    cells[(i & mask) * kWave + lane] =
        make_short2((short)(v.x > kLow ? v.x : kLow), (short)(v.y > kLow ? v.y : kLow));
  }
  __device__ __forceinline__ void SetInRange(int i, int2 v) {
    cells[(i & mask) * kWave + lane] = make_short2((short)v.x, (short)v.y);
  }
  __device__ __forceinline__ uint8_t &Code(int i) {
    return codes[(i & mask) * kWave + lane];
  }
};

struct LaneRingGlobal {
  static constexpr bool kShort = false;
  int2 *cells;     // [W]
  uint8_t *codes;  // [W]
  uint32_t mask;
  __device__ __forceinline__ int2 Get(int i) const { return cells[i & mask]; }
  __device__ __forceinline__ void Set(int i, int2 v) { cells[i & mask] = v; }
  __device__ __forceinline__ void SetInRange(int i, int2 v) { Set(i, v); }
  __device__ __forceinline__ uint8_t &Code(int i) { return codes[i & mask]; }
};

struct GappedParams {
  const uint64_t *queries;
  const uint64_t *database;
  uint32_t delim;
  int gap_open;
  int gap_ext;
  int cutoff;
};

// Returns false when the band would not fit the ring (width >= W).
template <bool kReverse, class Ring>
__device__ bool GappedOneSide(Ring &R, const GappedParams &P, const int8_t *sm,
                              uint32_t qpos, uint32_t dpos, int *best_score,
                              int *best_p0, int *best_p1, uint32_t *ncells) {
  const int inc = kReverse ? -1 : 1;
  const int W = (int)R.mask + 1;
  *best_score = 0;
  *best_p0 = -inc;
  *best_p1 = -inc;
  *ncells = 0;
  // column c (c >= 1) holds seq0 code at qpos + (c - 1) * inc
  auto load_code = [&](int col) -> uint32_t {
    return PackedCodeAt(P.queries, (uint32_t)((int)qpos + (col - 1) * inc));
  };
  typename std::conditional<kReverse, ReverseCursor, ForwardCursor>::type dc;
  dc.Init(P.database, dpos);
  uint32_t d = dc.Next();
  uint32_t c = load_code(1);
  if (c == P.delim || d == P.delim) return true;

  const int gap_init = P.gap_open + P.gap_ext;
  const int cutoff = P.cutoff;
  int array_start = 0, array_end = 0, ai;
  int score = -gap_init;
  R.Set(0, make_int2(0, -gap_init));
  int col = 1;
  for (ai = 1; c != P.delim; ++ai) {
    if (score < -cutoff) break;
    if (ai >= W) return false;
    R.Set(ai, make_int2(score, score - gap_init));
    R.Code(ai) = (uint8_t)c;
    score -= P.gap_ext;
    c = load_code(++col);
  }
  if (c != P.delim) {
    if (ai >= W) return false;
    R.Set(ai, make_int2(kInitScore, kInitScore));
    R.Code(ai) = (uint8_t)c;
    array_end = ai + 1;
  } else {
    array_end = ai;
  }
  int max_score = 0, max_p0 = -inc, max_p1 = -inc;
  uint32_t cells = 0;
  for (int row = 0; d != P.delim; ++row, d = dc.Next()) {
    // a row's values exceed its starting best by at most one matrix entry
    if (Ring::kShort && max_score > 16384) return false;
    const int8_t *mrow = sm + d * kTableDim;
    int2 first = R.Get(array_start);
    int prev = first.x;
    first.x = (max_score - first.y > cutoff) ? kInitScore : first.y;
    int gap_row = first.x - gap_init;
    first.y -= P.gap_ext;
    R.Set(array_start, first);
    int last = array_start;
    const int as0 = array_start;
    // the next column's code is read a column ahead (the ring index wraps,
    // so the read past the band is harmless and unused), so a cell's three
    // LDS reads are in flight together
    uint32_t code_next = R.Code(as0 + 1);
    // the row's best cell: the last column whose score rose above the best
    int best_ai = -1;
#pragma unroll 2
    for (ai = as0 + 1; ai < array_end; ++ai) {
      const int2 cell = R.Get(ai);
      const uint32_t code = code_next;
      code_next = R.Code(ai + 1);
      int s = prev + mrow[code];
      prev = cell.x;
      const int gap_col = cell.y;
      s = (s < gap_col) ? gap_col : s;
      s = (s < gap_row) ? gap_row : s;
      // the X-drop test and both of its outcomes as selects (lanes of a
      // wave split between them nearly every cell)
      const bool prune = max_score - s > cutoff;
      array_start += (prune && array_start + 1 == ai) ? 1 : 0;
      last = prune ? last : ai;
      const bool better = !prune && s > max_score;
      best_ai = better ? ai : best_ai;
      max_score = better ? s : max_score;
      const int ng = s - gap_init;
      const int gc = (ng > gap_col - P.gap_ext) ? ng : gap_col - P.gap_ext;
      R.SetInRange(ai, prune ? make_int2(kInitScore, kInitScore) : make_int2(s, gc));
      gap_row = prune ? kInitScore : ((ng > gap_row - P.gap_ext) ? ng : gap_row - P.gap_ext);
    }
    cells += (uint32_t)max(array_end - as0 - 1, 0);
    if (best_ai >= 0) {
      max_p0 = (best_ai - 1) * inc;
      max_p1 = row * inc;
    }
    if (array_start + 1 == array_end) break;
    // seq0 position after the row loop is column array_end (old value)
    int p0col = array_end;
    if (last < array_end - 1) {
      array_end = last + 1;
      c = load_code(p0col);
    } else {
      c = load_code(p0col);
      while (gap_row >= (max_score - cutoff) && c != P.delim) {
        if (array_end - array_start >= W) return false;
        R.Set(array_end, make_int2(gap_row, gap_row - gap_init));
        R.Code(array_end) = (uint8_t)c;
        gap_row -= P.gap_ext;
        ++array_end;
        c = load_code(++p0col);
      }
    }
    if (c != P.delim) {
      if (array_end - array_start >= W) return false;
      // the appended column's code: cached if it was inside the band,
      // otherwise it is the column just inspected
      if (array_end == p0col) R.Code(array_end) = (uint8_t)c;
      R.Set(array_end, make_int2(kInitScore, kInitScore));
      ++array_end;
    }
  }
  *best_score = max_score;
  *best_p0 = max_p0;
  *best_p1 = max_p1;
  *ncells = cells;
  return true;
}

// GappedOneSide cut into a start and one call per DP row, so that a lane can
// hold one extension's state across the rounds of a persistent kernel and
// begin its next extension while the other lanes of its wave are still in
// theirs.  Begin and Row run exactly the statements of GappedOneSide, in its
// order; Row does one pass of its row loop.
enum : int { kStepRunning = 0, kStepDone = 1, kStepTooWide = 2 };

template <bool kReverse, class Ring>
struct GappedStepper {
  typename std::conditional<kReverse, ReverseCursor, ForwardCursor>::type dc;
  uint32_t qpos, dpos, d, c;
  int array_start, array_end, max_score, max_p0, max_p1, row;
  uint32_t cells;

  __device__ __forceinline__ uint32_t Code(const GappedParams &P, int col) const {
    const int inc = kReverse ? -1 : 1;
    return PackedCodeAt(P.queries, (uint32_t)((int)qpos + (col - 1) * inc));
  }

  __device__ int Begin(Ring &R, const GappedParams &P, uint32_t q, uint32_t dp) {
    const int inc = kReverse ? -1 : 1;
    const int W = (int)R.mask + 1;
    qpos = q;
    dpos = dp;
    max_score = 0;
    max_p0 = -inc;
    max_p1 = -inc;
    cells = 0;
    row = 0;
    dc.Init(P.database, dp);
    d = dc.Next();
    c = Code(P, 1);
    if (c == P.delim || d == P.delim) return kStepDone;
    const int gap_init = P.gap_open + P.gap_ext;
    int score = -gap_init;
    array_start = 0;
    R.Set(0, make_int2(0, -gap_init));
    int col = 1, ai;
    for (ai = 1; c != P.delim; ++ai) {
      if (score < -P.cutoff) break;
      if (ai >= W) return kStepTooWide;
      R.Set(ai, make_int2(score, score - gap_init));
      R.Code(ai) = (uint8_t)c;
      score -= P.gap_ext;
      c = Code(P, ++col);
    }
    if (c != P.delim) {
      if (ai >= W) return kStepTooWide;
      R.Set(ai, make_int2(kInitScore, kInitScore));
      R.Code(ai) = (uint8_t)c;
      array_end = ai + 1;
    } else {
      array_end = ai;
    }
    return kStepRunning;
  }

  __device__ int Row(Ring &R, const GappedParams &P, const int8_t *sm) {
    const int inc = kReverse ? -1 : 1;
    const int W = (int)R.mask + 1;
    const int gap_init = P.gap_open + P.gap_ext;
    const int cutoff = P.cutoff;
    if (d == P.delim) return kStepDone;
    if (Ring::kShort && max_score > 16384) return kStepTooWide;
    const int8_t *mrow = sm + d * kTableDim;
    int2 first = R.Get(array_start);
    int prev = first.x;
    first.x = (max_score - first.y > cutoff) ? kInitScore : first.y;
    int gap_row = first.x - gap_init;
    first.y -= P.gap_ext;
    R.Set(array_start, first);
    int last = array_start;
    const int as0 = array_start;
    uint32_t code_next = R.Code(as0 + 1);
    int best_ai = -1;
#pragma unroll 2
    for (int ai = as0 + 1; ai < array_end; ++ai) {
      const int2 cell = R.Get(ai);
      const uint32_t code = code_next;
      code_next = R.Code(ai + 1);
      int s = prev + mrow[code];
      prev = cell.x;
      const int gap_col = cell.y;
      s = (s < gap_col) ? gap_col : s;
      s = (s < gap_row) ? gap_row : s;
      // the X-drop test and both of its outcomes as selects (lanes of a
      // wave split between them nearly every cell)
      const bool prune = max_score - s > cutoff;
      array_start += (prune && array_start + 1 == ai) ? 1 : 0;
      last = prune ? last : ai;
      const bool better = !prune && s > max_score;
      best_ai = better ? ai : best_ai;
      max_score = better ? s : max_score;
      const int ng = s - gap_init;
      const int gc = (ng > gap_col - P.gap_ext) ? ng : gap_col - P.gap_ext;
      R.SetInRange(ai, prune ? make_int2(kInitScore, kInitScore) : make_int2(s, gc));
      gap_row = prune ? kInitScore : ((ng > gap_row - P.gap_ext) ? ng : gap_row - P.gap_ext);
    }
    cells += (uint32_t)max(array_end - as0 - 1, 0);
    if (best_ai >= 0) {
      max_p0 = (best_ai - 1) * inc;
      max_p1 = row * inc;
    }
    if (array_start + 1 == array_end) return kStepDone;
    int p0col = array_end;
    if (last < array_end - 1) {
      array_end = last + 1;
      c = Code(P, p0col);
    } else {
      c = Code(P, p0col);
      while (gap_row >= (max_score - cutoff) && c != P.delim) {
        if (array_end - array_start >= W) return kStepTooWide;
        R.Set(array_end, make_int2(gap_row, gap_row - gap_init));
        R.Code(array_end) = (uint8_t)c;
        gap_row -= P.gap_ext;
        ++array_end;
        c = Code(P, ++p0col);
      }
    }
    if (c != P.delim) {
      if (array_end - array_start >= W) return kStepTooWide;
      if (array_end == p0col) R.Code(array_end) = (uint8_t)c;
      R.Set(array_end, make_int2(kInitScore, kInitScore));
      ++array_end;
    }
    ++row;
    d = dc.Next();
    return kStepRunning;
  }
};

constexpr int kGappedBlock = 64;

// Jobs [begin, end) (through `order` when set: slot k runs job order[k],
// which must be below `limit`), or the `*list_count` jobs of `list` when
// list is set; a job whose band outgrows the W-column ring is appended to
// the overflow list for the next, wider ring.  Results land at each job's
// own index.
template <bool kReverse, int W, bool kShort>
__global__ void __launch_bounds__(kGappedBlock)
    GappedKernel(size_t begin, size_t end, const uint32_t *list, const uint32_t *list_count,
                 const uint32_t *order, size_t limit, GappedParams P, const int *matrix,
                 uint32_t nl, uint32_t *qpos, uint32_t *dpos, int *scores, uint32_t *cells,
                 uint32_t *overflow_count, uint32_t *overflow_list, uint32_t *failed) {
  using Ring = typename std::conditional<kShort, LaneRingLds16, LaneRingLds>::type;
  using Cell = typename std::conditional<kShort, short2, int2>::type;
  __shared__ int8_t sm[kTableDim * kTableDim];
  __shared__ Cell ring_cells[W * kWave];
  __shared__ uint8_t ring_codes[W * kWave];
  StageScoreTable(sm, matrix, nl);
  __syncthreads();
  Ring R;
  R.cells = ring_cells;
  R.codes = ring_codes;
  R.mask = W - 1;
  R.lane = threadIdx.x;
  const size_t n = list ? (size_t)*list_count : end - begin;
  for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x) {
    const size_t i = list ? (size_t)list[k] : order ? (size_t)order[begin + k] : begin + k;
    if (i >= limit) {  // an order entry out of range: counted, never read
      atomicAdd(failed, 1u);
      continue;
    }
    const uint32_t q = qpos[i], d = dpos[i];
    int s, p0, p1;
    uint32_t nc;
    const bool ok = GappedOneSide<kReverse>(R, P, sm, q, d, &s, &p0, &p1, &nc);
    if (ok) {
      scores[i] = s;
      qpos[i] = (uint32_t)((int)q + p0);
      dpos[i] = (uint32_t)((int)d + p1);
      if (cells) cells[i] = nc;
    }
    // the lanes whose band outgrew the ring go to the overflow list, one
    // atomic per wave
    const uint64_t m = __ballot(!ok);
    if (m != 0) {
      const int lane = __lane_id();
      const int leader = __ffsll((unsigned long long)m) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(overflow_count, (uint32_t)__popcll(m));
      base = __shfl(base, leader, kWave);
      if (!ok) overflow_list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = (uint32_t)i;
    }
  }
}

// The batched form's persistent gapped extension: jobs come from a device
// work queue (slot k of the queue is job order[k], or k), handed to the
// waves 64 at a time with one atomic on the queue head; inside a wave each
// lane takes the wave's next job as soon as its own extension ends, so a
// wave runs as long as its lanes' summed work rather than its longest job
// per round.  Extensions whose band outgrows the ring go to the overflow
// list as in GappedKernel.  *head must be zero at the launch.
template <bool kReverse, int W, bool kShort>
__global__ void __launch_bounds__(kGappedBlock)
    GappedQueueKernel(const uint32_t *order, uint32_t n, uint32_t *head, GappedParams P,
                      const int *matrix, uint32_t nl, uint32_t *qpos, uint32_t *dpos,
                      int *scores, uint32_t *cells, uint32_t *overflow_count,
                      uint32_t *overflow_list, uint32_t *failed) {
  using Ring = typename std::conditional<kShort, LaneRingLds16, LaneRingLds>::type;
  using Cell = typename std::conditional<kShort, short2, int2>::type;
  __shared__ int8_t sm[kTableDim * kTableDim];
  __shared__ Cell ring_cells[W * kWave];
  __shared__ uint8_t ring_codes[W * kWave];
  StageScoreTable(sm, matrix, nl);
  __syncthreads();
  Ring R;
  R.cells = ring_cells;
  R.codes = ring_codes;
  R.mask = W - 1;
  R.lane = threadIdx.x;
  const int lane = __lane_id();
  GappedStepper<kReverse, Ring> st;
  // lane state: 0 wants a job, 1 runs one, 2 the queue is empty
  int state = 0;
  uint32_t job = 0;
  // the wave's current batch of queue slots: [batch, batch + kWave), of
  // which `used` are handed out (wave-uniform)
  uint32_t batch = 0, used = kWave;
  for (;;) {
    // hand out queue slots to the lanes that want a job
    uint64_t want = __ballot(state == 0);
    while (want != 0) {
      if (used == kWave) {
        uint32_t b = 0;
        if (lane == 0) b = atomicAdd(head, (uint32_t)kWave);
        batch = __shfl(b, 0, kWave);
        used = 0;
      }
      const uint32_t take = min((uint32_t)__popcll(want), kWave - used);
      const uint32_t rank = (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
      int r = kStepRunning;
      bool finished = false;
      if (((want >> lane) & 1ull) && rank < take) {
        const uint32_t k = batch + used + rank;
        if (k >= n) {
          state = 2;
        } else {
          job = order ? order[k] : k;
          if (job >= n) {  // an order entry out of range: counted, never read
            atomicAdd(failed, 1u);
            state = 1;
            finished = true;
            r = kStepRunning;
          } else {
            r = st.Begin(R, P, qpos[job], dpos[job]);
            state = 1;
            finished = r != kStepRunning;
          }
        }
      }
      used += take;
      // extensions that end in Begin (a delimiter at the seed, or a first
      // row wider than the ring) are written now and the lane asks again
      if (finished && r == kStepDone) {
        scores[job] = st.max_score;
        qpos[job] = (uint32_t)((int)st.qpos + st.max_p0);
        dpos[job] = (uint32_t)((int)st.dpos + st.max_p1);
        if (cells) cells[job] = st.cells;
      }
      const bool over = finished && r == kStepTooWide;
      const uint64_t om = __ballot(over);
      if (om != 0) {
        const int leader = __ffsll((unsigned long long)om) - 1;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(overflow_count, (uint32_t)__popcll(om));
        base = __shfl(base, leader, kWave);
        if (over) overflow_list[base + (uint32_t)__popcll(om & ((1ull << lane) - 1ull))] = job;
      }
      if (finished) state = 0;
      want = __ballot(state == 0);
    }
    if (!__any(state == 1)) break;  // every lane has found the queue empty
    int r = kStepRunning;
    if (state == 1) r = st.Row(R, P, sm);
    if (state == 1 && r == kStepDone) {
      scores[job] = st.max_score;
      qpos[job] = (uint32_t)((int)st.qpos + st.max_p0);
      dpos[job] = (uint32_t)((int)st.dpos + st.max_p1);
      if (cells) cells[job] = st.cells;
    }
    const bool over = state == 1 && r == kStepTooWide;
    const uint64_t om = __ballot(over);
    if (om != 0) {
      const int leader = __ffsll((unsigned long long)om) - 1;
      uint32_t base = 0;
      if (lane == leader) base = atomicAdd(overflow_count, (uint32_t)__popcll(om));
      base = __shfl(base, leader, kWave);
      if (over) overflow_list[base + (uint32_t)__popcll(om & ((1ull << lane) - 1ull))] = job;
    }
    if (state == 1 && r != kStepRunning) state = 0;
  }
}

template <bool kReverse>
__global__ void __launch_bounds__(kGappedBlock)
    GappedWideKernel(GappedParams P, const int *matrix, uint32_t nl,
                     uint32_t *qpos, uint32_t *dpos, int *scores,
                     uint32_t *cells, const uint32_t *overflow_count,
                     const uint32_t *overflow_list, int2 *scratch_cells,
                     uint8_t *scratch_codes, uint32_t wide, uint32_t *failed) {
  __shared__ int8_t sm[kTableDim * kTableDim];
  StageScoreTable(sm, matrix, nl);
  __syncthreads();
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t slots = gridDim.x * blockDim.x;
  LaneRingGlobal R;
  R.cells = scratch_cells + (size_t)slot * wide;
  R.codes = scratch_codes + (size_t)slot * wide;
  R.mask = wide - 1;
  const uint32_t n = *overflow_count;
  for (uint32_t j = slot; j < n; j += slots) {
    const uint32_t i = overflow_list[j];
    const uint32_t q = qpos[i], d = dpos[i];
    int s, p0, p1;
    uint32_t nc;
    if (GappedOneSide<kReverse>(R, P, sm, q, d, &s, &p0, &p1, &nc)) {
      scores[i] = s;
      qpos[i] = (uint32_t)((int)q + p0);
      dpos[i] = (uint32_t)((int)d + p1);
      if (cells) cells[i] = nc;
    } else {
      atomicAdd(failed, 1u);
    }
  }
}

// ---------------------------------------------------------------------------
// gapped X-drop, one wave per extension (bands past the 64-column ring)
// ---------------------------------------------------------------------------
// GappedOneSide with a row's columns spread over the wave's lanes, 64 per
// pass, the row kept in the wave's slot of the global ring.  Inside a row
// the lane recurrence carries two values from column to column: the running
// best score M and the horizontal gap g.  Both become wave scans:
//   * g: with the reference's reset of g to kInitScore after a pruned cell
//     dropped, g(a+1) = max(s'(a) - (open + ext), g(a) - ext), s'(a) being
//     the cell's score before the horizontal gap (the diagonal step or the
//     vertical gap), since a gap opened from a cell that a gap itself
//     reached is never better (open >= 0); so g is the row's entry gap
//     decayed, or an exclusive max-scan of s'(a) - (open + ext) + a ext
//     less (a - 1) ext;
//   * M before column a: the row's starting best or the exclusive max-scan
//     of the row's scores (a pruned score is below M, so whether it counts
//     changes nothing).
// Dropping the reset is exact when kInitScore < -cutoff (the cutoff below
// 1024): a cell is kept only with a score of at least M - cutoff >= -cutoff,
// and every gap that runs through a pruned cell -- the dropped one, which
// starts below that cell's score, and the reset one, which starts at
// kInitScore -- stays below M - cutoff from there on, so it neither keeps a
// cell nor sets a kept cell's score, nor the row's final gap when the band
// grows past its end (which happens only from a kept last column).  Pruning,
// kept scores, stored cells, the band bounds and the best cell are then
// those of the lane form; the best cell is the last column whose score is
// above every score before it (the first column holding the row's maximum).
// Returns false when the band would not fit the ring.
template <bool kReverse>
__device__ bool WaveOneSide(int2 *ring, uint32_t mask, const GappedParams &P, const int8_t *sm,
                            uint32_t qpos, uint32_t dpos, int *best_score, int *best_p0,
                            int *best_p1, uint32_t *ncells) {
  const int inc = kReverse ? -1 : 1;
  const int W = (int)mask + 1;
  const int lane = __lane_id();
  constexpr int kNone = -(1 << 29);  // below every score (the scans' identity)
  *best_score = 0;
  *best_p0 = -inc;
  *best_p1 = -inc;
  *ncells = 0;
  // column c (c >= 1) is the query code at qpos + (c - 1) * inc
  auto code_at = [&](int col) -> uint32_t {
    return PackedCodeAt(P.queries, (uint32_t)((int)qpos + (col - 1) * inc));
  };
  typename std::conditional<kReverse, ReverseCursor, ForwardCursor>::type dc;
  dc.Init(P.database, dpos);
  uint32_t d = dc.Next();
  if (code_at(1) == P.delim || d == P.delim) return true;
  const int gap_init = P.gap_open + P.gap_ext;
  const int ge = P.gap_ext;
  const int cutoff = P.cutoff;
  // the first row: columns 1, 2, ... while the gap score -gap_init - (a - 1)
  // ext is at least -cutoff and the column is not the query's delimiter;
  // the first column that stops (a*) is the band's end, after a pruned
  // column when it is not the delimiter.  A pass reads kPackPadCodes
  // columns: every column before it is a residue, so its codes lie at most
  // that far past a residue, inside the delimiter padding around the
  // packed sequences.
  if (lane == 0) ring[0] = make_int2(0, -gap_init);
  int stop_col = 0;
  bool stop_delim = false;
  constexpr int kPass = (int)kPackPadCodes;
  static_assert(kPass <= kWave, "one column per lane");
  for (int a0 = 1;; a0 += kPass) {
    if (a0 > W) return false;  // the loop would write column W
    const int a = a0 + lane;
    const bool mine = lane < kPass;
    const int score = -gap_init - (a - 1) * ge;
    const bool delim = mine && code_at(a) == P.delim;
    const uint64_t stop = __ballot(mine && (delim || score < -cutoff));
    const int k = stop ? __ffsll((unsigned long long)stop) - 1 : kPass;
    if (a0 + k - 1 >= W) return false;  // the largest column the pass writes
    if (lane < k) ring[a & mask] = make_int2(score, score - gap_init);
    if (k < kPass) {
      stop_col = a0 + k;
      stop_delim = __shfl(delim ? 1 : 0, k, kWave) != 0;
      break;
    }
  }
  int array_start = 0, array_end = stop_col;
  if (!stop_delim) {
    if (stop_col >= W) return false;
    if (lane == 0) ring[stop_col & mask] = make_int2(kInitScore, kInitScore);
    array_end = stop_col + 1;
  }
  int max_score = 0, max_p0 = -inc, max_p1 = -inc;
  uint32_t cells = 0;
  for (int row = 0; d != P.delim; ++row, d = dc.Next()) {
    // the previous row's cells (and the first row's) are written
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const int8_t *mrow = sm + d * kTableDim;
    int2 first = ring[array_start & mask];
    const int prev0 = first.x;
    first.x = (max_score - first.y > cutoff) ? kInitScore : first.y;
    int gap_row = first.x - gap_init;
    first.y -= ge;
    if (lane == 0) ring[array_start & mask] = first;
    const int as0 = array_start, ae = array_end;
    int M = max_score;      // best score before the pass's first column
    int gin = gap_row;      // horizontal gap into the pass's first column
    int diag = prev0;       // the old score of the column before the pass
    int first_kept = -1, last_kept = -1, best_col = -1;
    for (int a0 = as0 + 1; a0 < ae; a0 += kWave) {
      const int a = a0 + lane;
      const bool on = a < ae;
      const int2 cell = on ? ring[a & mask] : make_int2(kNone, kNone);
      const uint32_t code = on ? code_at(a) : 0u;
      const int prev = WaveShiftUp1(cell.x, diag);
      diag = __builtin_amdgcn_readlane(cell.x, kWave - 1);  // used only after a full pass
      int sp = prev + mrow[code];
      sp = sp < cell.y ? cell.y : sp;
      // horizontal gap into column a
      const int h = on ? sp - gap_init + lane * ge : kNone;
      const int hx = WaveShiftUp1(WaveInclusiveMaxDpp(h, kNone), kNone);
      int g = gin - lane * ge;
      if (lane > 0) {
        const int from = hx - (lane - 1) * ge;
        g = g < from ? from : g;
      }
      const int sc = sp < g ? g : sp;
      // best score before column a
      const int mi = WaveInclusiveMaxDpp(on ? sc : kNone, kNone);
      int mx = WaveShiftUp1(mi, M);
      mx = mx < M ? M : mx;
      const bool prune = mx - sc > cutoff;
      if (on) {
        const int ng = sc - gap_init;
        const int gc = (ng > cell.y - ge) ? ng : cell.y - ge;
        ring[a & mask] = prune ? make_int2(kInitScore, kInitScore) : make_int2(sc, gc);
      }
      const uint64_t kept = __ballot(on && !prune);
      if (kept) {
        if (first_kept < 0) first_kept = a0 + __ffsll((unsigned long long)kept) - 1;
        last_kept = a0 + 63 - __clzll((long long)kept);
      }
      const uint64_t better = __ballot(on && !prune && sc > mx);
      if (better) best_col = a0 + 63 - __clzll((long long)better);
      // the gap after column a, the pass's carries
      const int gout = (sc - gap_init > g - ge) ? sc - gap_init : g - ge;
      const int tail = (ae - a0 < kWave ? ae - a0 : kWave) - 1;
      gin = __builtin_amdgcn_readlane(gout, tail);
      const int pm = __builtin_amdgcn_readlane(mi, tail);
      M = M < pm ? pm : M;
    }
    if (ae > as0 + 1) gap_row = gin;
    cells += (uint32_t)(ae - as0 - 1 > 0 ? ae - as0 - 1 : 0);
    array_start = first_kept >= 0 ? first_kept - 1 : ae - 1;
    const int last = last_kept >= 0 ? last_kept : as0;
    max_score = M;
    if (best_col >= 0) {
      max_p0 = (best_col - 1) * inc;
      max_p1 = row * inc;
    }
    if (array_start + 1 == array_end) break;
    // the band's end for the next row (as in GappedOneSide)
    int p0col = array_end;
    uint32_t c;
    if (last < array_end - 1) {
      array_end = last + 1;
      c = code_at(p0col);
    } else {
      c = code_at(p0col);
      while (gap_row >= (max_score - cutoff) && c != P.delim) {
        if (array_end - array_start >= W) return false;
        if (lane == 0) ring[array_end & mask] = make_int2(gap_row, gap_row - gap_init);
        gap_row -= ge;
        ++array_end;
        c = code_at(++p0col);
      }
    }
    if (c != P.delim) {
      if (array_end - array_start >= W) return false;
      if (lane == 0) ring[array_end & mask] = make_int2(kInitScore, kInitScore);
      ++array_end;
    }
  }
  *best_score = max_score;
  *best_p0 = max_p0;
  *best_p1 = max_p1;
  *ncells = cells;
  return true;
}

// The global ring's jobs, one wave per job (one wave per block), the
// block's slot of the global ring as its row store.
template <bool kReverse>
__global__ void __launch_bounds__(kWave)
    GappedWaveKernel(GappedParams P, const int *matrix, uint32_t nl, uint32_t *qpos,
                     uint32_t *dpos, int *scores, uint32_t *cells, const uint32_t *list_count,
                     const uint32_t *list, int2 *scratch_cells, uint32_t wide,
                     uint32_t *failed) {
  __shared__ int8_t sm[kTableDim * kTableDim];
  StageScoreTable(sm, matrix, nl);
  __syncthreads();
  int2 *ring = scratch_cells + (size_t)blockIdx.x * wide;
  const uint32_t n = *list_count;
  for (uint32_t j = blockIdx.x; j < n; j += gridDim.x) {
    const uint32_t i = list[j];
    const uint32_t q = qpos[i], d = dpos[i];
    int s, p0, p1;
    uint32_t nc;
    const bool ok = WaveOneSide<kReverse>(ring, wide - 1, P, sm, q, d, &s, &p0, &p1, &nc);
    if (threadIdx.x == 0) {
      if (ok) {
        scores[i] = s;
        qpos[i] = (uint32_t)((int)q + p0);
        dpos[i] = (uint32_t)((int)d + p1);
        if (cells) cells[i] = nc;
      } else {
        atomicAdd(failed, 1u);
      }
    }
  }
}

__global__ void ConvertSeedsKernel(size_t size, const uint32_t *seed_positions,
                                   const uint32_t *qin, const uint32_t *din,
                                   uint32_t *qout, uint32_t *dout) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= size) return;
  const uint32_t s = seed_positions[i];
  qout[i] = qin[s];
  dout[i] = din[s];
}

// compute units of the current device
inline unsigned Cus() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus < 1)
    cus = 256;
  return (unsigned)cus;
}

inline unsigned GridFor(size_t n, unsigned block, unsigned cap = 65535u * 4) {
  size_t g = (n + block - 1) / block;
  if (g == 0) g = 1;
  return (unsigned)std::min<size_t>(g, cap);
}

}  // namespace gz

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace gz;

static inline hipStream_t S(gz_stream_t st) { return st ? st->s : 0; }

extern "C" {

const char *gz_last_error(void) { return LastError(); }

gz_status gz_set_device(int device) {
  GZ_HIP_CHECK(hipSetDevice(device));
  return GZ_OK;
}
gz_status gz_device_malloc(void **ptr, size_t bytes) {
  GZ_REQUIRE(ptr != nullptr, "gz_device_malloc: null out pointer");
  if (bytes == 0) bytes = 1;
  hipError_t e = gz::Malloc(ptr, bytes);
  if (e != hipSuccess) {
    SetError(std::string("hipMalloc: ") + hipGetErrorString(e));
    return GZ_ERROR_OUT_OF_MEMORY;
  }
  return GZ_OK;
}
gz_status gz_device_free(void *ptr) {
  if (ptr) GZ_HIP_CHECK(gz::Free(ptr));
  return GZ_OK;
}
gz_status gz_host_malloc(void **ptr, size_t bytes) {
  GZ_REQUIRE(ptr != nullptr, "gz_host_malloc: null out pointer");
  if (bytes == 0) bytes = 1;
  hipError_t e = gz::HostMalloc(ptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    SetError(std::string("hipHostMalloc: ") + hipGetErrorString(e));
    return GZ_ERROR_OUT_OF_MEMORY;
  }
  return GZ_OK;
}
gz_status gz_host_free(void *ptr) {
  if (ptr) GZ_HIP_CHECK(gz::HostFree(ptr));
  return GZ_OK;
}
gz_status gz_stream_create(gz_stream_t *stream) {
  GZ_REQUIRE(stream != nullptr, "gz_stream_create: null out pointer");
  gz_stream_s *s = new gz_stream_s();
  hipError_t e = hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete s;
    SetError(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    return GZ_ERROR_HIP;
  }
  *stream = s;
  return GZ_OK;
}
gz_status gz_stream_destroy(gz_stream_t stream) {
  if (!stream) return GZ_OK;
  hipError_t e = hipStreamDestroy(stream->s);
  delete stream;
  GZ_HIP_CHECK(e);
  return GZ_OK;
}
gz_status gz_stream_synchronize(gz_stream_t stream) {
  GZ_HIP_CHECK(hipStreamSynchronize(S(stream)));
  return GZ_OK;
}
gz_status gz_memcpy_async(void *dst, const void *src, size_t bytes,
                          gz_stream_t stream) {
  if (bytes == 0) return GZ_OK;
  GZ_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, S(stream)));
  return GZ_OK;
}
gz_status gz_memset_async(void *dst, int value, size_t bytes,
                          gz_stream_t stream) {
  if (bytes == 0) return GZ_OK;
  GZ_HIP_CHECK(hipMemsetAsync(dst, value, bytes, S(stream)));
  return GZ_OK;
}

size_t gz_packed_words(size_t codes) { return PackedWords(codes); }

// Streaming copy in 16-byte vectors: the reference point for HBM counters
// (bytes fetched / written by a kernel of known traffic).
__global__ void CopyKernel(const uint4 *src, uint4 *dst, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

gz_status gz_copy_kernel_async(void *dst, const void *src, size_t bytes, gz_stream_t stream) {
  GZ_REQUIRE(bytes % 16 == 0 && ((uintptr_t)dst % 16) == 0 && ((uintptr_t)src % 16) == 0,
             "gz_copy_kernel_async: 16-byte aligned sizes and pointers only");
  if (bytes == 0) return GZ_OK;
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(CopyKernel, dim3(GridFor(n, 256, 8192)), dim3(256), 0, S(stream),
                     (const uint4 *)src, (uint4 *)dst, n);
  GZ_HIP_CHECK(hipGetLastError());
  return GZ_OK;
}

gz_status gz_pack_sequence_async(const uint8_t *d_codes, size_t n,
                                 uint64_t *d_packed, gz_stream_t stream) {
  GZ_REQUIRE(d_packed != nullptr, "gz_pack_sequence_async: null output");
  const size_t words = PackedWords(n);
  hipLaunchKernelGGL(PackKernel, dim3(GridFor(words, 256, 1u << 30)), dim3(256), 0,
                     S(stream), d_codes, n, d_packed, words);
  GZ_HIP_CHECK(hipGetLastError());
  return GZ_OK;
}

uint8_t gz_packed_code_at(const uint64_t *packed, size_t position) {
  const size_t p = position + kPackPadCodes;
  return (uint8_t)((packed[p / kCodesPerWord] >> (kCodeBits * (p % kCodesPerWord))) & 31u);
}

// ---- DistanceCalculatorGpu ----
struct gz_distance_calculator_s {
  const uint64_t *queries = nullptr;
  const uint64_t *database = nullptr;
  const uint8_t *reduced_map = nullptr;
  uint32_t length = 0;
};

gz_status gz_distance_calculator_create(gz_distance_calculator_t *h) {
  GZ_REQUIRE(h != nullptr, "null handle pointer");
  *h = new gz_distance_calculator_s();
  return GZ_OK;
}
gz_status gz_distance_calculator_destroy(gz_distance_calculator_t h) {
  delete h;
  return GZ_OK;
}
gz_status gz_distance_calculator_set_queries(gz_distance_calculator_t h,
                                             const uint64_t *d) {
  GZ_REQUIRE(h, "null handle");
  h->queries = d;
  return GZ_OK;
}
gz_status gz_distance_calculator_set_database(gz_distance_calculator_t h,
                                              const uint64_t *d) {
  GZ_REQUIRE(h, "null handle");
  h->database = d;
  return GZ_OK;
}
gz_status gz_distance_calculator_set_reduced_code_map(gz_distance_calculator_t h,
                                                      const uint8_t *d) {
  GZ_REQUIRE(h, "null handle");
  h->reduced_map = d;
  return GZ_OK;
}
gz_status gz_distance_calculator_set_subsequence_length(gz_distance_calculator_t h,
                                                        uint32_t length) {
  GZ_REQUIRE(h, "null handle");
  h->length = length;
  return GZ_OK;
}
gz_status gz_distance_calculator_calculate_distances_async(
    gz_distance_calculator_t h, size_t size, const uint32_t *h_qc,
    const uint32_t *h_dc, char *h_dist, uint32_t *d_qc, uint32_t *d_dc,
    char *d_dist, gz_stream_t stream) {
  GZ_REQUIRE(h && h->queries && h->database && h->reduced_map,
             "distance calculator: queries, database and reduced map must be set");
  if (size == 0) return GZ_OK;
  GZ_REQUIRE(d_qc && d_dc && d_dist, "distance calculator: null device buffer");
  hipStream_t s = S(stream);
  if (h_qc) GZ_HIP_CHECK(hipMemcpyAsync(d_qc, h_qc, size * 4, hipMemcpyDefault, s));
  if (h_dc) GZ_HIP_CHECK(hipMemcpyAsync(d_dc, h_dc, size * 4, hipMemcpyDefault, s));
  hipLaunchKernelGGL(DistanceKernel, dim3(GridFor(size, 256)), dim3(256), 0, s,
                     size, d_qc, d_dc, h->queries, h->database, h->reduced_map,
                     h->length, d_dist);
  GZ_HIP_CHECK(hipGetLastError());
  if (h_dist) GZ_HIP_CHECK(hipMemcpyAsync(h_dist, d_dist, size, hipMemcpyDefault, s));
  return GZ_OK;
}

// ---- UngappedExtenderGpu ----
struct gz_ungapped_extender_s {
  uint32_t delimiter = kDelimiter;
  const uint32_t *seed_query_ids = nullptr;
  const uint32_t *seed_query_positions = nullptr;
  const uint64_t *queries = nullptr;
  const uint64_t *database = nullptr;
  const int *cutoffs = nullptr;
  const int *triggers = nullptr;
  const int *matrix = nullptr;
  uint32_t number_letters = 0;
};

gz_status gz_ungapped_extender_create(gz_ungapped_extender_t *h) {
  GZ_REQUIRE(h != nullptr, "null handle pointer");
  *h = new gz_ungapped_extender_s();
  return GZ_OK;
}
gz_status gz_ungapped_extender_destroy(gz_ungapped_extender_t h) {
  delete h;
  return GZ_OK;
}
gz_status gz_ungapped_extender_set_query_seed_data_list(gz_ungapped_extender_t h,
                                                        const uint32_t *ids,
                                                        const uint32_t *pos) {
  GZ_REQUIRE(h, "null handle");
  h->seed_query_ids = ids;
  h->seed_query_positions = pos;
  return GZ_OK;
}
gz_status gz_ungapped_extender_set_queries(gz_ungapped_extender_t h,
                                           uint8_t delim, const uint64_t *q,
                                           const int *cutoffs,
                                           const int *triggers) {
  GZ_REQUIRE(h, "null handle");
  h->delimiter = delim;
  h->queries = q;
  h->cutoffs = cutoffs;
  h->triggers = triggers;
  return GZ_OK;
}
gz_status gz_ungapped_extender_set_database(gz_ungapped_extender_t h,
                                            const uint64_t *d) {
  GZ_REQUIRE(h, "null handle");
  h->database = d;
  return GZ_OK;
}
gz_status gz_ungapped_extender_set_score_matrix(gz_ungapped_extender_t h,
                                                const int *m, uint32_t nl) {
  GZ_REQUIRE(h, "null handle");
  GZ_REQUIRE(nl <= kTableDim, "score matrix larger than 32 letters");
  h->matrix = m;
  h->number_letters = nl;
  return GZ_OK;
}

size_t gz_ungapped_extender_query_seeds_temp_bytes(size_t size) {
  return DeviceScanTempBytes<uint32_t>(size);
}

gz_status gz_ungapped_extender_set_query_seeds(
    gz_ungapped_extender_t h, size_t runs, size_t size, const uint32_t *h_ids,
    const uint32_t *h_starts, size_t temp_bytes, void *d_temp,
    uint32_t *d_ids, uint32_t *d_starts, uint32_t *d_qids, uint32_t *d_qpos,
    uint32_t *d_tmp, gz_stream_t stream) {
  GZ_REQUIRE(h && h->seed_query_ids && h->seed_query_positions,
             "ungapped extender: query seed data list not set");
  if (size == 0) return GZ_OK;
  GZ_REQUIRE(temp_bytes >= DeviceScanTempBytes<uint32_t>(size),
             "ungapped extender: temporary storage too small");
  hipStream_t s = S(stream);
  if (h_ids) GZ_HIP_CHECK(hipMemcpyAsync(d_ids, h_ids, runs * 4, hipMemcpyDefault, s));
  if (h_starts) GZ_HIP_CHECK(hipMemcpyAsync(d_starts, h_starts, runs * 4, hipMemcpyDefault, s));
  GZ_HIP_CHECK(hipMemsetAsync(d_tmp, 0, size * 4, s));
  hipLaunchKernelGGL(SeedScatterKernel, dim3(GridFor(runs, 256, 1u << 30)), dim3(256), 0,
                     s, runs, size, d_starts, d_tmp);
  GZ_HIP_CHECK(hipGetLastError());
  GZ_HIP_CHECK((DeviceScan<uint32_t, true>(d_tmp, d_tmp, size, true, d_temp, s)));
  hipLaunchKernelGGL(SeedGatherKernel, dim3(GridFor(size, 256, 1u << 30)), dim3(256), 0,
                     s, size, d_tmp, d_ids, h->seed_query_ids,
                     h->seed_query_positions, d_qids, d_qpos);
  GZ_HIP_CHECK(hipGetLastError());
  return GZ_OK;
}

gz_status gz_ungapped_extender_extend_with_trigger_async(
    gz_ungapped_extender_t h, size_t size, const uint32_t *h_qids,
    const uint32_t *h_qpos, const uint32_t *h_dpos, char *h_flags,
    uint32_t *d_qids, uint32_t *d_qpos, uint32_t *d_dpos, char *d_flags,
    int *d_temp, gz_stream_t stream) {
  (void)d_temp;
  GZ_REQUIRE(h && h->queries && h->database && h->matrix && h->cutoffs &&
                 h->triggers,
             "ungapped extender: queries, database and score matrix must be set");
  if (size == 0) return GZ_OK;
  hipStream_t s = S(stream);
  if (h_qids) GZ_HIP_CHECK(hipMemcpyAsync(d_qids, h_qids, size * 4, hipMemcpyDefault, s));
  if (h_qpos) GZ_HIP_CHECK(hipMemcpyAsync(d_qpos, h_qpos, size * 4, hipMemcpyDefault, s));
  if (h_dpos) GZ_HIP_CHECK(hipMemcpyAsync(d_dpos, h_dpos, size * 4, hipMemcpyDefault, s));
  hipLaunchKernelGGL(UngappedTriggerKernel, dim3(GridFor(size, 256)), dim3(256), 0,
                     s, size, d_qids, d_qpos, d_dpos, h->queries, h->database,
                     h->matrix, h->number_letters, (uint32_t)h->delimiter,
                     h->cutoffs, h->triggers, d_flags);
  GZ_HIP_CHECK(hipGetLastError());
  if (h_flags) GZ_HIP_CHECK(hipMemcpyAsync(h_flags, d_flags, size, hipMemcpyDefault, s));
  return GZ_OK;
}

// ---- GappedExtenderGpu ----
// LDS rings: most bands fit the narrow one (more waves per CU); the rest
// move to the 64-column ring, then to the global-memory ring
constexpr int kGappedRing = 64;
// lanes of the global ring: the allocation-free entry point's fixed count,
// and the handle's, one per job up to what the longest query allows within
// kWideBudget bytes (a long band runs on one lane, so the bands that leave
// the LDS rings need many lanes: Pareto-long protein queries leave
// thousands of them per call)
constexpr uint32_t kWideSlots = 256;
constexpr uint32_t kWideSlotsMax = 16384;
constexpr size_t kWideBudget = (size_t)8 << 30;

inline uint32_t WideSlotsFor(uint32_t wide) {
  const size_t per = (size_t)(wide ? wide : 1) * (sizeof(int2) + 1);
  size_t n = kWideBudget / per;
  n = n < kWideSlots ? kWideSlots : (n > kWideSlotsMax ? kWideSlotsMax : n);
  return (uint32_t)(n / 64 * 64);
}

enum : int { kWaveAuto = 0, kWaveNever = 1, kWaveDirect = 2 };

struct gz_gapped_extender_s {
  uint32_t delimiter = kDelimiter;
  const uint64_t *queries = nullptr;
  const uint64_t *database = nullptr;
  const int *matrix = nullptr;
  uint32_t number_letters = 0;
  int gap_open = 0, gap_ext = 0, cutoff = 0;
  uint32_t wide = 1u << 16;  // wide-ring columns (>= longest query + 2)
  uint32_t narrow = 32;      // columns of the first LDS ring (16 or 32)
  int persistent = 0;        // first ring: one persistent launch fed from a work queue
  // bands past the 64-column ring: kWaveAuto (a wave per band when exact),
  // kWaveNever (a lane per band), kWaveDirect (as kWaveAuto, and the first
  // ring's overflow skips the 64-column ring); GZ_GAPPED_WAVE=never|direct
  int wave_mode = kWaveAuto;
  // workspace
  // [0] jobs out of the narrow ring, [1] failures, [2] jobs out of the
  // 64-column ring, [3] the work queue's head
  uint32_t *counters = nullptr;
  uint32_t *overflow_list = nullptr;   // out of the narrow ring
  uint32_t *overflow_list2 = nullptr;  // out of the 64-column ring
  size_t overflow_capacity = 0;
  int2 *wide_cells = nullptr;
  uint8_t *wide_codes = nullptr;
  uint32_t wide_allocated = 0;
  uint32_t wide_slots = 0;  // lanes of the allocated global ring
  // the failure counter of the last run (its own workspace's, or the
  // allocation-free entry point's scratch on the device it ran on)
  uint32_t *failure_counter = nullptr;
  ~gz_gapped_extender_s() {
    if (counters) (void)gz::Free(counters);
    if (overflow_list) (void)gz::Free(overflow_list);
    if (overflow_list2) (void)gz::Free(overflow_list2);
    if (wide_cells) (void)gz::Free(wide_cells);
    if (wide_codes) (void)gz::Free(wide_codes);
  }
};

gz_status gz_gapped_extender_create(gz_gapped_extender_t *h) {
  GZ_REQUIRE(h != nullptr, "null handle pointer");
  *h = new gz_gapped_extender_s();
  const char *m = getenv("GZ_GAPPED_WAVE");
  if (m && strcmp(m, "never") == 0) (*h)->wave_mode = kWaveNever;
  if (m && strcmp(m, "direct") == 0) (*h)->wave_mode = kWaveDirect;
  return GZ_OK;
}
gz_status gz_gapped_extender_destroy(gz_gapped_extender_t h) {
  delete h;
  return GZ_OK;
}
gz_status gz_gapped_extender_set_queries(gz_gapped_extender_t h, uint8_t delim,
                                         const uint64_t *q) {
  GZ_REQUIRE(h, "null handle");
  h->delimiter = delim;
  h->queries = q;
  return GZ_OK;
}
gz_status gz_gapped_extender_set_database(gz_gapped_extender_t h,
                                          const uint64_t *d) {
  GZ_REQUIRE(h, "null handle");
  h->database = d;
  return GZ_OK;
}
gz_status gz_gapped_extender_set_score_parameters(gz_gapped_extender_t h,
                                                  const int *m, uint32_t nl,
                                                  int gap_open, int gap_ext,
                                                  int cutoff) {
  GZ_REQUIRE(h, "null handle");
  GZ_REQUIRE(nl <= kTableDim, "score matrix larger than 32 letters");
  h->matrix = m;
  h->number_letters = nl;
  h->gap_open = gap_open;
  h->gap_ext = gap_ext;
  h->cutoff = cutoff;
  return GZ_OK;
}

gz_status gz_gapped_extender_convert_to_gapped_extension_seeds_async(
    gz_gapped_extender_t h, size_t size, int reverse, const uint32_t *h_seed,
    const uint32_t *h_q, const uint32_t *h_d, uint32_t *d_seed,
    uint32_t *d_temp, uint32_t *d_q, uint32_t *d_d, gz_stream_t stream) {
  (void)reverse;
  GZ_REQUIRE(h, "null handle");
  if (size == 0) return GZ_OK;
  hipStream_t s = S(stream);
  // d_temp holds the un-permuted query positions, then database positions
  uint32_t *tq = d_temp, *td = d_temp + size;
  if (h_seed) GZ_HIP_CHECK(hipMemcpyAsync(d_seed, h_seed, size * 4, hipMemcpyDefault, s));
  GZ_HIP_CHECK(hipMemcpyAsync(tq, h_q ? h_q : d_q, size * 4, hipMemcpyDefault, s));
  GZ_HIP_CHECK(hipMemcpyAsync(td, h_d ? h_d : d_d, size * 4, hipMemcpyDefault, s));
  hipLaunchKernelGGL(ConvertSeedsKernel, dim3(GridFor(size, 256, 1u << 30)), dim3(256), 0,
                     s, size, d_seed, tq, td, d_q, d_d);
  GZ_HIP_CHECK(hipGetLastError());
  return GZ_OK;
}

extern "C++" {
// Device memory one gapped run uses beyond its jobs: counters ([0] jobs out
// of the narrow ring, [1] failures, [2] jobs out of the 64-column ring), the
// two overflow lists (room for every job) and the global ring (kWideSlots
// slots of `wide` cells and codes).
struct GappedWork {
  uint32_t *counters;
  uint32_t *overflow_list, *overflow_list2;
  int2 *wide_cells;
  uint8_t *wide_codes;
  uint32_t wide;
  uint32_t wide_slots;  // lanes of the global ring (a multiple of 64)
};

template <bool kReverse>
static void GappedLaunch(gz_gapped_extender_t h, const GappedWork &w, const GappedParams &P,
                         size_t size, size_t wide_first, const uint32_t *order, uint32_t *d_q,
                         uint32_t *d_d, int *d_scores, uint32_t *d_cells, hipStream_t s) {
  uint32_t *narrow_out = w.counters, *failures = w.counters + 1, *wide_out = w.counters + 2;
  // 16-bit cells in the LDS rings when the cutoff and gap costs allow them
  // (LaneRingLds16); bands whose scores grow past them move on like bands
  // that outgrow the ring
  const bool short_cells = P.cutoff <= 8192 && P.gap_ext >= 0 &&
                           P.gap_open + P.gap_ext <= 1024;
  // jobs [0, wide_first) start on the 64-column ring, the rest on the
  // narrow one; the 64-column ring then takes the narrow ring's overflow
#define GZ_RING64(SHORT, GRID, BEGIN, END, LIST, COUNT, ORDER)                                 \
  hipLaunchKernelGGL((GappedKernel<kReverse, kGappedRing, SHORT>), GRID, dim3(kGappedBlock), 0, \
                     s, (size_t)(BEGIN), (size_t)(END), (const uint32_t *)(LIST),              \
                     (const uint32_t *)(COUNT), ORDER, size, P, h->matrix, h->number_letters,  \
                     d_q, d_d, d_scores, d_cells, wide_out, w.overflow_list2, failures)
  if (wide_first > 0) {
    const dim3 grid(GridFor(wide_first, kGappedBlock, 256u * 32));
    if (short_cells) GZ_RING64(true, grid, 0, wide_first, nullptr, nullptr, order);
    else GZ_RING64(false, grid, 0, wide_first, nullptr, nullptr, order);
  }
  if (size > wide_first && h->persistent && wide_first == 0) {
    // one persistent launch over every job, sized to the waves the device
    // holds at once
#define GZ_QUEUE(WIDTH, SHORT)                                                               \
  {                                                                                          \
    int per_cu = 1;                                                                          \
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(                                        \
            &per_cu, (const void *)GappedQueueKernel<kReverse, WIDTH, SHORT>, kGappedBlock,   \
            0) != hipSuccess || per_cu < 1)                                                  \
      per_cu = 1;                                                                            \
    const size_t waves = std::min<size_t>((size_t)per_cu * Cus(), (size - 1) / kWave + 1);   \
    hipLaunchKernelGGL((GappedQueueKernel<kReverse, WIDTH, SHORT>), dim3((unsigned)waves),    \
                       dim3(kGappedBlock), 0, s, order, (uint32_t)size, w.counters + 3, P,    \
                       h->matrix, h->number_letters, d_q, d_d, d_scores, d_cells, narrow_out, \
                       w.overflow_list, failures);                                           \
  }
    if (h->narrow == 16) {
      if (short_cells) GZ_QUEUE(16, true) else GZ_QUEUE(16, false)
    } else {
      if (short_cells) GZ_QUEUE(32, true) else GZ_QUEUE(32, false)
    }
#undef GZ_QUEUE
  } else if (size > wide_first) {
    const dim3 grid(GridFor(size - wide_first, kGappedBlock, 256u * 64));
#define GZ_NARROW(WIDTH, SHORT)                                                              \
  hipLaunchKernelGGL((GappedKernel<kReverse, WIDTH, SHORT>), grid, dim3(kGappedBlock), 0, s, \
                     wide_first, size, (const uint32_t *)nullptr, (const uint32_t *)nullptr, \
                     order, size, P, h->matrix, h->number_letters, d_q, d_d, d_scores,       \
                     d_cells, narrow_out, w.overflow_list, failures)
    if (h->narrow == 16) {
      if (short_cells) GZ_NARROW(16, true);
      else GZ_NARROW(16, false);
    } else {
      if (short_cells) GZ_NARROW(32, true);
      else GZ_NARROW(32, false);
    }
#undef GZ_NARROW
  }
  // bands past the 64-column ring: one wave per band when that is exact
  // (the cutoff below -kInitScore and non-negative gap costs), else one lane
  // per band; a wave takes a lane's slot of the global ring
  const bool wave = h->wave_mode != kWaveNever && P.cutoff < -kInitScore && P.gap_open >= 0 &&
                    P.gap_ext >= 0;
  const dim3 wave_grid(std::min<uint32_t>(w.wide_slots, 4096u));
  if (wave && h->wave_mode == kWaveDirect) {
    // the narrow ring's overflow straight to the waves
    hipLaunchKernelGGL((GappedWaveKernel<kReverse>), wave_grid, dim3(kWave), 0, s, P, h->matrix,
                       h->number_letters, d_q, d_d, d_scores, d_cells, narrow_out,
                       w.overflow_list, w.wide_cells, w.wide, failures);
  } else {
    // the narrow ring's overflow; 16-bit cells leave LDS for more blocks
    const uint32_t *none = nullptr;
    if (short_cells) GZ_RING64(true, dim3(256u * 6), 0, 0, w.overflow_list, narrow_out, none);
    else GZ_RING64(false, dim3(256u * 4), 0, 0, w.overflow_list, narrow_out, none);
  }
#undef GZ_RING64
  if (wave)
    hipLaunchKernelGGL((GappedWaveKernel<kReverse>), wave_grid, dim3(kWave), 0, s, P, h->matrix,
                       h->number_letters, d_q, d_d, d_scores, d_cells, wide_out,
                       w.overflow_list2, w.wide_cells, w.wide, failures);
  else
    hipLaunchKernelGGL((GappedWideKernel<kReverse>), dim3(w.wide_slots / kGappedBlock),
                       dim3(kGappedBlock), 0, s, P, h->matrix, h->number_letters, d_q, d_d,
                       d_scores, d_cells, wide_out, w.overflow_list2, w.wide_cells,
                       w.wide_codes, w.wide, failures);
}
}  // extern "C++"

static gz_status GappedLaunchAll(gz_gapped_extender_t h, const GappedWork &w, size_t size,
                                 int reverse, uint32_t *d_q, uint32_t *d_d, int *d_scores,
                                 uint32_t *d_cells, hipStream_t s, size_t wide_first,
                                 const uint32_t *order) {
  GappedParams P;
  P.queries = h->queries;
  P.database = h->database;
  P.delim = h->delimiter;
  P.gap_open = h->gap_open;
  P.gap_ext = h->gap_ext;
  P.cutoff = h->cutoff;
  if (reverse) GappedLaunch<true>(h, w, P, size, wide_first, order, d_q, d_d, d_scores, d_cells, s);
  else GappedLaunch<false>(h, w, P, size, wide_first, order, d_q, d_d, d_scores, d_cells, s);
  GZ_HIP_CHECK(hipGetLastError());
  return GZ_OK;
}

// The run counters a launch sequence starts from: the overflow counts and
// the queue head are zeroed, the failure count ([1]) is kept until read.  A
// kernel, not a memset, so that the sequence can be captured in a graph.
__global__ void ResetRunCountersKernel(uint32_t *counters, int all) {
  if (threadIdx.x < 4 && (all || threadIdx.x != 1)) counters[threadIdx.x] = 0;
}

// Allocates the handle's workspace for runs of up to `size` jobs (no-op
// when it is large enough).
static gz_status GappedReserve(gz_gapped_extender_t h, size_t size, hipStream_t s) {
  if (!h->counters) {
    GZ_HIP_CHECK(gz::Malloc(&h->counters, 4 * sizeof(uint32_t)));
    hipLaunchKernelGGL(ResetRunCountersKernel, dim3(1), dim3(64), 0, s, h->counters, 1);
    GZ_HIP_CHECK(hipGetLastError());
  }
  if (h->overflow_capacity < size) {
    if (h->overflow_list) GZ_HIP_CHECK(gz::Free(h->overflow_list));
    if (h->overflow_list2) GZ_HIP_CHECK(gz::Free(h->overflow_list2));
    h->overflow_list = h->overflow_list2 = nullptr;
    h->overflow_capacity = 0;
    size_t cap = std::max<size_t>(size, 1 << 16);
    GZ_HIP_CHECK(gz::Malloc(&h->overflow_list, cap * sizeof(uint32_t)));
    GZ_HIP_CHECK(gz::Malloc(&h->overflow_list2, cap * sizeof(uint32_t)));
    h->overflow_capacity = cap;
  }
  // lanes: one per job at most, at least kWideSlots, within the budget
  const size_t by_jobs = (size + 63) / 64 * 64;
  const uint32_t want = (uint32_t)std::max<size_t>(
      kWideSlots, std::min<size_t>(WideSlotsFor(h->wide), by_jobs));
  if (h->wide_allocated < h->wide || h->wide_slots < want) {
    if (h->wide_cells) GZ_HIP_CHECK(gz::Free(h->wide_cells));
    if (h->wide_codes) GZ_HIP_CHECK(gz::Free(h->wide_codes));
    h->wide_cells = nullptr;
    h->wide_codes = nullptr;
    h->wide_allocated = 0;
    h->wide_slots = 0;
    const uint32_t slots = want;
    GZ_HIP_CHECK(gz::Malloc(&h->wide_cells, (size_t)slots * h->wide * sizeof(int2)));
    GZ_HIP_CHECK(gz::Malloc(&h->wide_codes, (size_t)slots * h->wide));
    h->wide_allocated = h->wide;
    h->wide_slots = slots;
  }
  return GZ_OK;
}

static gz_status GappedRun(gz_gapped_extender_t h, size_t size, int reverse,
                           uint32_t *d_q, uint32_t *d_d, int *d_scores,
                           uint32_t *d_cells, hipStream_t s, size_t wide_first = 0,
                           const uint32_t *order = nullptr) {
  GZ_REQUIRE(h->queries && h->database && h->matrix,
             "gapped extender: queries, database and score parameters must be set");
  GZ_REQUIRE(wide_first <= size, "gapped extender: wide_first beyond the jobs");
  GZ_REQUIRE(size < (1ull << 31), "gapped extender: at most 2^31 - 1 jobs per call");
  const gz_status rs = GappedReserve(h, size, s);
  if (rs != GZ_OK) return rs;
  // the overflow counts of this run start at zero; failures (counters[1])
  // are kept until gz_gapped_extender_failures reads them
  hipLaunchKernelGGL(ResetRunCountersKernel, dim3(1), dim3(64), 0, s, h->counters, 0);
  GZ_HIP_CHECK(hipGetLastError());
  const GappedWork w = {h->counters,   h->overflow_list, h->overflow_list2, h->wide_cells,
                        h->wide_codes, h->wide,          h->wide_slots};
  h->failure_counter = h->counters + 1;
  return GappedLaunchAll(h, w, size, reverse, d_q, d_d, d_scores, d_cells, s, wide_first, order);
}

gz_status gz_gapped_extender_extend_one_side_score_only_async(
    gz_gapped_extender_t h, size_t size, int reverse, uint32_t *h_q,
    uint32_t *h_d, int *h_scores, uint32_t *d_q, uint32_t *d_d, int *d_scores,
    gz_stream_t stream) {
  GZ_REQUIRE(h, "null handle");
  if (size == 0) return GZ_OK;
  hipStream_t s = S(stream);
  if (h_q) GZ_HIP_CHECK(hipMemcpyAsync(d_q, h_q, size * 4, hipMemcpyDefault, s));
  if (h_d) GZ_HIP_CHECK(hipMemcpyAsync(d_d, h_d, size * 4, hipMemcpyDefault, s));
  gz_status st = GappedRun(h, size, reverse, d_q, d_d, d_scores, nullptr, s);
  if (st != GZ_OK) return st;
  if (h_q) GZ_HIP_CHECK(hipMemcpyAsync(h_q, d_q, size * 4, hipMemcpyDefault, s));
  if (h_d) GZ_HIP_CHECK(hipMemcpyAsync(h_d, d_d, size * 4, hipMemcpyDefault, s));
  if (h_scores) GZ_HIP_CHECK(hipMemcpyAsync(h_scores, d_scores, size * 4, hipMemcpyDefault, s));
  return GZ_OK;
}

gz_status gz_gapped_extender_extend_one_side_score_only_cells_async(
    gz_gapped_extender_t h, size_t size, int reverse, uint32_t *d_q,
    uint32_t *d_d, int *d_scores, uint32_t *d_cells, gz_stream_t stream) {
  GZ_REQUIRE(h, "null handle");
  if (size == 0) return GZ_OK;
  return GappedRun(h, size, reverse, d_q, d_d, d_scores, d_cells, S(stream));
}

gz_status gz_gapped_extender_extend_one_side_score_only_binned_async(
    gz_gapped_extender_t h, size_t size, size_t wide_first, int reverse, uint32_t *d_q,
    uint32_t *d_d, int *d_scores, uint32_t *d_cells, gz_stream_t stream) {
  GZ_REQUIRE(h, "null handle");
  if (size == 0) return GZ_OK;
  return GappedRun(h, size, reverse, d_q, d_d, d_scores, d_cells, S(stream), wide_first);
}

gz_status gz_gapped_extender_extend_one_side_score_only_ordered_async(
    gz_gapped_extender_t h, size_t size, int reverse, const uint32_t *d_order, uint32_t *d_q,
    uint32_t *d_d, int *d_scores, uint32_t *d_cells, gz_stream_t stream) {
  GZ_REQUIRE(h && d_order, "null argument");
  if (size == 0) return GZ_OK;
  return GappedRun(h, size, reverse, d_q, d_d, d_scores, d_cells, S(stream), 0, d_order);
}

}  // extern "C"

// ---- the allocation-free entry point ----
// Its workspace is one file-scope device array per GPU: the three counters,
// the two overflow lists for up to kNoAllocMaxJobs jobs and the global ring
// for queries of up to kNoAllocMaxWide - 2 residues.
constexpr size_t kNoAllocMaxJobs = (size_t)1 << 20;
constexpr uint32_t kNoAllocMaxWide = 4096;
constexpr size_t kNoAllocCountersBytes = 256;
constexpr size_t kNoAllocListBytes = kNoAllocMaxJobs * sizeof(uint32_t);
constexpr size_t kNoAllocCellsBytes = (size_t)kWideSlots * kNoAllocMaxWide * sizeof(int2);
constexpr size_t kNoAllocCodesBytes = (size_t)kWideSlots * kNoAllocMaxWide;
constexpr size_t kNoAllocScratchBytes =
    kNoAllocCountersBytes + 2 * kNoAllocListBytes + kNoAllocCellsBytes + kNoAllocCodesBytes;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mutex;
void *g_scratch[kNoAllocMaxDevices] = {};

// The current device's copy of `scratch`, looked up once per device with
// that device current (host code must not pass the array's name itself).
gz_status ScratchOfCurrentDevice(uint8_t **out) {
  int dev = 0;
  GZ_HIP_CHECK(hipGetDevice(&dev));
  GZ_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "no-alloc gapped extension: device id");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    GZ_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    GZ_REQUIRE(p != nullptr, "no-alloc gapped extension: no scratch address");
    g_scratch[dev] = p;
  }
  *out = (uint8_t *)g_scratch[dev];
  return GZ_OK;
}

__global__ void ClearCountersKernel(uint32_t *counters) {
  if (threadIdx.x < 4) counters[threadIdx.x] = 0;
}

}  // namespace

extern "C" {

gz_status gz_gapped_extender_extend_one_side_score_only_noalloc_async(
    gz_gapped_extender_t h, size_t size, int reverse, uint32_t *d_q, uint32_t *d_d,
    int *d_scores, uint32_t *d_cells, gz_stream_t stream) {
  GZ_REQUIRE(h, "null handle");
  GZ_REQUIRE(h->queries && h->database && h->matrix,
             "gapped extender: queries, database and score parameters must be set");
  if (size > kNoAllocMaxJobs) {
    SetError("no-alloc gapped extension: more than 1048576 jobs in one call");
    return GZ_ERROR_CAPACITY;
  }
  if (h->wide > kNoAllocMaxWide) {
    SetError("no-alloc gapped extension: the maximum sequence length is above 4094 residues");
    return GZ_ERROR_CAPACITY;
  }
  if (size == 0) return GZ_OK;
  GZ_REQUIRE(d_q && d_d && d_scores, "null argument");
  uint8_t *base = nullptr;
  const gz_status st = ScratchOfCurrentDevice(&base);
  if (st != GZ_OK) return st;
  hipStream_t s = S(stream);
  GappedWork w;
  w.counters = (uint32_t *)base;
  w.overflow_list = (uint32_t *)(base + kNoAllocCountersBytes);
  w.overflow_list2 = (uint32_t *)(base + kNoAllocCountersBytes + kNoAllocListBytes);
  w.wide_cells = (int2 *)(base + kNoAllocCountersBytes + 2 * kNoAllocListBytes);
  w.wide_codes = base + kNoAllocCountersBytes + 2 * kNoAllocListBytes + kNoAllocCellsBytes;
  w.wide = h->wide;
  w.wide_slots = kWideSlots;
  // this call's counts and failures start at zero (a fill kernel, no memset
  // node, so that the call can be captured into a graph)
  hipLaunchKernelGGL(ClearCountersKernel, dim3(1), dim3(64), 0, s, w.counters);
  GZ_HIP_CHECK(hipGetLastError());
  h->failure_counter = w.counters + 1;
  return GappedLaunchAll(h, w, size, reverse, d_q, d_d, d_scores, d_cells, s, 0, nullptr);
}

gz_status gz_gapped_extender_set_max_sequence_length(gz_gapped_extender_t h,
                                                     uint32_t length) {
  GZ_REQUIRE(h, "null handle");
  uint32_t w = 64;
  while (w < length + 2) w <<= 1;
  h->wide = w;
  return GZ_OK;
}

gz_status gz_gapped_extender_reserve(gz_gapped_extender_t h, size_t size, gz_stream_t stream) {
  GZ_REQUIRE(h, "null handle");
  return GappedReserve(h, size, S(stream));
}

gz_status gz_gapped_extender_set_persistent(gz_gapped_extender_t h, int persistent) {
  GZ_REQUIRE(h, "null handle");
  h->persistent = persistent ? 1 : 0;
  return GZ_OK;
}

gz_status gz_gapped_extender_set_narrow_ring(gz_gapped_extender_t h, uint32_t columns) {
  GZ_REQUIRE(h, "null handle");
  GZ_REQUIRE(columns == 16 || columns == 32, "narrow ring: 16 or 32 columns");
  h->narrow = columns;
  return GZ_OK;
}

gz_status gz_gapped_extender_overflows(gz_gapped_extender_t h, uint32_t *narrow,
                                       uint32_t *ring64) {
  GZ_REQUIRE(h && narrow && ring64, "null argument");
  *narrow = *ring64 = 0;
  if (h->counters) {
    uint32_t c[3];
    GZ_HIP_CHECK(hipMemcpy(c, h->counters, sizeof(c), hipMemcpyDeviceToHost));
    *narrow = c[0];
    *ring64 = c[2];
  }
  return GZ_OK;
}

gz_status gz_gapped_extender_failures(gz_gapped_extender_t h, uint32_t *failures) {
  GZ_REQUIRE(h && failures, "null argument");
  *failures = 0;
  if (h->failure_counter) {
    uint32_t c = 0;
    GZ_HIP_CHECK(hipMemcpy(&c, h->failure_counter, sizeof(c), hipMemcpyDeviceToHost));
    *failures = c;
    if (c != 0) GZ_HIP_CHECK(hipMemset(h->failure_counter, 0, sizeof(uint32_t)));
  }
  return GZ_OK;
}

}  // extern "C"
