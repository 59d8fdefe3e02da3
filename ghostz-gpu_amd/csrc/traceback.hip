// traceback.hip -- the result build (AlignerBuildResultsThread) on the device.
//
// For every presearched result the reference re-extends both sides with the
// result cutoff and a traceback (GappedExtender::ExtendOneSide with
// EditBlocks, src/gapped_extender.cpp:70-304,473-531), then counts length,
// mismatches and gap openings / extensions over the joined edit script
// (AlignerCommon::BuildResult, src/aligner_common.h:276-343).  One lane runs
// one result.  The op store follows the reference's gap_states_ layout: rows
// are laid back to back in one byte arena ([0] holds the first op, row r
// starts where row r-1 ends), so the traceback's reads outside a row's band
// see the same neighbouring bytes.  Cells an extension does not write hold a
// marker; reading one (or reading outside the written range) means the
// reference would have read a byte left by an earlier extension, and the
// result is flagged so the host replays the chunk in the reference order.
// Results whose scratch needs exceed the per-lane capacities are flagged for
// the host as well.
#include <hip/hip_runtime.h>

#include <climits>

#include "gz_common.h"
#include "gz_device.h"
#include "traceback.h"
#include "ungapped.h"

namespace gz {

namespace {

constexpr uint8_t kSub = 0, kGap0 = 1, kGap1 = 2, kExt0 = 4, kExt1 = 8;
constexpr uint8_t kEditGap0 = 0, kEditSub = 1, kEditGap1 = 2;
constexpr uint8_t kUnwritten = 0xFF;
constexpr int kTbThreads = 64;
// Query sides of up to kTbLdsCols - 2 residues keep their DP columns in LDS,
// column-major over the block's lanes ([column][lane]): 16-bit cells (16 KB
// per block) when the cutoff and gap costs allow them, else 32-bit ones.
constexpr uint32_t kTbLdsCols = 64;
constexpr int kTbLow = -16384;

__shared__ int2 tb_dp[kTbLdsCols * kTbThreads];
// 16-bit columns.  Exact when the cutoff is at most 8192 and gap open +
// extension at most 1024 (the launch checks): a side of at most 62 residues
// scores at most 62 x 127, and values stored below -16384 (the band's first
// gap column keeps decreasing) are clamped there, which changes no comparison
// and no recorded op, since every value compared with them is above
// -16384 + the gap extension.
__shared__ short2 tb_dp16[kTbLdsCols * kTbThreads];
// The query side's codes by DP column: column a (a >= 1) holds the code at
// q0 + (a - 1) inc, the delimiter closing the side included.
__shared__ uint8_t tb_qc[kTbLdsCols * kTbThreads];

// DP column access: the lane's LDS column set (32- or 16-bit cells), or its
// global scratch.
template <bool kLds, bool kShort>
struct DpCells {
  int2 *global;
  __device__ __forceinline__ int2 Get(int i) const {
    if constexpr (kLds && kShort) {
      const short2 v = tb_dp16[i * kTbThreads + (int)threadIdx.x];
      return make_int2(v.x, v.y);
    } else if constexpr (kLds) {
      return tb_dp[i * kTbThreads + (int)threadIdx.x];
    } else {
      return global[i];
    }
  }
  __device__ __forceinline__ void Set(int i, int2 v) const {
    if constexpr (kLds && kShort) {
      tb_dp16[i * kTbThreads + (int)threadIdx.x] =
          make_short2((short)(v.x > kTbLow ? v.x : kTbLow), (short)(v.y > kTbLow ? v.y : kTbLow));
    } else if constexpr (kLds) {
      tb_dp[i * kTbThreads + (int)threadIdx.x] = v;
    } else {
      global[i] = v;
    }
  }
};

struct Lane {
  int2 *dp;
  uint8_t *arena;
  int2 *rows;  // (base, off)
  uint8_t *ops;
  uint32_t dp_cap, arena_cap, row_cap, op_cap;
};

struct Side {
  int score, p0, p1;
  uint32_t nops;  // ops generated (traceback order), in ops[base ..)
};

__device__ __forceinline__ uint32_t CodeAt(const uint64_t *packed, int64_t pos) {
  return PackedCodeAt(packed, (uint32_t)pos);
}

// GappedExtender::ExtendOneSide with traceback (the DP and band bookkeeping
// after the delimiter checks); ops are generated from the best cell back to
// the origin.  Returns flags (kTracebackStale/Overflow).
template <bool kLds, bool kShort>
__device__ __forceinline__ uint32_t ExtendBody(const uint64_t *pq, int64_t q0, uint32_t s0len,
                                               const uint64_t *pd, int64_t d0, int inc,
                                               const int8_t *sm, int gap_open, int gap_ext,
                                               int cutoff, Lane &L, uint32_t op_base,
                                               Side *out) {
  const DpCells<kLds, kShort> dp{L.dp};
  // the query code at q0 + p, p = (a - 1) inc for DP column a
  auto qcode = [&](int p) -> uint32_t {
    if constexpr (kLds) return tb_qc[(p * inc + 1) * kTbThreads + (int)threadIdx.x];
    else return CodeAt(pq, q0 + p);
  };
  const uint32_t extra =
      (gap_ext > 0 && cutoff != INT_MAX) ? (uint32_t)(cutoff / gap_ext + 3) : s0len + 3;
  uint32_t used = 0;
  int64_t written = 0;  // highest arena index written
  {
    const uint32_t l = extra + extra / 3;
    if ((uint64_t)used + l >= L.arena_cap) return kTracebackOverflow;
  }
  L.arena[0] = kSub;
  const int gi = gap_open + gap_ext;
  int as = 0, ae = 0, ai = 0, last = 0;
  int score = -gi;
  dp.Set(0, make_int2(0, -gi));
  int p0 = 0;
  for (ai = 1; qcode(p0) != kDelimiter; ++ai, p0 += inc) {
    if (score < -cutoff) break;
    dp.Set(ai, make_int2(score, score - gi));
    score -= gap_ext;
  }
  if (qcode(p0) != kDelimiter) {
    dp.Set(ai, make_int2(kInitScore, kInitScore));
    ae = ai + 1;
  } else {
    ae = ai;
  }
  int max_score = 0, mp0 = -inc, mp1 = -inc;
  uint32_t nrows = 0;
  uint32_t c1_next = CodeAt(pd, d0);
  for (int p1 = 0;; p1 += inc) {
    const uint32_t c1 = c1_next;
    if (c1 == kDelimiter) break;
    // the next row's code, read ahead (in bounds: c1 is a residue)
    c1_next = CodeAt(pd, d0 + p1 + inc);
    const uint32_t length = gap_ext > 0 ? (uint32_t)(ae - as) + extra : extra - (uint32_t)as;
    const uint32_t l = length + length / 3;
    if ((uint64_t)used + l >= L.arena_cap || nrows >= L.row_cap) return kTracebackOverflow;
    const int base = (int)used + 1;
    const int off = as;
    L.rows[nrows++] = make_int2(base, off);
    const int zero = base - off;
    const int8_t *row = sm + c1 * kTableDim;
    p0 = as * inc;
    int prev = dp.Get(as).x;
    {
      int2 c = dp.Get(as);
      c.x = (max_score - c.y > cutoff) ? kInitScore : c.y;
      dp.Set(as, make_int2(c.x, c.y - gap_ext));
    }
    int gap_row = dp.Get(as).x - gi;
    L.arena[zero + as] = kGap0 + kExt0;
    written = max(written, (int64_t)(zero + as));
    last = as;
    for (ai = as + 1; ai < ae; ++ai, p0 += inc) {
      int s = prev + row[qcode(p0)];
      const int2 cell = dp.Get(ai);
      prev = cell.x;
      int gap_col = cell.y;
      uint8_t op = kSub;
      if (s < gap_col) {
        op = kGap0;
        s = gap_col;
      }
      if (s < gap_row) {
        op = kGap1;
        s = gap_row;
      }
      if (max_score - s > cutoff) {
        if (as + 1 == ai) ++as;
        dp.Set(ai, make_int2(kInitScore, kInitScore));
        gap_row = kInitScore;
        L.arena[zero + ai] = kUnwritten;
      } else {
        last = ai;
        if (s > max_score) {
          max_score = s;
          mp0 = p0;
          mp1 = p1;
        }
        gap_col -= gap_ext;
        gap_row -= gap_ext;
        int gap_keep;
        if (gap_col < s - gi) {
          gap_keep = s - gi;
        } else {
          gap_keep = gap_col;
          op += kExt0;
        }
        if (gap_row < s - gi) gap_row = s - gi; else op += kExt1;
        L.arena[zero + ai] = op;
        dp.Set(ai, make_int2(s, gap_keep));
      }
      written = max(written, (int64_t)(zero + ai));
    }
    if (as + 1 == ae) break;
    if (last < ae - 1) {
      ae = last + 1;
    } else {
      while (gap_row >= (max_score - cutoff) && qcode(p0) != kDelimiter) {
        dp.Set(ae, make_int2(gap_row, gap_row - gi));
        L.arena[zero + ae] = kGap0 + kExt0;
        written = max(written, (int64_t)(zero + ae));
        gap_row -= gap_ext;
        ++ae;
        p0 += inc;
      }
    }
    if (qcode(p0) != kDelimiter) {
      dp.Set(ae, make_int2(kInitScore, kInitScore));
      L.arena[zero + ae] = kUnwritten;
      written = max(written, (int64_t)(zero + ae));
      ++ae;
    }
    used += (uint32_t)(ae - off);
  }
  out->score = max_score;
  out->p0 = mp0;
  out->p1 = mp1;
  // PickOptimalPath: generated from the best cell back to the origin
  uint32_t flags = 0;
  uint32_t n = op_base;
  int q = mp0, d = mp1;
  uint8_t op = kSub;
  auto emit = [&](uint8_t e, int count) {
    for (int k = 0; k < count; ++k) {
      if (n >= L.op_cap) {
        flags |= kTracebackOverflow;
        return;
      }
      L.ops[n++] = e;
    }
  };
  while (d * inc >= 0 && q * inc >= 0) {
    const int r = d * inc;
    const int2 rb = L.rows[r];
    const int64_t idx = (int64_t)rb.x + q * inc - rb.y + 1;
    uint8_t next = 0;
    if (idx < 0 || idx > written) {
      flags |= kTracebackStale;
    } else {
      next = L.arena[idx];
      if (next == kUnwritten) {
        flags |= kTracebackStale;
        next = 0;
      }
    }
    if (op == kGap0) {
      op = next & 3;
      if (next & kExt0) op = kGap0;
    } else if (op == kGap1) {
      op = next & 3;
      if (next & kExt1) op = kGap1;
    } else {
      op = next & 3;
    }
    if (op == kGap0) {
      emit(kEditGap0, 1);
      d -= inc;
    } else if (op == kGap1) {
      emit(kEditGap1, 1);
      q -= inc;
    } else {
      emit(kEditSub, 1);
      d -= inc;
      q -= inc;
    }
    if (flags & kTracebackOverflow) return flags;
  }
  if (d * inc >= 0) emit(kEditGap0, d * inc + 1);
  else if (q * inc >= 0) emit(kEditGap1, q * inc + 1);
  out->nops = n - op_base;
  return flags;
}

template <bool kShort>
__device__ uint32_t Extend(const uint64_t *pq, int64_t q0, uint32_t s0len, const uint64_t *pd,
                           int64_t d0, bool reversed, const int8_t *sm, int gap_open,
                           int gap_ext, int cutoff, Lane &L, uint32_t op_base, Side *out) {
  const int inc = reversed ? -1 : 1;
  out->score = 0;
  out->p0 = -inc;
  out->p1 = -inc;
  out->nops = 0;
  if (CodeAt(pq, q0) == kDelimiter || CodeAt(pd, d0) == kDelimiter) return 0;
  // n0 = residues before the delimiter, twelve codes per packed word pair;
  // the codes go to the side's LDS columns on the way.  A window starts at a
  // residue, so its words lie inside the padded array.
  uint32_t n0 = 0;
  bool found = false;
  for (uint32_t base = 0; !found; base += kCodesPerWord) {
    const uint32_t first =
        (uint32_t)(reversed ? q0 - (int64_t)base - (int64_t)(kCodesPerWord - 1)
                            : q0 + (int64_t)base) + kPackPadCodes;
    const uint32_t w = first / kCodesPerWord;
    const uint64_t win = Window12(pq[w] & kCodeBitsMask, pq[w + 1] & kCodeBitsMask,
                                  first - w * kCodesPerWord);
#pragma unroll
    for (int j = 0; j < (int)kCodesPerWord; ++j) {
      const uint32_t code =
          (uint32_t)(win >> (kCodeBits * (reversed ? (int)kCodesPerWord - 1 - j : j))) & 31u;
      if (!found) {
        const uint32_t a = base + (uint32_t)j + 1;
        if (a < kTbLdsCols) tb_qc[a * kTbThreads + threadIdx.x] = (uint8_t)code;
        if (code == kDelimiter) {
          found = true;
          n0 = base + (uint32_t)j;
        }
      }
    }
  }
  // columns 0 .. n0 + 1 are touched (the band may step one past the last
  // residue); s0len, the distance to the query's far end, can exceed n0 when
  // the query holds several frames
  if (n0 + 2 <= kTbLdsCols)
    return ExtendBody<true, kShort>(pq, q0, s0len, pd, d0, inc, sm, gap_open, gap_ext, cutoff,
                                    L, op_base, out);
  if (max(s0len + 1, n0 + 2) > L.dp_cap) return kTracebackOverflow;
  return ExtendBody<false, false>(pq, q0, s0len, pd, d0, inc, sm, gap_open, gap_ext, cutoff,
                                  L, op_base, out);
}

template <bool kShort>
__global__ void __launch_bounds__(kTbThreads)
    TracebackKernel(const uint64_t *pq, const uint64_t *pd, const TracebackJob *jobs,
                    uint32_t n, const int *matrix, uint32_t nl, int gap_open, int gap_ext,
                    int cutoff, uint8_t *scratch, LaneCaps caps, TracebackOut *outs) {
  __shared__ int8_t sm[kTableDim * kTableDim];
  StageScoreTable(sm, matrix, nl);
  __syncthreads();
  const uint32_t lane_id = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lanes = gridDim.x * blockDim.x;
  Lane L;
  uint8_t *mine = scratch + (size_t)lane_id * caps.bytes_per_lane;
  L.dp = reinterpret_cast<int2 *>(mine);
  L.rows = reinterpret_cast<int2 *>(mine + caps.dp * sizeof(int2));
  L.arena = mine + (caps.dp + caps.rows) * sizeof(int2);
  L.ops = L.arena + caps.arena;
  L.dp_cap = caps.dp;
  L.row_cap = caps.rows;
  L.arena_cap = caps.arena;
  L.op_cap = caps.ops;
  for (uint32_t j = lane_id; j < n; j += lanes) {
    const TracebackJob job = jobs[j];
    TracebackOut o = {};
    Side rev, fwd;
    // reverse side from (hit - 1), forward side from the hit
    uint32_t flags = Extend<kShort>(pq, (int64_t)job.query_offset + job.hit_q - 1, job.hit_q - 1, pd,
                            (int64_t)job.hit_d - 1, true, sm, gap_open, gap_ext, cutoff, L, 0,
                            &rev);
    if (!(flags & kTracebackOverflow))
      flags |= Extend<kShort>(pq, (int64_t)job.query_offset + job.hit_q, job.query_length - job.hit_q,
                      pd, (int64_t)job.hit_d, false, sm, gap_open, gap_ext, cutoff, L, rev.nops,
                      &fwd);
    o.flags = flags;
    if (!(flags & kTracebackOverflow)) {
      o.score = rev.score + fwd.score;
      o.start_q = job.hit_q - 1 + rev.p0;
      o.start_d = job.hit_d - 1 + rev.p1;
      o.end_q = job.hit_q + fwd.p0;
      o.end_d = job.hit_d + fwd.p1;
      // the joined script: both generated sequences reversed, reverse side first
      uint32_t qp = o.start_q, dp = o.start_d, m = 0, op_open = 0, ext = 0, len = 0;
      uint8_t prev = kEditSub;
      for (int part = 0; part < 2; ++part) {
        const uint32_t b = part == 0 ? 0 : rev.nops;
        const uint32_t cnt = part == 0 ? rev.nops : fwd.nops;
        for (uint32_t k = cnt; k-- > 0;) {
          const uint8_t e = L.ops[b + k];
          if (e == kEditSub) {
            if (CodeAt(pq, (int64_t)job.query_offset + qp) != CodeAt(pd, dp)) ++m;
            ++qp;
            ++dp;
          } else if (e == kEditGap0) {
            if (prev != kEditGap0) ++op_open; else ++ext;
            ++dp;
          } else {
            if (prev != kEditGap1) ++op_open; else ++ext;
            ++qp;
          }
          prev = e;
          ++len;
        }
      }
      o.length = len;
      o.mismatches = m;
      o.gap_openings = op_open;
      o.gap_extensions = ext;
    }
    outs[j] = o;
  }
}

}  // namespace

LaneCaps TracebackCaps(uint32_t max_query_length, uint32_t max_subject_length,
                       uint32_t arena) {
  LaneCaps c;
  c.dp = max_query_length + 4;
  c.rows = max_subject_length + 4;
  c.arena = arena;
  c.ops = 2 * (max_query_length + max_subject_length) + 16;
  size_t b = (size_t)(c.dp + c.rows) * sizeof(int2) + c.arena + c.ops;
  c.bytes_per_lane = (b + 255) / 256 * 256;
  return c;
}

hipError_t TracebackLaunch(const uint64_t *pq, const uint64_t *pd, const TracebackJob *jobs,
                           uint32_t n, const int *matrix, uint32_t nl, int gap_open,
                           int gap_ext, int cutoff, uint8_t *scratch, const LaneCaps &caps,
                           uint32_t lanes, TracebackOut *outs, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = (lanes + kTbThreads - 1) / kTbThreads;
  const bool short_cells = cutoff <= 8192 && gap_ext >= 0 && gap_open + gap_ext <= 1024;
  if (short_cells)
    hipLaunchKernelGGL(TracebackKernel<true>, dim3(blocks), dim3(kTbThreads), 0, s, pq, pd, jobs,
                       n, matrix, nl, gap_open, gap_ext, cutoff, scratch, caps, outs);
  else
    hipLaunchKernelGGL(TracebackKernel<false>, dim3(blocks), dim3(kTbThreads), 0, s, pq, pd,
                       jobs, n, matrix, nl, gap_open, gap_ext, cutoff, scratch, caps, outs);
  return hipGetLastError();
}

}  // namespace gz
