// search_device.hip -- see search_device.h.
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cstring>
#include <vector>

#include "gz_common.h"
#include "gz_runtime.h"
#include "gz_device.h"
#include "gz_scan.h"
#include "search_device.h"
#include "ungapped.h"

namespace gz {

namespace {

constexpr int kSeedThreads = 256;
constexpr int kSeedItems = 2;
constexpr uint32_t kSeedTile = kSeedThreads * kSeedItems;
constexpr uint64_t kUnitsPerLaunch = 1ull << 28;
constexpr size_t kInitialHitCapacity = 1u << 25;

// The hit buffer's first size: kInitialHitCapacity, or GZ_SEED_HIT_CAPACITY
// seeds when that variable is set (a small value makes every launch take the
// grow-and-redo path, which the tests use).
size_t InitialHitCapacity() {
  const char *v = getenv("GZ_SEED_HIT_CAPACITY");
  if (v && *v) {
    const unsigned long long n = strtoull(v, nullptr, 10);
    if (n >= 64) return (size_t)n;
  }
  return kInitialHitCapacity;
}

// per list: database representatives and values with the list's hash, and
// the list's entries that take part in this run
struct ListInfo {
  uint32_t rep_begin, nrep, val_begin, nval, ent_begin, nent;
};

// counters: [0] hits, [1] distance pairs, [2] ungapped extensions
constexpr int kCounters = 4;

__global__ void LookupKernel(DeviceQueryLists q, DeviceDbChunk db, int clustering,
                             ListInfo *info) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= q.nlists) return;
  const uint32_t h = q.hash[l];
  uint32_t lo = 0, hi = db.nhash;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (db.hashes[mid] < h) lo = mid + 1; else hi = mid;
  }
  ListInfo li = {0, 0, 0, 0, 0, 0};
  if (lo < db.nhash && db.hashes[lo] == h) {
    if (clustering) {
      li.rep_begin = db.rep_off[lo];
      li.nrep = db.rep_off[lo + 1] - li.rep_begin;
    }
    li.val_begin = db.val_off[lo];
    li.nval = db.val_off[lo + 1] - li.val_begin;
  }
  info[l] = li;
}

// Entries of queries [qb, qe) in every list and the list's units.  Entries
// are in appearance order, which visits queries in blocks of 64, so with qb
// and qe multiples of 64 (or qe past the last query) they are one run.
__global__ void BlockUnitsKernel(DeviceQueryLists q, const ListInfo *base, uint32_t qb,
                                 uint32_t qe, ListInfo *info, uint64_t *units_rep,
                                 uint64_t *units_val) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= q.nlists) return;
  ListInfo li = base[l];
  const uint32_t eb = q.entry_off[l], ee = q.entry_off[l + 1];
  uint32_t lo = eb, hi = ee;
  if (qb > 0 || qe < 0xffffffffu) {
    const uint32_t gb = qb >> 6, ge = qe == 0xffffffffu ? 0xffffffffu : qe >> 6;
    uint32_t a = eb, b = ee;  // first entry with query block >= gb
    while (a < b) {
      const uint32_t mid = (a + b) >> 1;
      if ((q.entry_qid[mid] >> 6) < gb) a = mid + 1; else b = mid;
    }
    lo = a;
    b = ee;  // first entry with query block >= ge
    while (a < b) {
      const uint32_t mid = (a + b) >> 1;
      if ((q.entry_qid[mid] >> 6) < ge) a = mid + 1; else b = mid;
    }
    hi = a;
  }
  li.ent_begin = lo;
  li.nent = hi - lo;
  info[l] = li;
  units_rep[l] = (uint64_t)li.nrep * li.nent;
  units_val[l] = (uint64_t)li.nval * li.nent;
}

__global__ void ActiveFlagsKernel(uint32_t n, const uint64_t *units, uint32_t *flags) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < n) flags[l] = units[l] > 0 ? 1u : 0u;
}

__global__ void ActiveScatterKernel(uint32_t n, const uint64_t *units,
                                    const uint64_t *offsets, const uint32_t *slot,
                                    uint32_t *active, uint64_t *active_off) {
  const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n) return;
  if (units[l] > 0) {
    active[slot[l]] = l;
    active_off[slot[l]] = offsets[l];
  }
  if (l == n - 1) {
    // sentinel: total units after the last active list
    const uint32_t count = slot[l] + (units[l] > 0 ? 1u : 0u);
    active_off[count] = offsets[l] + units[l];
  }
}

// Hits staged per wave in LDS and written in batches: one atomic on the
// shared hit counter per batch of up to kWaveStage hits instead of one per
// append.  `count` is wave-uniform; appends and flushes are wave-wide calls.
constexpr uint32_t kWaveStage = 64;

struct HitStage {
  SeedHit *buf;  // this wave's kWaveStage slots
  uint32_t count;
};

__device__ __forceinline__ void StageFlush(HitStage &st, SeedHit *out, uint32_t *counter,
                                           uint32_t capacity) {
  if (st.count == 0) return;
  const uint32_t lane = (uint32_t)__lane_id();
  uint32_t base = 0;
  if (lane == 0) base = atomicAdd(counter, st.count);
  base = __shfl(base, 0, kWave);
  if (lane < st.count && base + lane < capacity) out[base + lane] = st.buf[lane];
  st.count = 0;
}

__device__ __forceinline__ void StageAppend(bool pred, const SeedHit &h, HitStage &st,
                                            SeedHit *out, uint32_t *counter,
                                            uint32_t capacity) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return;
  const uint32_t n = (uint32_t)__popcll(mask);
  if (st.count + n > kWaveStage) StageFlush(st, out, counter, capacity);
  if (pred) {
    const int lane = __lane_id();
    st.buf[st.count + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull))] = h;
  }
  st.count += n;
}


// first and last active list of every tile of kSeedTile units starting at
// unit_begin (upper bounds over the active lists' unit offsets)
__global__ void TileRangeKernel(const uint64_t *active_off, uint32_t nactive,
                                uint64_t unit_begin, uint64_t unit_end, uint32_t tiles,
                                uint2 *range) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= tiles) return;
  const uint64_t tile = unit_begin + (uint64_t)t * kSeedTile;
  const uint64_t tile_last = min(tile + (uint64_t)kSeedTile, unit_end) - 1;
  uint32_t lo = 0, hi = nactive + 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (active_off[mid] <= tile) lo = mid + 1; else hi = mid;
  }
  const uint32_t a = lo - 1;
  hi = nactive + 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (active_off[mid] <= tile_last) lo = mid + 1; else hi = mid;
  }
  range[t] = make_uint2(a, lo - a);  // lists a .. lo-1
}

// Item `second` of a lane's two windows, field by field (a select of whole
// structs would go through private memory).
__device__ __forceinline__ SeedWindow PickWindow(bool second, const SeedWindow (&s)[kSeedItems]) {
  SeedWindow r;
  r.w0 = second ? s[1].w0 : s[0].w0;
  r.w1 = second ? s[1].w1 : s[0].w1;
  r.w2 = second ? s[1].w2 : s[0].w2;
  r.wi = second ? s[1].wi : s[0].wi;
  r.k = second ? s[1].k : s[0].k;
  return r;
}

// The `kReverse` sides of the lane's items whose flag is set, one query word
// per round: the lane starts its next item as soon as the current walk
// stops.  Rounds repeat while any lane of the wave has a walk left.
template <bool kReverse, typename T>
__device__ __forceinline__ void WalkItems(const uint64_t *qw, const uint64_t *dw,
                                          const PairTables<T> *tab,
                                          const SeedWindow (&sq)[kSeedItems],
                                          const SeedWindow (&sd)[kSeedItems],
                                          const bool (&todo)[kSeedItems],
                                          const int (&cutoff)[kSeedItems],
                                          int (&score)[kSeedItems]) {
  static_assert(kSeedItems == 2, "the walk switches between two items");
  score[0] = 0;
  score[1] = 0;
  int cur = todo[0] ? 0 : (todo[1] ? 1 : 2);
  PairWalk w = PairStart<kReverse>(PickWindow(cur == 1, sq), PickWindow(cur == 1, sd));
  int cut = cur == 1 ? cutoff[1] : cutoff[0];
  for (;;) {
    const bool active = cur < 2;
    if (!__any(active)) break;
    if (active && !PairStep<kReverse>(qw, dw, tab, cut, w)) {
      if (cur == 0) {
        score[0] = w.best;
        cur = todo[1] ? 1 : 2;
        if (cur == 1) {
          w = PairStart<kReverse>(sq[1], sd[1]);
          cut = cutoff[1];
        }
      } else {
        score[1] = w.best;
        cur = 2;
      }
    }
  }
}

// Register budget: four waves per SIMD (at five the value kernel spills
// its per-tile state).
template <bool kValues, typename T>
__global__ void __launch_bounds__(kSeedThreads)
    __attribute__((amdgpu_waves_per_eu(4)))
    SeedKernel(DeviceQueryLists q, DeviceDbChunk db, SeedStageParams p,
               const ListInfo *info, const uint32_t *active,
               const uint64_t *active_off, const uint2 *tile_range, uint64_t unit_begin,
               uint64_t unit_end, SeedHit *hits, uint32_t capacity, uint32_t *counters) {
  __shared__ PairTables<T> ptab;
  __shared__ uint8_t differs[kTableDim * kTableDim];
  // the tile's lists: unit offset, list rank, first entry, entries, first
  // representative / value, representatives
  __shared__ uint64_t loff[kSeedTile + 1];
  __shared__ uint32_t lrank[kSeedTile], leb[kSeedTile], lp[kSeedTile], lbeg[kSeedTile],
      lnrep[kSeedTile];
  __shared__ double linv[kSeedTile];        // 1 / entries, for the unit division
  __shared__ uint32_t owner[kSeedTile];     // the tile list of each unit of the tile
  __shared__ uint32_t scan_lds[kSeedThreads / kWave];
  __shared__ SeedHit stage_lds[kSeedThreads / kWave][kWaveStage];
  __shared__ uint4 rep_ex[kValues ? 1 : kSeedThreads / kWave][kValues ? 1 : kSeedItems * kWave];
  HitStage stage{stage_lds[threadIdx.x / kWave], 0u};
  for (uint32_t i = threadIdx.x; i < kTableDim * kTableDim; i += blockDim.x) {
    PairTablesFill(ptab, p.matrix, p.nl, i);
    if (!kValues) differs[i] = DiffersEntry(p.reduced_map, i);
  }
  const int lane = threadIdx.x & (kWave - 1);
  for (uint64_t tile = unit_begin + (uint64_t)blockIdx.x * kSeedTile; tile < unit_end;
       tile += (uint64_t)gridDim.x * kSeedTile) {
    const uint64_t tile_end = min(tile + (uint64_t)kSeedTile, unit_end);
    const uint2 tr = tile_range[(tile - unit_begin) / kSeedTile];
    const uint32_t first = tr.x, count = tr.y;
    __syncthreads();  // the previous tile's table is no longer read
    for (uint32_t i = threadIdx.x; i < kSeedTile; i += kSeedThreads) owner[i] = 0;
    for (uint32_t i = threadIdx.x; i <= count; i += kSeedThreads) {
      loff[i] = active_off[first + i];
      if (i < count) {
        const uint32_t l = active[first + i];
        const ListInfo li = info[l];
        lrank[i] = l;
        leb[i] = li.ent_begin;
        lp[i] = li.nent;
        linv[i] = 1.0 / (double)(li.nent ? li.nent : 1u);
        lbeg[i] = kValues ? li.val_begin : li.rep_begin;
        lnrep[i] = li.nrep;
      }
    }
    __syncthreads();
    // owner[x] = the list holding unit tile + x: mark where each list starts
    // inside the tile, then an inclusive max-scan over the tile
    for (uint32_t i = threadIdx.x + 1; i < count; i += kSeedThreads)
      owner[loff[i] - tile] = i;  // lists after the first start inside the tile
    __syncthreads();
    {
      static_assert(kSeedItems == 2, "owner scan covers two units per thread");
      const uint32_t a = owner[2 * threadIdx.x], b = owner[2 * threadIdx.x + 1];
      const uint32_t mine = a > b ? a : b;
      uint32_t incl, excl;
      BlockScan<uint32_t, true>(mine, scan_lds, &incl, &excl, nullptr);
      owner[2 * threadIdx.x] = a > excl ? a : excl;
      owner[2 * threadIdx.x + 1] = incl;
    }
    __syncthreads();
    // all items' indices and first loads are issued before any is evaluated:
    // the entry record and database position, then the packed words around
    // both seed positions
    bool valid[kSeedItems];
    uint32_t lidx[kSeedItems], ridx[kSeedItems], eidx[kSeedItems], dposv[kSeedItems];
    uint4 rec[kSeedItems];
#pragma unroll
    for (int j = 0; j < kSeedItems; ++j) {
      const uint64_t u = tile + (uint64_t)j * kSeedThreads + threadIdx.x;
      valid[j] = u < tile_end;
      const uint32_t lo = valid[j] ? owner[u - tile] : 0u;
      const uint64_t local = valid[j] ? u - loff[lo] : 0;
      const uint32_t P = lp[lo] ? lp[lo] : 1u;
      uint32_t idx, e;
      if (local < 0xffffffffull) {
        // quotient from the double reciprocal (53 bits: off by at most one)
        const uint32_t n = (uint32_t)local;
        uint32_t qt = (uint32_t)((double)n * linv[lo]);
        int64_t r = (int64_t)n - (int64_t)qt * P;
        if (r < 0) {
          --qt;
          r += P;
        } else if (r >= (int64_t)P) {
          ++qt;
          r -= P;
        }
        idx = qt;
        e = (uint32_t)r;
      } else {
        idx = (uint32_t)(local / P);
        e = (uint32_t)(local - (uint64_t)idx * P);
      }
      lidx[j] = lo;
      ridx[j] = idx;
      eidx[j] = e;
      rec[j] = make_uint4(0, 0, 0, 0);
      dposv[j] = 0;
      if (valid[j]) {
        rec[j] = q.entry_rec[leb[lo] + e];
        const uint32_t ri = lbeg[lo] + idx;
        dposv[j] = kValues ? db.val_pos[ri] : db.rep_pos[ri];
      }
    }
    SeedWindow wq[kSeedItems], wd[kSeedItems];
#pragma unroll
    for (int j = 0; j < kSeedItems; ++j) {
      wq[j] = SeedWindow{0, 0, 0, 0, 0};
      wd[j] = SeedWindow{0, 0, 0, 0, 0};
      // (representatives: only the distance check reads them here)
      if (valid[j] && (kValues || ((rec[j].y >> 31) != 0 &&
                                   p.subsequence_length <= kCodesPerWord))) {
        wq[j] = LoadSeedWindow(q.packed, rec[j].x);
        wd[j] = LoadSeedWindow(db.packed, dposv[j]);
      }
    }
    // which items extend: every valid value unit; a representative unit
    // when its reduced-alphabet distance is within the threshold
    bool need[kSeedItems];
    int cutv[kSeedItems], trigv[kSeedItems], dist[kSeedItems];
#pragma unroll
    for (int j = 0; j < kSeedItems; ++j) {
      cutv[j] = (int)rec[j].z;
      trigv[j] = (int)rec[j].w;
      dist[j] = 0;
      need[j] = valid[j];
      if (!kValues && valid[j] && (rec[j].y >> 31) != 0) {
        dist[j] = p.subsequence_length <= kCodesPerWord
                      ? DistanceFrom(wq[j], wd[j], p.subsequence_length, differs)
                      : DistanceWindow(q.packed, rec[j].x, db.packed, dposv[j],
                                       p.subsequence_length, differs);
        need[j] = dist[j] <= p.distance_threshold;
      }
    }
    if constexpr (kValues) {
      // the reverse sides of the lane's items, then the forward sides of
      // those whose reverse score is within the trigger: a lane moves on to
      // its next item as soon as the current walk stops, so a wave runs as
      // long as its longest sum of walks rather than the sum of its longest
      // walks
      int sum[kSeedItems];
      WalkItems<true>(q.packed, db.packed, &ptab, wq, wd, need, cutv, sum);
      bool fwd[kSeedItems];
      int fsum[kSeedItems];
#pragma unroll
      for (int j = 0; j < kSeedItems; ++j) fwd[j] = need[j] && sum[j] <= trigv[j];
      WalkItems<false>(q.packed, db.packed, &ptab, wq, wd, fwd, cutv, fsum);
#pragma unroll
      for (int j = 0; j < kSeedItems; ++j) {
        if (fwd[j]) sum[j] += fsum[j];
        SeedHit h = {0, 0, 0, 0, 0, 0, 0};
        if (valid[j]) {
          h.dpos = dposv[j];
          h.qid = rec[j].y & 0x7fffffffu;
          h.qpos = rec[j].x;
          h.list = lrank[lidx[j]];
          h.entry = eidx[j];
          h.rv = lnrep[lidx[j]] + ridx[j];
        }
        StageAppend(need[j] && sum[j] > trigv[j], h, stage, hits, counters, capacity);
      }
    } else {
      // Representative units: a representative within the distance
      // threshold is walked itself, and every member of one within the
      // threshold plus the allowed mismatches is walked at the
      // representative's query position.  Item j of a lane expands to cnt[j]
      // walks (the representative's own first when it has one); the walks
      // of the wave's 128 items are dealt out 64 at a time, one per lane,
      // from a per-wave LDS record of each item (tile list, entry,
      // representative index, first member).
      uint4 *ex = rep_ex[threadIdx.x / kWave];
      uint32_t cnt[kSeedItems];
#pragma unroll
      for (int j = 0; j < kSeedItems; ++j) {
        bool member_ok = false;
        uint32_t mb = 0, me = 0;
        if (valid[j]) {
          const bool check = (rec[j].y >> 31) != 0;
          member_ok = !check || max(dist[j] - p.max_mismatches, 0) <= p.distance_threshold;
          if (member_ok) {
            const uint32_t ri = lbeg[lidx[j]] + ridx[j];
            mb = db.mem_off[ri];
            me = db.mem_off[ri + 1];
          }
        }
        cnt[j] = (need[j] ? 1u : 0u) + (member_ok ? me - mb : 0u);
        ex[2 * lane + j] = make_uint4(lidx[j], eidx[j], ridx[j], mb);
      }
      const uint32_t mine = cnt[0] + cnt[1];
      const uint32_t incl = WaveInclusiveSum(mine);
      const uint32_t total = __shfl(incl, kWave - 1, kWave);
      // item 0's walk count, and which items walk their representative
      const uint32_t split = cnt[0] << 2 | (need[0] ? 1u : 0u) | (need[1] ? 2u : 0u);
      for (uint32_t t0 = 0; t0 < total; t0 += kWave) {
        const uint32_t t = t0 + (uint32_t)lane;
        int owner = 0;  // first lane with incl > t
#pragma unroll
        for (int step = kWave / 2; step >= 1; step >>= 1) {
          const uint32_t v = __shfl(incl, owner + step - 1, kWave);
          if (v <= t) owner += step;
        }
        owner = min(owner, kWave - 1);
        const uint32_t first = __shfl(incl - mine, owner, kWave);
        const uint32_t osplit = __shfl(split, owner, kWave);
        SeedHit h = {0, 0, 0, 0, 0, 0, 0};
        bool pass = false;
        if (t < total) {
          uint32_t r = t - first;  // the walk's index among the owner's
          const uint32_t c0 = osplit >> 2;
          const int j = r >= c0 ? 1 : 0;
          if (j) r -= c0;
          const uint32_t own = (osplit >> j) & 1u;  // the item walks its representative
          const uint4 x = ex[2 * owner + j];        // tile list, entry, index, first member
          const uint4 er = q.entry_rec[leb[x.x] + x.y];
          const bool self = own != 0 && r == 0;
          h.dpos = self ? db.rep_pos[lbeg[x.x] + x.z] : db.mem_pos[x.w + r - own];
          h.member = self ? 0u : 1u + r - own;
          h.qid = er.y & 0x7fffffffu;
          h.qpos = er.x;
          h.list = lrank[x.x];
          h.rv = x.z;
          h.entry = x.y;
          const int cut = (int)er.z, trig = (int)er.w;
          pass = UngappedWithTrigger(q.packed, h.qpos, db.packed, h.dpos, &ptab, cut, trig) > trig;
        }
        StageAppend(pass, h, stage, hits, counters, capacity);
      }
    }
  }
  StageFlush(stage, hits, counters, capacity);
}

// One 16-byte record per list entry: query position, query id (bit 31: the
// distance check), ungapped cutoff and trigger of the query.
__global__ void EntryRecordKernel(DeviceQueryLists q, uint32_t n, uint4 *rec) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t qid = q.entry_qid[i];
  rec[i] = make_uint4(q.entry_pos[i], qid | (q.entry_check[i] ? 0x80000000u : 0u),
                      (uint32_t)q.cutoffs[qid], (uint32_t)q.triggers[qid]);
}

__global__ void FillU32Kernel(uint32_t *p, size_t n, uint32_t v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void QueryCountKernel(const SeedHit *hits, size_t n, uint32_t qb, uint32_t *count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) atomicAdd(&count[hits[i].qid - qb], 1u);
}

// (positions, when set, receives each grouped hit's query and database
// positions at the same index)
__global__ void QueryScatterKernel(const SeedHit *hits, size_t n, uint32_t qb,
                                   uint32_t *cursor, SeedHit *out, HitPosition *positions) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const SeedHit h = hits[i];
    const uint32_t k = atomicAdd(&cursor[h.qid - qb], 1u);
    out[k] = h;
    if (positions) positions[k] = HitPosition{h.qpos, h.dpos};
  }
}

}  // namespace

SeedStage::~SeedStage() {
  void *ps[] = {lookup_,     units_,      scan_tmp_,  active_,    hits_,
                grouped_[0], grouped_[1], positions_, qcount_,    tile_range_,
                base_info_,  entry_rec_};
  for (void *p : ps)
    if (p) (void)gz::Free(p);
  for (int k = 0; k < 2; ++k) {
    if (host_hits_[k]) (void)gz::HostFree(host_hits_[k]);
    if (host_qstart_[k]) (void)gz::HostFree(host_qstart_[k]);
  }
  if (counters_) (void)gz::Free(counters_);
  if (pinned_) (void)gz::HostFree(pinned_);
  for (hipEvent_t e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : launch_ev_)
    if (e) (void)hipEventDestroy(e);
  if (fetch_stream_) (void)hipStreamDestroy(fetch_stream_);
  if (fetch_host_) (void)gz::HostFree(fetch_host_);
}

hipError_t SeedStage::Reserve(void **p, size_t *cap, size_t bytes) {
  if (*cap >= bytes && *p) return hipSuccess;
  if (*p) {
    hipError_t e = gz::Free(*p);
    if (e != hipSuccess) return e;
    *p = nullptr;
  }
  bytes = std::max<size_t>(bytes, 256);
  hipError_t e = gz::Malloc(p, bytes);
  if (e == hipSuccess) *cap = bytes;
  return e;
}

#define GZ_TRY(x)                      \
  do {                                 \
    hipError_t _e = (x);               \
    if (_e != hipSuccess) return _e;   \
  } while (0)

hipError_t SeedStage::PrepareEntries(DeviceQueryLists *q, uint64_t entries, hipStream_t s) {
  q->entry_rec = nullptr;
  if (entries >= 0xffffffffull) return hipErrorInvalidValue;
  GZ_TRY(Reserve(&entry_rec_, &entry_rec_cap_, std::max<uint64_t>(entries, 1) * sizeof(uint4)));
  if (entries > 0) {
    hipLaunchKernelGGL(EntryRecordKernel, dim3((unsigned)((entries + 255) / 256)), dim3(256), 0,
                       s, *q, (uint32_t)entries, (uint4 *)entry_rec_);
    GZ_TRY(hipGetLastError());
  }
  q->entry_rec = (const uint4 *)entry_rec_;
  return hipSuccess;
}

hipError_t SeedStage::Lookup(const DeviceQueryLists &q, const DeviceDbChunk &db,
                             bool clustering, hipStream_t s) {
  looked_up_ = false;
  const uint32_t n = q.nlists;
  if (n > 0) {
    GZ_TRY(Reserve(&base_info_, &base_info_cap_, (size_t)n * sizeof(ListInfo)));
    hipLaunchKernelGGL(LookupKernel, dim3((n + 255) / 256), dim3(256), 0, s, q, db,
                       clustering ? 1 : 0, (ListInfo *)base_info_);
    GZ_TRY(hipGetLastError());
  }
  looked_up_ = true;
  return hipSuccess;
}

hipError_t SeedStage::Run(const DeviceQueryLists &q, const DeviceDbChunk &db,
                          const SeedStageParams &p, uint32_t qb, uint32_t qe, hipStream_t s,
                          SeedHitView *out, SeedStageStats *stats, unsigned kinds,
                          bool compact) {
  *out = SeedHitView();
  if (!looked_up_ || qb > qe || qe > q.nqueries || (qb & 63) != 0 ||
      ((qe & 63) != 0 && qe != q.nqueries))
    return hipErrorInvalidValue;
  const uint32_t nq = qe - qb;
  const int slot = flip_;
  flip_ ^= 1;
  if (!counters_) GZ_TRY(gz::Malloc(&counters_, kCounters * sizeof(uint32_t)));
  if (!pinned_) {
    GZ_TRY(gz::HostMalloc(&pinned_, 64, hipHostMallocDefault));
    pinned_cap_ = 64;
  }
  if (!ev_[0]) {
    GZ_TRY(hipEventCreate(&ev_[0]));
    GZ_TRY(hipEventCreate(&ev_[1]));
  }
  // per-query counts, also the qstart array handed back
  GZ_TRY(Reserve(&qcount_, &qcount_cap_, 2 * AlignUp((size_t)(nq + 1) * 4, 256) +
                                             DeviceScanTempBytes<uint32_t>(nq + 1)));
  if (host_qstart_cap_[slot] < (size_t)(nq + 1) * 4) {
    if (host_qstart_[slot]) GZ_TRY(gz::HostFree(host_qstart_[slot]));
    host_qstart_[slot] = nullptr;
    host_qstart_cap_[slot] = 0;
    GZ_TRY(gz::HostMalloc(&host_qstart_[slot], (size_t)(nq + 1) * 4, hipHostMallocDefault));
    host_qstart_cap_[slot] = (size_t)(nq + 1) * 4;
  }
  uint32_t *h_qstart = (uint32_t *)host_qstart_[slot];
  const uint32_t n = q.nlists;
  uint64_t total_hits = 0;
  if (n > 0 && nq > 0) {
    // layout: info | units_rep | units_val | off_rep | off_val | flags | slots
    const size_t sz_info = AlignUp(n * sizeof(ListInfo), 256);
    const size_t sz_u64 = AlignUp((size_t)(n + 1) * 8, 256);
    const size_t sz_u32 = AlignUp((size_t)(n + 1) * 4, 256);
    GZ_TRY(Reserve(&lookup_, &lookup_cap_, sz_info + 4 * sz_u64 + 2 * sz_u32));
    char *base = (char *)lookup_;
    ListInfo *info = (ListInfo *)base;
    uint64_t *units[2] = {(uint64_t *)(base + sz_info), (uint64_t *)(base + sz_info + sz_u64)};
    uint64_t *offs[2] = {(uint64_t *)(base + sz_info + 2 * sz_u64),
                         (uint64_t *)(base + sz_info + 3 * sz_u64)};
    uint32_t *flags = (uint32_t *)(base + sz_info + 4 * sz_u64);
    uint32_t *slots = (uint32_t *)(base + sz_info + 4 * sz_u64 + sz_u32);
    const size_t tmp =
        std::max(DeviceScanTempBytes<uint64_t>(n), DeviceScanTempBytes<uint32_t>(n));
    GZ_TRY(Reserve(&scan_tmp_, &scan_tmp_cap_, tmp));
    GZ_TRY(Reserve(&active_, &active_cap_, 2 * (AlignUp((size_t)(n + 1) * 4, 256) +
                                                AlignUp((size_t)(n + 1) * 8, 256))));
    if (!hits_) GZ_TRY(Reserve(&hits_, &hits_cap_, InitialHitCapacity() * sizeof(SeedHit)));

    const unsigned grid_n = (n + 255) / 256;
    hipLaunchKernelGGL(BlockUnitsKernel, dim3(grid_n), dim3(256), 0, s, q,
                       (const ListInfo *)base_info_, qb,
                       qe == q.nqueries ? 0xffffffffu : qe, info, units[0], units[1]);
    GZ_TRY(hipGetLastError());

    uint32_t *act[2];
    uint64_t *act_off[2];
    {
      char *a = (char *)active_;
      const size_t s32 = AlignUp((size_t)(n + 1) * 4, 256), s64 = AlignUp((size_t)(n + 1) * 8, 256);
      act[0] = (uint32_t *)a;
      act_off[0] = (uint64_t *)(a + s32);
      act[1] = (uint32_t *)(a + s32 + s64);
      act_off[1] = (uint64_t *)(a + 2 * s32 + s64);
    }
    uint64_t totals[2];
    uint32_t nact[2];
    for (int k = 0; k < 2; ++k) {
      GZ_TRY((DeviceScan<uint64_t, false>(units[k], offs[k], n, false, scan_tmp_, s)));
      hipLaunchKernelGGL(ActiveFlagsKernel, dim3(grid_n), dim3(256), 0, s, n, units[k], flags);
      GZ_TRY(hipGetLastError());
      GZ_TRY((DeviceScan<uint32_t, false>(flags, slots, n, false, scan_tmp_, s)));
      hipLaunchKernelGGL(ActiveScatterKernel, dim3(grid_n), dim3(256), 0, s, n, units[k],
                         offs[k], slots, act[k], act_off[k]);
      GZ_TRY(hipGetLastError());
      // totals: last offset + last units, last slot + last flag
      uint64_t *hp = (uint64_t *)pinned_;
      GZ_TRY(hipMemcpyAsync(hp, offs[k] + (n - 1), 8, hipMemcpyDeviceToHost, s));
      GZ_TRY(hipMemcpyAsync(hp + 1, units[k] + (n - 1), 8, hipMemcpyDeviceToHost, s));
      GZ_TRY(hipMemcpyAsync((uint32_t *)(hp + 2), slots + (n - 1), 4, hipMemcpyDeviceToHost, s));
      GZ_TRY(hipMemcpyAsync((uint32_t *)(hp + 2) + 1, flags + (n - 1), 4, hipMemcpyDeviceToHost,
                            s));
      GZ_TRY(hipStreamSynchronize(s));
      totals[k] = hp[0] + hp[1];
      nact[k] = ((uint32_t *)(hp + 2))[0] + ((uint32_t *)(hp + 2))[1];
    }
    if (stats) {
      stats->units_rep += totals[0];
      stats->units_value += totals[1];
    }
    int dev = 0;
    GZ_TRY(hipGetDevice(&dev));
    int cus = 256;
    GZ_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // passing seeds accumulate in hits_; counters_[0] is their running
    // count.  Every range of both kinds is queued without a host round trip
    // (a wait between launches would hand the seed kernels' compute units to
    // the other streams' long kernels); the kernels never write past the
    // buffer, and a run whose count passed it is redone whole with a buffer
    // that holds it.
    uint64_t most_tiles = 1;
    for (int k = 0; k < 2; ++k)
      if (totals[k] && ((kinds >> k) & 1))
        most_tiles = std::max<uint64_t>(
            most_tiles, (std::min<uint64_t>(totals[k], kUnitsPerLaunch) + kSeedTile - 1) / kSeedTile);
    GZ_TRY(Reserve(&tile_range_, &tile_range_cap_, most_tiles * sizeof(uint2)));
    for (;;) {
      GZ_TRY(hipMemsetAsync(counters_, 0, kCounters * sizeof(uint32_t), s));
      const uint32_t cap = (uint32_t)std::min<size_t>(hits_cap_ / sizeof(SeedHit), 0xffffffffu);
      size_t nev = 0;
      int kind_of[2] = {0, 0};
      for (int k = 0; k < 2; ++k) {
        if (totals[k] == 0 || !((kinds >> k) & 1)) continue;
        for (uint64_t u0 = 0; u0 < totals[k];) {
          const uint64_t u1 = std::min(totals[k], u0 + kUnitsPerLaunch);
          const uint64_t tiles = (u1 - u0 + kSeedTile - 1) / kSeedTile;
          const unsigned grid = (unsigned)std::min<uint64_t>(tiles, (uint64_t)cus * 8);
          hipLaunchKernelGGL(TileRangeKernel, dim3((unsigned)((tiles + 255) / 256)), dim3(256), 0,
                             s, act_off[k], nact[k], u0, u1, (uint32_t)tiles,
                             (uint2 *)tile_range_);
          GZ_TRY(hipGetLastError());
          if (stats) {
            while (launch_ev_.size() < 2 * nev + 2) {
              hipEvent_t e = nullptr;
              GZ_TRY(hipEventCreate(&e));
              launch_ev_.push_back(e);
            }
            GZ_TRY(hipEventRecord(launch_ev_[2 * nev], s));
          }
          // byte score tables when every cutoff and score fits them
          const uint2 *tr = (const uint2 *)tile_range_;
          if (k == 0 && p.byte_tables) {
            hipLaunchKernelGGL((SeedKernel<false, int8_t>), dim3(grid), dim3(kSeedThreads), 0,
                               s, q, db, p, info, act[0], act_off[0], tr, u0, u1,
                               (SeedHit *)hits_, cap, counters_);
          } else if (k == 0) {
            hipLaunchKernelGGL((SeedKernel<false, int16_t>), dim3(grid), dim3(kSeedThreads), 0,
                               s, q, db, p, info, act[0], act_off[0], tr, u0, u1,
                               (SeedHit *)hits_, cap, counters_);
          } else if (p.byte_tables) {
            hipLaunchKernelGGL((SeedKernel<true, int8_t>), dim3(grid), dim3(kSeedThreads), 0, s,
                               q, db, p, info, act[1], act_off[1], tr, u0, u1,
                               (SeedHit *)hits_, cap, counters_);
          } else {
            hipLaunchKernelGGL((SeedKernel<true, int16_t>), dim3(grid), dim3(kSeedThreads), 0,
                               s, q, db, p, info, act[1], act_off[1], tr, u0, u1,
                               (SeedHit *)hits_, cap, counters_);
          }
          GZ_TRY(hipGetLastError());
          if (stats) {
            GZ_TRY(hipEventRecord(launch_ev_[2 * nev + 1], s));
            ++nev;
            ++kind_of[k];
          }
          u0 = u1;
        }
      }
      uint32_t *hc = (uint32_t *)pinned_;
      // the count's copy is queued once the seed kernels are done: a
      // device-to-host copy queued behind a running kernel holds up the
      // other streams' copies until that kernel ends
      GZ_TRY(hipStreamSynchronize(s));
      GZ_TRY(hipMemcpyAsync(hc, counters_, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
      GZ_TRY(hipStreamSynchronize(s));
      const uint32_t count = hc[0];
      if (count > cap) {
        // the buffer overflowed: grow it to hold the run and redo it
        GZ_TRY(gz::Free(hits_));
        hits_ = nullptr;
        hits_cap_ = 0;
        const size_t want = ((size_t)count + (size_t)count / 4 + 1024) * sizeof(SeedHit);
        GZ_TRY(gz::Malloc(&hits_, want));
        hits_cap_ = want;
        continue;
      }
      total_hits = count;
      if (stats) {
        size_t e = 0;
        for (int k = 0; k < 2; ++k)
          for (int j = 0; j < kind_of[k]; ++j, ++e) {
            float ms = 0;
            GZ_TRY(hipEventElapsedTime(&ms, launch_ev_[2 * e], launch_ev_[2 * e + 1]));
            stats->kernel_ms[k] += ms;
            stats->launches[k] += 1;
          }
      }
      break;
    }
  }
  if (stats) stats->hits += total_hits;
  // group by query: counts, exclusive scan, scatter
  uint32_t *qcount = (uint32_t *)qcount_;
  uint32_t *qstart = (uint32_t *)((char *)qcount_ + AlignUp((size_t)(nq + 1) * 4, 256));
  void *qtemp = (char *)qcount_ + 2 * AlignUp((size_t)(nq + 1) * 4, 256);
  hipLaunchKernelGGL(FillU32Kernel, dim3((nq + 1 + 255) / 256), dim3(256), 0, s, qcount,
                     (size_t)nq + 1, 0u);
  if (total_hits) {
    GZ_TRY(Reserve(&grouped_[slot], &grouped_cap_[slot], total_hits * sizeof(SeedHit)));
    if (compact) GZ_TRY(Reserve(&positions_, &positions_cap_, total_hits * sizeof(HitPosition)));
    const unsigned g = (unsigned)((total_hits + 255) / 256);
    hipLaunchKernelGGL(QueryCountKernel, dim3(g), dim3(256), 0, s, (const SeedHit *)hits_,
                       (size_t)total_hits, qb, qcount);
  }
  GZ_TRY((DeviceScan<uint32_t, false>(qcount, qstart, nq + 1, false, qtemp, s)));
  GZ_TRY(hipMemcpyAsync(h_qstart, qstart, (size_t)(nq + 1) * 4, hipMemcpyDeviceToHost, s));
  if (total_hits) {
    // qcount becomes the scatter cursor
    GZ_TRY(hipMemcpyAsync(qcount, qstart, (size_t)nq * 4, hipMemcpyDeviceToDevice, s));
    const unsigned g = (unsigned)((total_hits + 255) / 256);
    hipLaunchKernelGGL(QueryScatterKernel, dim3(g), dim3(256), 0, s, (const SeedHit *)hits_,
                       (size_t)total_hits, qb, qcount, (SeedHit *)grouped_[slot],
                       compact ? (HitPosition *)positions_ : nullptr);
    GZ_TRY(hipGetLastError());
    // a compact view moves 8 bytes per hit to the host instead of 28
    const size_t bytes = total_hits * (compact ? sizeof(HitPosition) : sizeof(SeedHit));
    if (host_hits_cap_[slot] < bytes) {
      if (host_hits_[slot]) GZ_TRY(gz::HostFree(host_hits_[slot]));
      host_hits_[slot] = nullptr;
      host_hits_cap_[slot] = 0;
      const size_t want = bytes + bytes / 4;
      GZ_TRY(gz::HostMalloc(&host_hits_[slot], want, hipHostMallocDefault));
      host_hits_cap_[slot] = want;
    }
    GZ_TRY(hipMemcpyAsync(host_hits_[slot], compact ? positions_ : grouped_[slot], bytes,
                          hipMemcpyDeviceToHost, s));
  }
  GZ_TRY(hipStreamSynchronize(s));
  if (compact) {
    out->positions = (const HitPosition *)host_hits_[slot];
    const SeedHit *dev = (const SeedHit *)grouped_[slot];
    // every range of one call is copied into pinned staging before a single
    // wait, and the records are read from there
    out->fetch = [this, dev](size_t n, const size_t *begin, const size_t *count,
                             const SeedHit **records) -> int {
      std::lock_guard<std::mutex> lock(fetch_mutex_);
      hipError_t e = hipSuccess;
      if (!fetch_stream_) e = hipStreamCreateWithFlags(&fetch_stream_, hipStreamNonBlocking);
      size_t total = 0;
      for (size_t i = 0; i < n; ++i) total += count[i];
      const size_t bytes = total * sizeof(SeedHit);
      if (e == hipSuccess && fetch_host_cap_ < bytes) {
        if (fetch_host_) e = gz::HostFree(fetch_host_);
        fetch_host_ = nullptr;
        fetch_host_cap_ = 0;
        if (e == hipSuccess) e = gz::HostMalloc(&fetch_host_, bytes + bytes / 4, hipHostMallocDefault);
        if (e == hipSuccess) fetch_host_cap_ = bytes + bytes / 4;
      }
      SeedHit *stage = (SeedHit *)fetch_host_;
      for (size_t i = 0, at = 0; i < n && e == hipSuccess; at += count[i], ++i)
        if (count[i])
          e = hipMemcpyAsync(stage + at, dev + begin[i], count[i] * sizeof(SeedHit),
                             hipMemcpyDeviceToHost, fetch_stream_);
      if (e == hipSuccess) e = hipStreamSynchronize(fetch_stream_);
      *records = stage;
      return (int)e;
    };
  } else {
    out->hits = (const SeedHit *)host_hits_[slot];
  }
  out->count = total_hits;
  out->qstart = h_qstart;
  return hipSuccess;
}

}  // namespace gz
