// query_lists_device.hip -- see query_lists_device.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "gz_common.h"
#include "gz_runtime.h"
#include "gz_device.h"
#include "gz_scan.h"
#include "query_lists_device.h"
#include "radix_sort.h"
#include "ungapped.h"

namespace gz {

namespace {

constexpr uint32_t kMaxHashCodes = 32;  // code_map / scores table size

struct HashParams {
  uint8_t code_map[kMaxHashCodes];  // protein code -> shifted reduced code
  int16_t scores[kMaxHashCodes];    // per shifted code
  uint32_t max_code, max_length, shift;
  int threshold;
};

inline unsigned Blocks(size_t n, unsigned t = 256) { return (unsigned)((n + t - 1) / t); }

// fwd[p] = p + 1 at a delimiter; rev[n - 1 - p] = n - p at a delimiter; 0
// elsewhere.  After inclusive max-scans, fwd[p] is the first position of
// p's segment and n - rev[n - 1 - p] the delimiter that ends it.
__global__ void DelimiterMarksKernel(const uint64_t *packed, uint32_t n, uint32_t *fwd,
                                     uint32_t *rev) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const bool d = PackedCodeAt(packed, p) == kDelimiter;
  fwd[p] = d ? p + 1 : 0u;
  rev[n - 1 - p] = d ? n - p : 0u;
}

struct Located {
  uint32_t qid, pos;
};

// appearance index j -> (query, concatenated position): queries in the
// appearance order, each over its framed length
__device__ __forceinline__ Located Locate(uint32_t j, const uint32_t *app_start,
                                          const uint32_t *app_qid, const uint32_t *coff,
                                          uint32_t nq) {
  uint32_t lo = 0, hi = nq;  // largest k with app_start[k] <= j
  while (lo + 1 < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (app_start[mid] <= j) lo = mid; else hi = mid;
  }
  Located r;
  r.qid = app_qid[lo];
  r.pos = coff[r.qid] + (j - app_start[lo]);
  return r;
}

__device__ __forceinline__ bool HashWindow(uint64_t w, const uint8_t *code_map,
                                           const int16_t *scores, const HashParams &hp,
                                           uint32_t *out) {
  uint32_t h = 0;
  int score = 0;
  for (uint32_t l = 0; l < hp.max_length; ++l) {
    const uint32_t c0 = (uint32_t)(w >> (kCodeBits * l)) & 31u;
    if (c0 < kNumberLetters && code_map[c0] <= hp.max_code) {
      if (l >= 6 && score >= hp.threshold) break;
      h = (h << hp.shift) + code_map[c0];
      score += scores[code_map[c0]];
    } else {
      return false;
    }
  }
  *out = h;
  return true;
}

__device__ __forceinline__ uint64_t WindowAt(const uint64_t *packed, uint32_t p) {
  const uint32_t c = p + kPackPadCodes, w = c / kCodesPerWord;
  return Window12(packed[w] & kCodeBitsMask, packed[w + 1] & kCodeBitsMask,
                  c - w * kCodesPerWord);
}

// Per appearance index: entry flag and hash.
__global__ void EntryFlagsKernel(const uint64_t *packed, uint32_t n_codes,
                                 const uint32_t *app_start, const uint32_t *app_qid,
                                 const uint32_t *coff, uint32_t nq, uint32_t total,
                                 const uint32_t *fwd, const uint32_t *rev, HashParams hp,
                                 uint32_t L, uint32_t *flag, uint32_t *hashv) {
  __shared__ uint8_t code_map[kMaxHashCodes];
  __shared__ int16_t scores[kMaxHashCodes];
  if (threadIdx.x < kMaxHashCodes) {
    code_map[threadIdx.x] = hp.code_map[threadIdx.x];
    scores[threadIdx.x] = hp.scores[threadIdx.x];
  }
  __syncthreads();
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total) return;
  const Located at = Locate(j, app_start, app_qid, coff, nq);
  const uint32_t p = at.pos;
  const uint32_t s = fwd[p], e = n_codes - rev[n_codes - 1 - p];
  uint32_t h = 0;
  bool valid = false;
  if (p >= s && p < e && e - s >= L) valid = HashWindow(WindowAt(packed, p), code_map, scores, hp, &h);
  flag[j] = valid ? 1u : 0u;
  hashv[j] = h;
}

// Entries in appearance order: sort key (hash), value (entry index),
// position, query and check flag.
__global__ void CompactEntriesKernel(uint32_t n_codes, const uint32_t *app_start,
                                     const uint32_t *app_qid, const uint32_t *coff, uint32_t nq,
                                     uint32_t total, const uint32_t *fwd, const uint32_t *rev,
                                     uint32_t L, const uint32_t *flag, const uint32_t *hashv,
                                     const uint32_t *index, uint32_t *keys, uint32_t *values,
                                     uint32_t *pos, uint32_t *qid, uint8_t *check) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total || !flag[j]) return;
  const Located at = Locate(j, app_start, app_qid, coff, nq);
  const uint32_t p = at.pos;
  const uint32_t s = fwd[p], e = n_codes - rev[n_codes - 1 - p];
  const uint32_t o = index[j];
  keys[o] = hashv[j];
  values[o] = o;
  pos[o] = p;
  qid[o] = at.qid;
  check[o] = (s + L / 2 <= p && p < e - L / 2) ? 1 : 0;
}

__global__ void GroupHeadsKernel(const uint32_t *keys, uint32_t n, uint32_t *head) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// gstart[g]; the group's first appearance (its first entry's index) as the
// key of the group sort.  heads: inclusive scan of the group heads, so the
// group of item i is heads[i] - 1.
__global__ void GroupStartsKernel(const uint32_t *keys, const uint32_t *values, uint32_t n,
                                  const uint32_t *heads, uint32_t *gstart, uint32_t *gfirst,
                                  uint32_t *gorder) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0 || keys[i] != keys[i - 1]) {
    const uint32_t g = heads[i] - 1;
    gstart[g] = i;
    gfirst[g] = values[i];
    gorder[g] = g;
  }
}

// group_of_rank: first-appearance order reversed inside blocks of 256
__global__ void RankKernel(const uint32_t *by_appearance, uint32_t ng, const uint32_t *gstart,
                           const uint32_t *keys, uint32_t n, uint32_t *rank_of_group,
                           uint32_t *list_size, uint32_t *out_hash) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ng) return;
  const uint32_t b = r & ~255u, e = min(ng, b + 256u);
  const uint32_t g = by_appearance[e - 1 - (r - b)];
  rank_of_group[g] = r;
  const uint32_t s = gstart[g], t = g + 1 < ng ? gstart[g + 1] : n;
  list_size[r] = t - s;
  out_hash[r] = keys[s];
}

__global__ void ScatterEntriesKernel(const uint32_t *values, uint32_t n, const uint32_t *heads,
                                     const uint32_t *gstart, const uint32_t *rank_of_group,
                                     const uint32_t *entry_off, const uint32_t *pos,
                                     const uint32_t *qid, const uint8_t *check,
                                     uint32_t *out_pos, uint32_t *out_qid, uint8_t *out_check) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t g = heads[i] - 1;
  const uint32_t o = entry_off[rank_of_group[g]] + (i - gstart[g]);
  const uint32_t e = values[i];
  out_pos[o] = pos[e];
  out_qid[o] = qid[e];
  out_check[o] = check[e];
}

}  // namespace

#define GZ_TRY(x)                      \
  do {                                 \
    hipError_t _e = (x);               \
    if (_e != hipSuccess) return _e;   \
  } while (0)

DeviceListBuilder::~DeviceListBuilder() {
  Buffer *bs[] = {&host_arrays_, &seg_fwd_, &seg_rev_, &flags_, &hashes_, &index_,
                  &scan_tmp_, &sort_tmp_, &keys_, &values_, &keys_alt_, &values_alt_,
                  &pos_, &qid_, &check_, &gid_, &gstart_, &gfirst_, &gorder_, &galt_,
                  &galt2_, &rank_of_group_, &list_size_, &out_hash_, &out_off_, &out_pos_,
                  &out_qid_, &out_check_};
  for (Buffer *b : bs)
    if (b->p) (void)gz::Free(b->p);
  if (pinned_) (void)gz::HostFree(pinned_);
}

hipError_t DeviceListBuilder::Reserve(Buffer &b, size_t bytes) {
  bytes = std::max<size_t>(AlignUp(bytes, 256), 256);
  if (b.cap >= bytes && b.p) return hipSuccess;
  if (b.p) GZ_TRY(gz::Free(b.p));
  b.p = nullptr;
  b.cap = 0;
  GZ_TRY(gz::Malloc(&b.p, bytes));
  b.cap = bytes;
  return hipSuccess;
}

hipError_t DeviceListBuilder::Build(const uint64_t *packed, uint64_t n_codes, const QuerySet &q,
                                    const std::vector<uint64_t> &coff, const HashFunction &hf,
                                    uint32_t L, hipStream_t s, DeviceQueryLists *out,
                                    uint64_t *entries) {
  const size_t nq = q.size();
  out->nlists = 0;
  *entries = 0;
  if (hf.max_length > kCodesPerWord || hf.scores.size() > kMaxHashCodes ||
      hf.max_code >= kMaxHashCodes || n_codes >= 0xffffffffull)
    return hipErrorInvalidValue;
  HashParams hp;
  memset(&hp, 0, sizeof(hp));
  for (uint32_t c = 0; c < kNumberLetters; ++c) hp.code_map[c] = hf.code_map[c];
  for (size_t c = 0; c < hf.scores.size(); ++c) hp.scores[c] = (int16_t)hf.scores[c];
  hp.max_code = hf.max_code;
  hp.max_length = hf.max_length;
  hp.shift = hf.shift;
  hp.threshold = hf.threshold;
  if (!pinned_) GZ_TRY(gz::HostMalloc(&pinned_, 64, hipHostMallocDefault));
  uint32_t *hp32 = (uint32_t *)pinned_;

  // appearance order (blocks of 64 queries, reversed inside a block) and the
  // appearance index where each query starts
  std::vector<uint32_t> host(3 * nq + 1);
  uint32_t *app_start = host.data(), *app_qid = app_start + nq + 1, *qoff = app_qid + nq;
  uint64_t total = 0;
  for (size_t b = 0; b < nq; b += 64) {
    const size_t e = std::min(nq, b + 64);
    for (size_t k = 0; k < e - b; ++k) {
      const uint32_t qi = (uint32_t)(e - 1 - k);
      app_start[b + k] = (uint32_t)total;
      app_qid[b + k] = qi;
      total += q.Length(qi);
    }
  }
  app_start[nq] = (uint32_t)total;
  for (size_t i = 0; i < nq; ++i) qoff[i] = (uint32_t)coff[i];
  if (total >= 0xffffffffull) return hipErrorInvalidValue;
  if (nq == 0 || total == 0) return hipSuccess;
  GZ_TRY(Reserve(host_arrays_, host.size() * 4));
  uint32_t *d_app_start = (uint32_t *)host_arrays_.p;
  uint32_t *d_app_qid = d_app_start + nq + 1, *d_coff = d_app_qid + nq;
  GZ_TRY(hipMemcpyAsync(host_arrays_.p, host.data(), host.size() * 4, hipMemcpyHostToDevice, s));

  // segment bounds
  const uint32_t n = (uint32_t)n_codes;
  GZ_TRY(Reserve(seg_fwd_, (size_t)n * 4));
  GZ_TRY(Reserve(seg_rev_, (size_t)n * 4));
  const size_t scan_bytes = std::max(DeviceScanTempBytes<uint32_t>(n),
                                     DeviceScanTempBytes<uint32_t>(total + 1));
  GZ_TRY(Reserve(scan_tmp_, scan_bytes));
  uint32_t *fwd = (uint32_t *)seg_fwd_.p, *rev = (uint32_t *)seg_rev_.p;
  hipLaunchKernelGGL(DelimiterMarksKernel, dim3(Blocks(n)), dim3(256), 0, s, packed, n, fwd, rev);
  GZ_TRY(hipGetLastError());
  GZ_TRY((DeviceScan<uint32_t, true>(fwd, fwd, n, true, scan_tmp_.p, s)));
  GZ_TRY((DeviceScan<uint32_t, true>(rev, rev, n, true, scan_tmp_.p, s)));

  // entry flags and hashes in appearance order, then compaction
  const uint32_t T = (uint32_t)total;
  GZ_TRY(Reserve(flags_, (size_t)T * 4));
  GZ_TRY(Reserve(hashes_, (size_t)T * 4));
  GZ_TRY(Reserve(index_, (size_t)T * 4));
  uint32_t *flag = (uint32_t *)flags_.p, *hashv = (uint32_t *)hashes_.p;
  uint32_t *index = (uint32_t *)index_.p;
  hipLaunchKernelGGL(EntryFlagsKernel, dim3(Blocks(T)), dim3(256), 0, s, packed, n, d_app_start,
                     d_app_qid, d_coff, (uint32_t)nq, T, fwd, rev, hp, L, flag, hashv);
  GZ_TRY(hipGetLastError());
  GZ_TRY((DeviceScan<uint32_t, false>(flag, index, T, false, scan_tmp_.p, s)));
  GZ_TRY(hipMemcpyAsync(hp32, index + (T - 1), 4, hipMemcpyDeviceToHost, s));
  GZ_TRY(hipMemcpyAsync(hp32 + 1, flag + (T - 1), 4, hipMemcpyDeviceToHost, s));
  GZ_TRY(hipStreamSynchronize(s));
  const uint32_t ne = hp32[0] + hp32[1];
  *entries = ne;
  if (ne == 0) return hipSuccess;
  GZ_TRY(Reserve(keys_, (size_t)ne * 4));
  GZ_TRY(Reserve(values_, (size_t)ne * 4));
  GZ_TRY(Reserve(keys_alt_, (size_t)ne * 4));
  GZ_TRY(Reserve(values_alt_, (size_t)ne * 4));
  GZ_TRY(Reserve(pos_, (size_t)ne * 4));
  GZ_TRY(Reserve(qid_, (size_t)ne * 4));
  GZ_TRY(Reserve(check_, (size_t)ne));
  GZ_TRY(Reserve(gid_, (size_t)ne * 4));
  uint32_t *keys = (uint32_t *)keys_.p, *values = (uint32_t *)values_.p;
  uint32_t *pos = (uint32_t *)pos_.p, *qid = (uint32_t *)qid_.p, *gid = (uint32_t *)gid_.p;
  uint8_t *check = (uint8_t *)check_.p;
  hipLaunchKernelGGL(CompactEntriesKernel, dim3(Blocks(T)), dim3(256), 0, s, n, d_app_start,
                     d_app_qid, d_coff, (uint32_t)nq, T, fwd, rev, L, flag, hashv, index, keys,
                     values, pos, qid, check);
  GZ_TRY(hipGetLastError());

  // stable sort by hash: groups of equal hashes, appearance order inside
  GZ_TRY(Reserve(sort_tmp_, RadixSortTempBytes(ne)));
  GZ_TRY(RadixSortPairs(keys, values, (uint32_t *)keys_alt_.p, (uint32_t *)values_alt_.p, ne,
                        32, sort_tmp_.p, s));
  GZ_TRY(Reserve(scan_tmp_, std::max(scan_bytes, DeviceScanTempBytes<uint32_t>(ne))));
  uint32_t *head = (uint32_t *)keys_alt_.p;  // free after the sort
  hipLaunchKernelGGL(GroupHeadsKernel, dim3(Blocks(ne)), dim3(256), 0, s, keys, ne, head);
  GZ_TRY(hipGetLastError());
  // gid: inclusive scan of the heads (group of item i = gid[i] - 1)
  GZ_TRY((DeviceScan<uint32_t, false>(head, gid, ne, true, scan_tmp_.p, s)));
  GZ_TRY(hipMemcpyAsync(hp32, gid + (ne - 1), 4, hipMemcpyDeviceToHost, s));
  GZ_TRY(hipStreamSynchronize(s));
  const uint32_t ng = hp32[0];
  GZ_TRY(Reserve(gstart_, (size_t)ng * 4));
  GZ_TRY(Reserve(gfirst_, (size_t)ng * 4));
  GZ_TRY(Reserve(gorder_, (size_t)ng * 4));
  GZ_TRY(Reserve(galt_, (size_t)ng * 4));
  GZ_TRY(Reserve(galt2_, (size_t)ng * 4));
  GZ_TRY(Reserve(rank_of_group_, (size_t)ng * 4));
  GZ_TRY(Reserve(list_size_, (size_t)(ng + 1) * 4));
  GZ_TRY(Reserve(out_hash_, (size_t)ng * 4));
  GZ_TRY(Reserve(out_off_, (size_t)(ng + 1) * 4));
  GZ_TRY(Reserve(out_pos_, (size_t)ne * 4));
  GZ_TRY(Reserve(out_qid_, (size_t)ne * 4));
  GZ_TRY(Reserve(out_check_, (size_t)ne));
  uint32_t *gstart = (uint32_t *)gstart_.p, *gfirst = (uint32_t *)gfirst_.p;
  uint32_t *gorder = (uint32_t *)gorder_.p;
  hipLaunchKernelGGL(GroupStartsKernel, dim3(Blocks(ne)), dim3(256), 0, s, keys, values, ne, gid,
                     gstart, gfirst, gorder);
  GZ_TRY(hipGetLastError());
  // groups by first appearance (distinct keys below 2^32)
  GZ_TRY(Reserve(sort_tmp_, std::max(RadixSortTempBytes(ne), RadixSortTempBytes(ng))));
  GZ_TRY(RadixSortPairs(gfirst, gorder, (uint32_t *)galt_.p, (uint32_t *)galt2_.p, ng, 32,
                        sort_tmp_.p, s));
  uint32_t *rank_of_group = (uint32_t *)rank_of_group_.p, *list_size = (uint32_t *)list_size_.p;
  uint32_t *out_hash = (uint32_t *)out_hash_.p, *out_off = (uint32_t *)out_off_.p;
  hipLaunchKernelGGL(RankKernel, dim3(Blocks(ng)), dim3(256), 0, s, gorder, ng, gstart, keys, ne,
                     rank_of_group, list_size, out_hash);
  GZ_TRY(hipGetLastError());
  // entry offsets in rank order (the last one is the entry count)
  GZ_TRY(hipMemsetAsync(list_size + ng, 0, 4, s));
  GZ_TRY((DeviceScan<uint32_t, false>(list_size, out_off, (size_t)ng + 1, false, scan_tmp_.p, s)));
  hipLaunchKernelGGL(ScatterEntriesKernel, dim3(Blocks(ne)), dim3(256), 0, s, values, ne, gid,
                     gstart, rank_of_group, out_off, pos, qid, check, (uint32_t *)out_pos_.p,
                     (uint32_t *)out_qid_.p, (uint8_t *)out_check_.p);
  GZ_TRY(hipGetLastError());
  GZ_TRY(hipStreamSynchronize(s));
  out->nlists = ng;
  out->hash = out_hash;
  out->entry_off = out_off;
  out->entry_pos = (const uint32_t *)out_pos_.p;
  out->entry_qid = (const uint32_t *)out_qid_.p;
  out->entry_check = (const uint8_t *)out_check_.p;
  return hipSuccess;
}

}  // namespace gz
