// query_lists_device.h -- the query hash lists of a query chunk built on the
// device (SeedSearcherGpuQueryParameters::Build /
// SeedSearcherQueryParameters::Build, src/seed_searcher_query_parameters.cpp
// :17-272, restated on the host in query_lists.cpp).
//
// The lists are identical to the host's: every query position whose
// delimiter-free segment is at least the subsequence length and whose
// variable-length reduced hash is defined becomes an entry (check flag: the
// position lies at least L/2 inside its segment); entries are grouped by
// hash, in appearance order inside a group (queries visited in blocks of 64,
// reversed inside a block); the groups are numbered by first appearance and
// then reversed in blocks of 256, which is the list rank.
//
// Device steps: segment bounds by max-scans over delimiter positions; one
// thread per position in appearance order hashes it (12 codes of the packed
// concatenation); compaction; a stable radix sort by hash; group heads and
// starts; a sort of the groups by first appearance; ranks; the lists laid out
// in rank order by one scatter.
#pragma once

#include <stdint.h>

#include <hip/hip_runtime.h>

#include <vector>

#include "database.h"
#include "queries.h"
#include "search_device.h"

namespace gz {

class DeviceListBuilder {
 public:
  ~DeviceListBuilder();
  // packed: the padded packed concatenation of the chunk's queries (n_codes
  // codes, shared delimiters); coff: each query's offset in it.  Fills the
  // hash / entry arrays and nlists of *out (device memory owned by the
  // builder, valid until the next Build); *entries receives the entry count.
  hipError_t Build(const uint64_t *packed, uint64_t n_codes, const QuerySet &q,
                   const std::vector<uint64_t> &coff, const HashFunction &hash,
                   uint32_t subsequence_length, hipStream_t stream, DeviceQueryLists *out,
                   uint64_t *entries);

 private:
  struct Buffer {
    void *p = nullptr;
    size_t cap = 0;
  };
  hipError_t Reserve(Buffer &b, size_t bytes);
  Buffer host_arrays_, seg_fwd_, seg_rev_, flags_, hashes_, index_, scan_tmp_, sort_tmp_;
  Buffer keys_, values_, keys_alt_, values_alt_, pos_, qid_, check_, gid_;
  Buffer gstart_, gfirst_, gorder_, galt_, galt2_, rank_of_group_, list_size_;
  Buffer out_hash_, out_off_, out_pos_, out_qid_, out_check_;
  void *pinned_ = nullptr;
};

}  // namespace gz
