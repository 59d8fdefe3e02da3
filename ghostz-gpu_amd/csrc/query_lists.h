// query_lists.h -- per-hash lists of query seed positions
// (SeedSearcherGpuQueryParameters, src/seed_searcher_gpu_query_parameters.cpp
// :17-272, with the list order of SeedSearcherQueryParameters::Build,
// src/seed_searcher_query_parameters.cpp:17-219).
#pragma once

#include <stdint.h>

#include <vector>

#include "database.h"
#include "queries.h"

namespace gz {

// Lists are stored in the order the presearch visits them (hash lists are
// handed out in blocks of 256 and taken from the back of each block,
// src/aligner_presearch_thread.cpp:140-152); `rank` is that order.  Entries
// of a list keep their order of first appearance (queries visited in
// reverse within blocks of 64, positions ascending).
struct QueryLists {
  std::vector<uint32_t> hash;       // per list
  std::vector<uint32_t> entry_off;  // per list + 1
  std::vector<uint32_t> entry_pos;  // concatenated query position
  std::vector<uint32_t> entry_qid;
  std::vector<uint8_t> entry_check; // similarity (distance) check required
  size_t NumberLists() const { return hash.size(); }
};

// Offsets of each query in the shared-delimiter concatenation the device
// uses (AlignerGpu::SetQueriesData, src/aligner_gpu.cpp:213-242).
void ConcatenatedOffsets(const QuerySet &q, std::vector<uint64_t> *offsets,
                         std::vector<uint8_t> *concatenated);

void BuildQueryLists(const QuerySet &q, const std::vector<uint64_t> &concat_offsets,
                     const Database &db, int threads, QueryLists *out);

}  // namespace gz
