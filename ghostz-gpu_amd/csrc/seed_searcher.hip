// seed_searcher.hip -- SeedSearcherGpu's entry points
// (src/seed_searcher_gpu.h:28-115, src/seed_searcher_gpu.cpp:48-548) on the
// device seed stage.
//
// The reference enumerates candidate pairs on the host, list by list, into
// seed lists that its GPU kernels score (distance, then ungapped extension
// with trigger) before the host compacts the passing seeds by query
// (src/aligner_gpu_presearch_thread.cpp:822-896).  Here Reset makes the query
// chunk and one database chunk resident and builds and looks up the hash
// lists on the device; a search call runs the fused device stage (lookup,
// distance, ungapped extension, compaction) for a range of queries and
// returns the passing seeds -- the hits the reference's whole chain hands to
// the chain filter -- per query in the one-thread CPU enumeration order.
// The distance-pair enumeration of one list is kept too, for callers that
// drive the distance calculator themselves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <string>
#include <vector>

#include "aligner.h"
#include "capi_types.h"
#include "ghostz_hip.h"
#include "gz_common.h"
#include "presearch.h"
#include "query_lists.h"
#include "statistics.h"

using namespace gz;

struct gz_seed_searcher_s {
  std::vector<int> matrix;
  AlignerOptions opt;
  std::unique_ptr<Aligner> dev;
  const gz_database_s *db = nullptr;
  const gz_queries_s *queries = nullptr;
  uint32_t chunk = 0;
  bool resident = false;
  std::vector<uint8_t> concat;
  std::vector<uint64_t> coff;
  AlignStats stats;
  // the last search's hits
  std::vector<uint32_t> qoff, qpos, dpos;
  ~gz_seed_searcher_s() {
    if (dev) dev->EndQueryChunk();
  }
};

namespace {

gz_status Search(gz_seed_searcher_t h, uint32_t qb, uint32_t qe, unsigned kinds,
                 gz_seed_hits *out) {
  GZ_REQUIRE(h && out, "seed search: null argument");
  GZ_REQUIRE(h->resident, "seed search: Reset first");
  const uint32_t nq = (uint32_t)h->queries->q.size();
  GZ_REQUIRE(qb <= qe && qe <= nq && (qb & 63) == 0 && ((qe & 63) == 0 || qe == nq),
             "seed search: queries [qb, qe) must start at a multiple of 64 and end at one "
             "or at the last query");
  h->dev->SetSeedKinds(kinds);
  SeedHitView v;
  const int r = h->dev->SeedHits(h->chunk, qb, qe, &v, &h->stats);
  h->dev->SetSeedKinds(3);
  if (r != 0) {
    SetError(std::string("seed search: ") + hipGetErrorString((hipError_t)r));
    return GZ_ERROR_HIP;
  }
  // the full records (a compact view keeps them on the device)
  const SeedHit *hits = v.hits;
  if (!hits && v.count) {
    const size_t zero = 0;
    const int f = v.fetch(1, &zero, &v.count, &hits);
    if (f != 0) {
      SetError(std::string("seed search: ") + hipGetErrorString((hipError_t)f));
      return GZ_ERROR_HIP;
    }
  }
  // per query, the one-thread enumeration order: by (list rank,
  // representative or value index, member, entry), which are distinct
  const uint32_t n = qe - qb;
  h->qoff.assign(n + 1, 0);
  h->qpos.resize(v.count);
  h->dpos.resize(v.count);
  struct Keyed {
    uint64_t hi, lo;
    uint32_t q, d;
    bool operator<(const Keyed &o) const { return hi != o.hi ? hi < o.hi : lo < o.lo; }
  };
  std::vector<Keyed> keyed;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t s0 = v.qstart[i], s1 = v.qstart[i + 1];
    h->qoff[i] = s0;
    keyed.resize(s1 - s0);
    for (uint32_t k = s0; k < s1; ++k) {
      const SeedHit &x = hits[k];
      keyed[k - s0] = {((uint64_t)x.list << 32) | x.rv, ((uint64_t)x.member << 32) | x.entry,
                       x.qpos - (uint32_t)h->coff[qb + i], x.dpos};
    }
    std::sort(keyed.begin(), keyed.end());
    for (uint32_t k = s0; k < s1; ++k) {
      h->qpos[k] = keyed[k - s0].q;
      h->dpos[k] = keyed[k - s0].d;
    }
  }
  h->qoff[n] = (uint32_t)v.count;
  out->number_hits = v.count;
  out->number_queries = n;
  out->query_offsets = h->qoff.data();
  out->query_positions = h->qpos.data();
  out->database_positions = h->dpos.data();
  return GZ_OK;
}

}  // namespace

extern "C" {

gz_status gz_seed_searcher_create(const gz_aligner_parameters *p, gz_seed_searcher_t *out) {
  GZ_REQUIRE(out, "gz_seed_searcher_create: null argument");
  gz_aligner_parameters d;
  gz_aligner_default_parameters(&d);
  if (!p) p = &d;
  std::unique_ptr<gz_seed_searcher_s> h(new gz_seed_searcher_s());
  h->matrix.assign(25 * 25, 0);
  uint32_t nl = 0;
  const gz_status st = gz_score_matrix_blosum62(h->matrix.data(), &nl);
  if (st != GZ_OK) return st;
  SearchParameters &s = h->opt.search;
  s.matrix = h->matrix.data();
  s.nl = nl;
  s.gap_open = p->gap_open;
  s.gap_extension = p->gap_extension;
  s.normalized_ungapped_cutoff = p->normalized_ungapped_cutoff;
  s.normalized_gapped_trigger = p->normalized_gapped_trigger;
  s.normalized_presearch_cutoff = p->normalized_presearch_cutoff;
  s.normalized_result_cutoff = p->normalized_result_cutoff;
  s.max_results = p->max_results;
  s.max_results_per_subject = p->max_results_per_subject;
  h->opt.threads = p->threads;
  h->dev.reset(new Aligner(h->opt));
  *out = h.release();
  return GZ_OK;
}

gz_status gz_seed_searcher_destroy(gz_seed_searcher_t h) {
  delete h;
  return GZ_OK;
}

gz_status gz_seed_searcher_reset(gz_seed_searcher_t h, gz_queries_t queries,
                                 gz_database_t database, uint32_t chunk,
                                 uint32_t *number_hash_lists) {
  GZ_REQUIRE(h && queries && database, "gz_seed_searcher_reset: null argument");
  GZ_REQUIRE(chunk < database->db.chunks.size(), "gz_seed_searcher_reset: no such chunk");
  GZ_REQUIRE(queries->q.size() > 0, "gz_seed_searcher_reset: no queries");
  if (h->db != database) {
    h->resident = false;
    h->queries = nullptr;
    GZ_HIP_CHECK(h->dev->SetDatabase(&database->db));
    h->db = database;
  }
  if (h->queries != queries) {
    // the query chunk: cutoffs and triggers (AlignerGpu::SetQueriesData,
    // src/aligner_gpu.cpp:213-242), the shared-delimiter concatenation and
    // its hash lists
    h->resident = false;
    const QuerySet &q = queries->q;
    const SearchParameters &P = h->opt.search;
    const KarlinParameters gk = Statistics::GappedBlosum62();
    const int presearch_cutoff = Statistics::Normalized2Nominal(P.normalized_presearch_cutoff, gk);
    std::vector<int> cut(q.size()), trig(q.size());
    for (size_t k = 0; k < q.size(); ++k) {
      KarlinParameters kp = {q.lambda[k], q.K[k], 0};
      cut[k] = Statistics::NormalizedCutoff2NominalCutoff(P.normalized_ungapped_cutoff, kp);
      trig[k] = Statistics::Normalized2Nominal(P.normalized_gapped_trigger, kp);
    }
    ConcatenatedOffsets(q, &h->coff, &h->concat);
    const int r = h->dev->BeginQueryChunk(q, h->concat, h->coff, cut, trig, presearch_cutoff,
                                          &h->stats);
    if (r != 0) {
      SetError(std::string("gz_seed_searcher_reset: ") + hipGetErrorString((hipError_t)r));
      return GZ_ERROR_HIP;
    }
    h->queries = queries;
  }
  h->chunk = chunk;
  h->resident = true;
  if (number_hash_lists) *number_hash_lists = h->dev->QueryLists().nlists;
  return GZ_OK;
}

gz_status gz_seed_searcher_search_seeds_from_representatives_for_similarity_filtering(
    gz_seed_searcher_t h, uint32_t list, uint32_t *query_concatenated_positions,
    uint32_t *database_positions, size_t capacity, size_t *count) {
  GZ_REQUIRE(h && count, "representative search: null argument");
  GZ_REQUIRE(h->resident, "representative search: Reset first");
  const DeviceQueryLists &L = h->dev->QueryLists();
  GZ_REQUIRE(list < L.nlists, "representative search: no such list");
  *count = 0;
  const DbChunk &c = h->db->db.chunks[h->chunk];
  if (!c.clustering) return GZ_OK;
  // the list's hash and its entries (query concatenated positions, check
  // flags), read from the device lists
  uint32_t hash = 0, range[2] = {0, 0};
  GZ_HIP_CHECK(hipMemcpy(&hash, L.hash + list, 4, hipMemcpyDeviceToHost));
  GZ_HIP_CHECK(hipMemcpy(range, L.entry_off + list, 8, hipMemcpyDeviceToHost));
  const uint32_t ne = range[1] - range[0];
  std::vector<uint32_t> pos(ne);
  std::vector<uint8_t> check(ne);
  if (ne) {
    GZ_HIP_CHECK(hipMemcpy(pos.data(), L.entry_pos + range[0], ne * 4, hipMemcpyDeviceToHost));
    GZ_HIP_CHECK(hipMemcpy(check.data(), L.entry_check + range[0], ne, hipMemcpyDeviceToHost));
  }
  const auto it = std::lower_bound(c.hashes.begin(), c.hashes.end(), hash);
  if (it == c.hashes.end() || *it != hash) return GZ_OK;
  const size_t hi = (size_t)(it - c.hashes.begin());
  // SearchSeedsFromRepresentativesForSimilarityFiltering
  // (src/seed_searcher_gpu.cpp:187-233): every representative against every
  // checked entry, representatives outer
  size_t n = 0;
  for (uint32_t r = c.rep_off[hi]; r < c.rep_off[hi + 1]; ++r)
    for (uint32_t k = 0; k < ne; ++k)
      if (check[k]) ++n;
  *count = n;
  if (n > capacity) {
    SetError("representative search: more pairs than the capacity (count holds the need)");
    return GZ_ERROR_CAPACITY;
  }
  GZ_REQUIRE(n == 0 || (query_concatenated_positions && database_positions),
             "representative search: null output");
  size_t o = 0;
  for (uint32_t r = c.rep_off[hi]; r < c.rep_off[hi + 1]; ++r)
    for (uint32_t k = 0; k < ne; ++k)
      if (check[k]) {
        query_concatenated_positions[o] = pos[k];
        database_positions[o] = c.rep_pos[r];
        ++o;
      }
  return GZ_OK;
}

gz_status gz_seed_searcher_search_seeds_from_clustering_subsequence(gz_seed_searcher_t h,
                                                                    uint32_t query_begin,
                                                                    uint32_t query_end,
                                                                    gz_seed_hits *out) {
  return Search(h, query_begin, query_end, 1u, out);
}

gz_status gz_seed_searcher_search_seeds_from_non_clustering_subsequence(gz_seed_searcher_t h,
                                                                        uint32_t query_begin,
                                                                        uint32_t query_end,
                                                                        gz_seed_hits *out) {
  return Search(h, query_begin, query_end, 2u, out);
}

gz_status gz_seed_searcher_search_seeds(gz_seed_searcher_t h, uint32_t query_begin,
                                        uint32_t query_end, gz_seed_hits *out) {
  return Search(h, query_begin, query_end, 3u, out);
}

}  // extern "C"
