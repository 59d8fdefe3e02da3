// shim_driver.cpp -- drives the reference-named shim classes
// (reference_shims.h) the way AlignerGpuPresearchThread does
// (src/aligner_gpu_presearch_thread.cpp:150-410): per database chunk, Reset
// the seed searcher, search the seeds, run distance calculations, ungapped
// extensions with trigger and gapped extensions through the
// GpuStreamController, and print every result for the test to check against
// the oracle.
//
//   shim_driver QUERIES.fasta DB.fasta TRANSLATED(0|1) OUT.txt [CHUNK_RESIDUES]
//
// Output lines (positions in the shared-delimiter query concatenation "qc"
// or query-local "q", database chunk positions "d"):
//   C chunk lists
//   H chunk query q d                      seed search hit, enumeration order
//   D chunk list qc d distance             representative distance pair
//   U chunk qc d flag                      ungapped extension with trigger
//   G chunk reverse qc d score end_qc end_d  one-sided gapped extension
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "reference_shims.h"

using shim::Check;

namespace {

struct DeviceArray {
  void *p = nullptr;
  explicit DeviceArray(size_t bytes) { Check(gz_device_malloc(&p, bytes ? bytes : 16)); }
  ~DeviceArray() { gz_device_free(p); }
};

// pack host codes on the device
uint64_t *Pack(const std::vector<uint8_t> &codes, std::vector<DeviceArray *> *keep,
               gz_stream_t s) {
  DeviceArray *raw = new DeviceArray(codes.size());
  DeviceArray *packed = new DeviceArray(gz_packed_words(codes.size()) * 8);
  keep->push_back(raw);
  keep->push_back(packed);
  Check(gz_memcpy_async(raw->p, codes.data(), codes.size(), s));
  Check(gz_pack_sequence_async((const uint8_t *)raw->p, codes.size(), (uint64_t *)packed->p, s));
  Check(gz_stream_synchronize(s));
  return (uint64_t *)packed->p;
}

template <class T>
T *Upload(const std::vector<T> &v, std::vector<DeviceArray *> *keep, gz_stream_t s) {
  DeviceArray *a = new DeviceArray(v.size() * sizeof(T));
  keep->push_back(a);
  if (!v.empty()) Check(gz_memcpy_async(a->p, v.data(), v.size() * sizeof(T), s));
  Check(gz_stream_synchronize(s));
  return (T *)a->p;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 5 && argc != 6) {
    fprintf(stderr, "usage: %s QUERIES DB TRANSLATED OUT [CHUNK_RESIDUES]\n", argv[0]);
    return 2;
  }
  Check(gz_set_device(0));
  FILE *out = fopen(argv[4], "w");
  if (!out) return 2;
  gz_query_options qo = {atoi(argv[3]), 1, 0};
  gz_queries_t queries = nullptr;
  Check(gz_queries_read_fasta(argv[1], &qo, &queries));
  gz_database_options dbo = {1, 10, 39, argc == 6 ? strtoull(argv[5], nullptr, 10) : 0, 0};
  gz_database_t db = nullptr;
  Check(gz_database_read_fasta(argv[2], &dbo, &db));
  gz_queries_view_t qv;
  Check(gz_queries_view(queries, &qv));
  gz_database_info_t info;
  Check(gz_database_info(db, &info));
  const uint32_t nq = (uint32_t)qv.number_queries;

  // the shared-delimiter concatenation of the queries (AlignerGpu::
  // SetQueriesData): each query's framed codes without its closing delimiter
  std::vector<uint64_t> coff(nq);
  std::vector<uint8_t> concat;
  for (uint32_t i = 0; i < nq; ++i) {
    coff[i] = concat.size();
    concat.insert(concat.end(), qv.codes + qv.offsets[i], qv.codes + qv.offsets[i + 1] - 1);
  }
  concat.push_back(25);
  std::vector<int> cutoffs(nq), triggers(nq);
  int gapped_cutoff = 0;
  Check(gz_queries_cutoffs(queries, nullptr, cutoffs.data(), triggers.data(), &gapped_cutoff));
  std::vector<int> matrix(25 * 25);
  uint32_t nl = 0;
  Check(gz_score_matrix_blosum62(matrix.data(), &nl));

  gz_stream_t stream = nullptr;
  Check(gz_stream_create(&stream));
  std::vector<DeviceArray *> keep;
  const uint64_t *d_queries = Pack(concat, &keep, stream);
  const int *d_matrix = Upload(matrix, &keep, stream);
  const int *d_cutoffs = Upload(cutoffs, &keep, stream);
  const int *d_triggers = Upload(triggers, &keep, stream);
  std::vector<uint8_t> rmap(info.reduced_code_map, info.reduced_code_map + 25);
  const uint8_t *d_rmap = Upload(rmap, &keep, stream);

  shim::DistanceCalculatorGpu distance;
  shim::UngappedExtenderGpu ungapped;
  shim::GappedExtenderGpu gapped;
  distance.SetQueries(d_queries);
  distance.SetReducedCodeMap(d_rmap);
  distance.SetSubsequenceLength(info.subsequence_length);
  ungapped.SetQueries(25, d_queries, d_cutoffs, d_triggers);
  ungapped.SetScoreMatrix(d_matrix, nl);
  gapped.SetQueries(25, d_queries);
  gapped.SetScoreParameters(d_matrix, nl, 11, 1, gapped_cutoff);
  uint32_t max_length = 0;
  for (uint32_t i = 0; i < nq; ++i)
    max_length = std::max<uint32_t>(max_length, (uint32_t)(qv.offsets[i + 1] - qv.offsets[i]));
  Check(gz_gapped_extender_set_max_sequence_length(gapped.handle(), max_length));

  const size_t kBuffer = 1 << 16;
  shim::GpuStreamController controller;
  controller.Init(kBuffer);
  controller.SetDistanceCalculator(&distance);
  controller.SetUngappedExtender(&ungapped);
  controller.SetGappedExtender(&gapped);
  shim::HostSeedsMemory memory0(0, kBuffer), memory1(1, kBuffer);
  shim::HostSeedsMemory *memories[2] = {&memory0, &memory1};

  shim::SeedSearcherGpu searcher;
  for (uint32_t ci = 0; ci < info.number_chunks; ++ci) {
    gz_database_chunk_view_t cv;
    Check(gz_database_chunk(db, ci, &cv));
    std::vector<uint8_t> chunk_codes(cv.sequence, cv.sequence + cv.sequence_length);
    const uint64_t *d_db = Pack(chunk_codes, &keep, stream);
    distance.SetDatabase(d_db);
    ungapped.SetDatabase(d_db);
    gapped.SetDatabase(d_db);

    searcher.Reset(queries, db, ci);
    const uint32_t lists = searcher.GetNumberHashPositionDataLists();
    fprintf(out, "C %u %u\n", ci, lists);

    // seed search over every query, in two query ranges to cover the range
    // arguments (the first ends at a multiple of 64)
    std::vector<std::vector<shim::SeedSearcherGpu::Hit> > hits;
    const uint32_t split = std::min<uint32_t>(nq, 64);
    searcher.SearchSeeds(0, split, &hits);
    if (split < nq) searcher.SearchSeeds(split, nq, &hits);
    std::vector<std::pair<uint32_t, uint32_t> > flat;  // (qc, d)
    for (uint32_t i = 0; i < nq && i < hits.size(); ++i)
      for (const auto &h : hits[i]) {
        fprintf(out, "H %u %u %u %u\n", ci, i, h.query_sequence_position,
                h.database_sequence_position);
        flat.push_back({(uint32_t)(coff[i] + h.query_sequence_position),
                        h.database_sequence_position});
      }

    // distance pairs of the first lists, through the controller, two
    // resources alternating as the presearch thread uses them
    int mem = 0;
    for (uint32_t li = 0; li < std::min<uint32_t>(lists, 300); ++li) {
      std::vector<uint32_t> qp, dp;
      searcher.SearchSeedsFromRepresentativesForSimilarityFiltering(li, &qp, &dp);
      for (size_t b = 0; b < qp.size(); b += kBuffer) {
        const size_t n = std::min(kBuffer, qp.size() - b);
        shim::HostSeedsMemory &m = *memories[mem];
        memcpy(m.GetQueryConcatenatedPositions(), qp.data() + b, n * 4);
        memcpy(m.GetDatabasePositions(), dp.data() + b, n * 4);
        shim::GpuStreamController::DistanceCalculationSeedsParameters p = {n};
        const int id = controller.Run(p, m);
        controller.WaitRun(id);
        const char *dist = (const char *)m.GetResultValues();
        for (size_t k = 0; k < n; ++k)
          fprintf(out, "D %u %u %u %u %d\n", ci, li, qp[b + k], dp[b + k], (int)dist[k]);
        controller.ReleaseResource(id);
        mem ^= 1;
      }
    }

    // ungapped extension with trigger of every hit and of every hit shifted
    // by 7 on the database (mostly failing seeds): each seed is a run of
    // one query-seed entry
    {
      std::vector<uint32_t> sq, sd;
      for (const auto &f : flat) {
        sq.push_back(f.first);
        sd.push_back(f.second);
        if (f.second + 7 < cv.sequence_length) {
          sq.push_back(f.first);
          sd.push_back(f.second + 7);
        }
      }
      for (size_t b = 0; b < sq.size(); b += kBuffer) {
        const size_t n = std::min(kBuffer, sq.size() - b);
        // query seed data list: entry k = (query id, concatenated position)
        std::vector<uint32_t> qid(n), qpos(sq.begin() + b, sq.begin() + b + n);
        for (size_t k = 0; k < n; ++k)
          qid[k] = (uint32_t)(std::upper_bound(coff.begin(), coff.end(), (uint64_t)qpos[k]) -
                              coff.begin() - 1);
        std::vector<DeviceArray *> lk;
        ungapped.SetQuerySeedDataList(Upload(qid, &lk, stream), Upload(qpos, &lk, stream));
        shim::HostSeedsMemory &m = *memories[mem];
        for (size_t k = 0; k < n; ++k) {
          m.GetQuerySeedIds()[k] = (uint32_t)k;
          m.GetQuerySeedStarts()[k] = (uint32_t)k;
          m.GetDatabasePositions()[k] = sd[b + k];
        }
        shim::GpuStreamController::UngappedExtensionWithTriggerSeedsParameters p = {n, n};
        const int id = controller.Run(p, m);
        controller.WaitRun(id);
        const char *flag = (const char *)m.GetResultValues();
        for (size_t k = 0; k < n; ++k)
          fprintf(out, "U %u %u %u %d\n", ci, sq[b + k], sd[b + k], (int)flag[k]);
        controller.ReleaseResource(id);
        for (DeviceArray *a : lk) delete a;
        mem ^= 1;
      }
    }

    // gapped extension of both sides of every hit: reverse seeds start one
    // before the hit, forward seeds at it
    for (size_t b = 0; b < flat.size(); b += kBuffer / 2) {
      const size_t n = std::min(kBuffer / 2, flat.size() - b);
      shim::HostSeedsMemory &m = *memories[mem];
      for (size_t k = 0; k < n; ++k) {
        m.GetQuerySeedIds()[k] = (uint32_t)k;
        m.GetQueryConcatenatedPositions()[k] = flat[b + k].first - 1;
        m.GetDatabasePositions()[k] = flat[b + k].second - 1;
        m.GetQuerySeedIds()[n + k] = (uint32_t)k;
        m.GetQueryConcatenatedPositions()[n + k] = flat[b + k].first;
        m.GetDatabasePositions()[n + k] = flat[b + k].second;
      }
      shim::GpuStreamController::GappedExtensionSeesParameters p = {0, n, n, n};
      const int id = controller.Run(p, m);
      controller.WaitRun(id);
      for (size_t k = 0; k < 2 * n; ++k) {
        const int rev = k < n ? 1 : 0;
        const auto &f = flat[b + (k % n)];
        fprintf(out, "G %u %d %u %u %d %u %u\n", ci, rev, f.first - rev, f.second - rev,
                m.GetResultValues()[k], m.GetQueryConcatenatedPositions()[k],
                m.GetDatabasePositions()[k]);
      }
      controller.ReleaseResource(id);
      mem ^= 1;
    }
    uint32_t failures = 0;
    Check(gz_gapped_extender_failures(gapped.handle(), &failures));
    if (failures) {
      fprintf(stderr, "gapped failures %u\n", failures);
      return 1;
    }
  }
  controller.WaitRun();
  for (DeviceArray *a : keep) delete a;
  gz_stream_destroy(stream);
  gz_queries_destroy(queries);
  gz_database_destroy(db);
  fclose(out);
  printf("shim driver: %u queries, %u chunks\n", nq, info.number_chunks);
  return 0;
}
