/*
 * ref_driver.cpp -- C ABI over the reference's own leaf sources.
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile, only when the
 * reference tree is present, together with the reference sources it calls
 * (compiled unmodified from where they lie); the output goes to oracle/_ref/.
 * Tests use it to pin the oracle restatement and the product's host-side
 * preparation (translation, SEG, Karlin, seed index) bit-exactly against the
 * reference on randomised inputs, and to write the golden fixtures under
 * tests/golden/.  Nothing here is shipped or measured.
 */
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>
#include <algorithm>
#include <map>

#include "alphabet_coder.h"
#include "dna_sequence.h"
#include "edit_blocks.h"
#include "gapped_extender.h"
#include "distance_calculator.h"
#include "protein_sequence.h"
#include "protein_type.h"
#include "reduced_alphabet_coder.h"
#include "reduced_alphabet_file_reader.h"
#include "reduced_alphabet_variable_hash_function.h"
#include "score_matrix.h"
#include "score_matrix_reader.h"
#include "seed_searcher_database_parameters.h"
#include "sequence_seg_filter.h"
#include "statistics.h"
#include "translator.h"
#include "ungapped_extender.h"
#include "../ext/karlin/src/karlin.h"

extern "C" {

/* ScoreMatrixReader::Read on a matrix text (score_matrix_reader.cpp). */
int gzr_read_matrix(const char *text, int *matrix_out, uint32_t *nl_out) {
  ProteinType t;
  std::istringstream in(text);
  ScoreMatrixReader r;
  std::vector<int> m;
  unsigned int nl = 0;
  r.Read(in, t, m, nl);
  for (size_t i = 0; i < m.size(); ++i) matrix_out[i] = m[i];
  *nl_out = nl;
  return 0;
}

/* AlphabetCoder(ProteinType)::Encode for every byte value. */
void gzr_protein_encode_table(uint8_t *table256) {
  ProteinType t;
  AlphabetCoder coder(t);
  for (int c = 0; c < 256; ++c) {
    table256[c] = (c < 128) ? coder.Encode((char)c) : 255;
  }
}

/* The hash and similarity reduced maps Aligner::BuildDatabaseParameters
 * derives (src/aligner.cpp:114-168) from an alphabet-set string. */
void gzr_reduced_map(const char *sets, uint8_t *map25, uint32_t *max_regular) {
  ProteinType t;
  AlphabetCoder coder(t);
  std::istringstream in(sets);
  std::vector<std::string> alphabet_sets;
  ReducedAlphabetFileReader reader;
  reader.Read(in, alphabet_sets);
  ReducedAlphabetCoder rc(t, alphabet_sets);
  for (AlphabetCoder::Code c = coder.GetMinCode(); c <= coder.GetMaxCode(); ++c)
    map25[c] = rc.Encode(coder.Decode(c));
  *max_regular = rc.GetMaxRegularLetterCode();
}

int gzr_ungapped_one_side(const uint8_t *s0, const uint8_t *s1, uint8_t delim,
                          int reversed, const int *matrix, uint32_t nl,
                          int cutoff) {
  ScoreMatrix sm("m", const_cast<int *>(matrix), nl);
  int best = 0;
  UngappedExtender::ExtendOneSideScoreOnly(s0, s1, delim, reversed != 0, sm,
                                           cutoff, &best);
  return best;
}

int gzr_ungapped_one_side_positions(const uint8_t *s0, const uint8_t *s1,
                                    uint8_t delim, int reversed,
                                    const int *matrix, uint32_t nl, int cutoff,
                                    int *p0, int *p1) {
  ScoreMatrix sm("m", const_cast<int *>(matrix), nl);
  int best = 0;
  UngappedExtender::ExtendOneSide(s0, s1, delim, reversed != 0, sm, cutoff,
                                  &best, p0, p1, NULL);
  return best;
}

int gzr_gapped_one_side(const uint8_t *s0, uint32_t s0len, const uint8_t *s1,
                        uint8_t delim, int reversed, const int *matrix,
                        uint32_t nl, int gap_open, int gap_ext, int cutoff,
                        int *score, int *p0, int *p1, uint8_t *edits,
                        uint32_t *nedits) {
  ScoreMatrix sm("m", const_cast<int *>(matrix), nl);
  GappedExtender e;
  if (edits == NULL) {
    e.ExtendOneSide(s0, s0len, s1, delim, reversed != 0, sm, gap_open, gap_ext,
                    cutoff, score, p0, p1, NULL);
    if (nedits) *nedits = 0;
    return 0;
  }
  EditBlocks eb;
  e.ExtendOneSide(s0, s0len, s1, delim, reversed != 0, sm, gap_open, gap_ext,
                  cutoff, score, p0, p1, &eb);
  std::vector<EditBlocks::EditOpType> v = eb.ToVector();
  for (size_t i = 0; i < v.size(); ++i) edits[i] = (uint8_t)v[i];
  *nedits = (uint32_t)v.size();
  return 0;
}

int gzr_distance(const uint8_t *q, uint32_t qc, const uint8_t *d, uint32_t dc,
                 uint32_t L, const uint8_t *map) {
  DistanceCalculator c;
  c.SetQueries(q);
  c.SetDatabase(d);
  c.SetReducedCodeMap(map);
  return c.CalculateDistance(qc, dc, L);
}

/* ReducedAlphabetVariableHashFunction built as the db builder does
 * (seed_searcher_database_parameters.cpp:41-78). */
int gzr_hash(const uint8_t *seq, const uint8_t *hash_map25,
             uint32_t max_indexing_code, int threshold, uint32_t max_length,
             const int *code_scores, uint32_t *hash, uint32_t *hlen) {
  std::vector<AlphabetCoder::Code> m(hash_map25, hash_map25 + 25);
  std::vector<int> cs(code_scores, code_scores + max_indexing_code + 1);
  ReducedAlphabetVariableHashFunction f(max_indexing_code, threshold,
                                        max_length, m, cs);
  return f.CalculateHash(seq, hash, hlen);
}

void gzr_karlin_blk(double *probs_at_min, int min, int max, float *lambda,
                    float *K, float *H) {
  BlastKarlinBlkCalc(probs_at_min - min, min, max, lambda, K, H);
}

/* Statistics::CalculateUngappedKarlinParameters for an encoded, framed
 * query (protein_query.cpp:52-61); returns the bool result. */
int gzr_query_karlin(const uint8_t *seq, uint32_t len, const int *matrix,
                     uint32_t nl, float *lambda, float *K, float *H) {
  ProteinType t;
  Statistics s(t);
  ScoreMatrix sm("BLOSUM62", const_cast<int *>(matrix), nl);
  Statistics::KarlinParameters p;
  bool ok = s.CalculateUngappedKarlinParameters(seq, len, sm, &p);
  *lambda = p.lambda;
  *K = p.K;
  *H = p.H;
  return ok ? 1 : 0;
}

void gzr_ideal_karlin(const int *matrix, uint32_t nl, float *lambda, float *K,
                      float *H) {
  ProteinType t;
  Statistics s(t);
  ScoreMatrix sm("BLOSUM62", const_cast<int *>(matrix), nl);
  Statistics::KarlinParameters p;
  s.CalculateUngappedIdealKarlinParameters(sm, &p);
  *lambda = p.lambda;
  *K = p.K;
  *H = p.H;
}

int gzr_normalized2nominal(float x, float lambda, float K) {
  Statistics::KarlinParameters p = {lambda, K, 0};
  return Statistics::Normalized2Nominal(x, p);
}
int gzr_normalized_cutoff2nominal(float x, float lambda, float K) {
  Statistics::KarlinParameters p = {lambda, K, 0};
  return Statistics::NormalizedCutoff2NominalCutoff(x, p);
}
float gzr_nominal2normalized(int s, float lambda, float K) {
  Statistics::KarlinParameters p = {lambda, K, 0};
  return Statistics::Nominal2Normalized(s, p);
}
double gzr_nominal2evalue(int s, uint64_t space, float lambda, float K) {
  Statistics::KarlinParameters p = {lambda, K, 0};
  return Statistics::Nominal2EValue(s, space, p);
}
uint64_t gzr_search_space(uint64_t qlen, uint64_t dblen, uint64_t dbn) {
  ProteinType t;
  Statistics s(t);
  Statistics::KarlinParameters p = {0.267f, 0.041f, 0.14f};
  return s.CalculateSearchSpace(qlen, dblen, dbn, p, 1.9, -30);
}

/* Translator::Translate: six frames, written back to back into out, with
 * frame lengths in lens[6]. */
int gzr_translate(const char *dna, char *out, uint32_t *lens) {
  Translator tr;
  DnaSequence d("q", dna);
  std::vector<ProteinSequence> p;
  tr.Translate(d, p);
  size_t o = 0;
  for (size_t i = 0; i < 6; ++i) {
    if (i < p.size()) {
      const std::string &s = p[i].GetSequenceData();
      memcpy(out + o, s.data(), s.size());
      lens[i] = (uint32_t)s.size();
      o += s.size();
    } else {
      lens[i] = 0;
    }
  }
  return (int)p.size();
}

/* SequenceSegFilter::Filter (window 12): returns the filter status;
 * out receives the masked sequence when the status is 0. */
int gzr_seg(const char *seq, char *out) {
  SequenceSegFilter f;
  std::string masked;
  int r = f.Filter(seq, &masked);
  if (r == 0) memcpy(out, masked.data(), masked.size());
  return r;
}

/* SeedSearcherDatabaseParameters::Build on one concatenated chunk; the
 * index is returned through its own Save() stream, re-read here into a
 * canonical form: per hash (ascending) the representatives (in stored
 * order) with their members, and the non-clustered values. */
struct gzr_index {
  std::vector<uint32_t> hashes, rep_off, rep_pos, mem_off, mem_pos, val_off,
      val_pos;
};

static void ReadDense(std::istream &is, std::map<uint32_t, uint32_t> &keys,
                      std::vector<uint32_t> &offsets, size_t value_size,
                      std::vector<char> &values) {
  size_t n = 0;
  is.read((char *)&n, sizeof(n));
  for (size_t i = 0; i < n; ++i) {
    uint32_t k, v;
    is.read((char *)&k, sizeof(k));
    is.read((char *)&v, sizeof(v));
    keys[k] = v;
  }
  size_t no = 0;
  is.read((char *)&no, sizeof(no));
  offsets.resize(no);
  is.read((char *)&offsets[0], sizeof(uint32_t) * no);
  size_t nv = 0;
  is.read((char *)&nv, sizeof(nv));
  values.resize(nv * value_size);
  if (nv) is.read(&values[0], value_size * nv);
}

static void SkipHashFunction(std::istream &is) {
  ReducedAlphabetVariableHashFunction f;
  f.Load(is);
}

static void IndexBuildParameters(int clustering, uint32_t L, int seed_threshold,
                                 const int *matrix, uint32_t nl, const char *sets,
                                 SeedSearcherDatabaseParameters::BuildParameters &bp) {
  ProteinType t;
  AlphabetCoder coder(t);
  bp.clustering = clustering != 0;
  bp.subsequence_length = L;
  bp.sequence_delimiter = coder.GetMaxCode() + 1;
  bp.seed_threshold = seed_threshold;
  bp.number_threads = 1;
  bp.score_matrix = ScoreMatrix("BLOSUM62", const_cast<int *>(matrix), nl);
  std::istringstream in(sets);
  std::vector<std::string> alphabet_sets;
  ReducedAlphabetFileReader reader;
  reader.Read(in, alphabet_sets);
  ReducedAlphabetCoder rc(t, alphabet_sets);
  bp.hash_code_map.resize(coder.GetMaxCode() + 1);
  bp.similairty_code_map.resize(coder.GetMaxCode() + 1);
  for (AlphabetCoder::Code c = coder.GetMinCode(); c <= coder.GetMaxCode(); ++c) {
    bp.hash_code_map[c] = rc.Encode(coder.Decode(c));
    bp.similairty_code_map[c] = rc.Encode(coder.Decode(c));
  }
  bp.max_indexing_code = rc.GetMaxRegularLetterCode();
}

/* The bytes SeedSearcherDatabaseParameters::Save (the chunk's .sdp file)
 * and SaveCommonParameters (its .scp file) write for one concatenated chunk
 * (malloc'd; gzr_free_bytes). */
void gzr_index_files(const uint8_t *seq, uint32_t len, int clustering, uint32_t L,
                     int seed_threshold, const int *matrix, uint32_t nl, const char *sets,
                     char **sdp, uint64_t *sdp_len, char **scp, uint64_t *scp_len) {
  SeedSearcherDatabaseParameters::BuildParameters bp;
  IndexBuildParameters(clustering, L, seed_threshold, matrix, nl, sets, bp);
  std::vector<uint8_t> copy(seq, seq + len);
  SeedSearcherDatabaseParameters p;
  p.Build(&copy[0], len, bp);
  std::stringstream a, b;
  p.Save(a);
  p.SaveCommonParameters(b);
  const std::string sa = a.str(), sb = b.str();
  *sdp = (char *)malloc(sa.size() + 1);
  *scp = (char *)malloc(sb.size() + 1);
  memcpy(*sdp, sa.data(), sa.size());
  memcpy(*scp, sb.data(), sb.size());
  *sdp_len = sa.size();
  *scp_len = sb.size();
}

void gzr_free_bytes(char *p) { free(p); }

void *gzr_build_index(const uint8_t *seq, uint32_t len, int clustering,
                      uint32_t L, int seed_threshold, const int *matrix,
                      uint32_t nl, const char *sets) {
  SeedSearcherDatabaseParameters::BuildParameters bp;
  IndexBuildParameters(clustering, L, seed_threshold, matrix, nl, sets, bp);
  std::vector<uint8_t> copy(seq, seq + len);
  SeedSearcherDatabaseParameters p;
  p.Build(&copy[0], len, bp);
  std::stringstream ss;
  p.Save(ss);
  bool cl;
  uint8_t dl;
  uint32_t sl;
  ss.read((char *)&cl, sizeof(cl));
  ss.read((char *)&dl, sizeof(dl));
  ss.read((char *)&sl, sizeof(sl));
  // sequences index
  SkipHashFunction(ss);
  std::map<uint32_t, uint32_t> vkeys;
  std::vector<uint32_t> voff;
  std::vector<char> vvals;
  ReadDense(ss, vkeys, voff, sizeof(uint32_t), vvals);
  // clustering index: hash function, reps (pair<cluster,pos>), members
  SkipHashFunction(ss);
  std::map<uint32_t, uint32_t> rkeys;
  std::vector<uint32_t> roff;
  std::vector<char> rvals;
  ReadDense(ss, rkeys, roff, sizeof(std::pair<uint32_t, uint32_t>), rvals);
  uint32_t max_key = 0;
  ss.read((char *)&max_key, sizeof(max_key));
  size_t nmo = 0;
  ss.read((char *)&nmo, sizeof(nmo));
  std::vector<uint32_t> moff(nmo);
  if (nmo) ss.read((char *)&moff[0], sizeof(uint32_t) * nmo);
  size_t nmv = 0;
  ss.read((char *)&nmv, sizeof(nmv));
  std::vector<uint32_t> mval(nmv);
  if (nmv) ss.read((char *)&mval[0], sizeof(uint32_t) * nmv);

  gzr_index *ix = new gzr_index();
  std::vector<uint32_t> all;
  for (std::map<uint32_t, uint32_t>::iterator it = vkeys.begin(); it != vkeys.end(); ++it)
    all.push_back(it->first);
  for (std::map<uint32_t, uint32_t>::iterator it = rkeys.begin(); it != rkeys.end(); ++it)
    all.push_back(it->first);
  std::sort(all.begin(), all.end());
  all.erase(std::unique(all.begin(), all.end()), all.end());
  const uint32_t *vv = (const uint32_t *)(vvals.empty() ? NULL : &vvals[0]);
  const std::pair<uint32_t, uint32_t> *rv =
      (const std::pair<uint32_t, uint32_t> *)(rvals.empty() ? NULL : &rvals[0]);
  ix->rep_off.push_back(0);
  ix->val_off.push_back(0);
  ix->mem_off.push_back(0);
  for (size_t i = 0; i < all.size(); ++i) {
    uint32_t h = all[i];
    ix->hashes.push_back(h);
    std::map<uint32_t, uint32_t>::iterator r = rkeys.find(h);
    if (r != rkeys.end()) {
      for (uint32_t k = roff[r->second]; k < roff[r->second + 1]; ++k) {
        ix->rep_pos.push_back(rv[k].second);
        uint32_t cid = rv[k].first;
        if (cid <= max_key)
          for (uint32_t m = moff[cid]; m < moff[cid + 1]; ++m)
            ix->mem_pos.push_back(mval[m]);
        ix->mem_off.push_back((uint32_t)ix->mem_pos.size());
      }
    }
    ix->rep_off.push_back((uint32_t)ix->rep_pos.size());
    std::map<uint32_t, uint32_t>::iterator v = vkeys.find(h);
    if (v != vkeys.end()) {
      for (uint32_t k = voff[v->second]; k < voff[v->second + 1]; ++k)
        ix->val_pos.push_back(vv[k]);
    }
    ix->val_off.push_back((uint32_t)ix->val_pos.size());
  }
  return ix;
}

/* sizes[0..6] = hashes, rep_off, rep_pos, mem_off, mem_pos, val_off, val_pos */
void gzr_index_sizes(void *h, uint64_t *sizes) {
  gzr_index *ix = (gzr_index *)h;
  sizes[0] = ix->hashes.size();
  sizes[1] = ix->rep_off.size();
  sizes[2] = ix->rep_pos.size();
  sizes[3] = ix->mem_off.size();
  sizes[4] = ix->mem_pos.size();
  sizes[5] = ix->val_off.size();
  sizes[6] = ix->val_pos.size();
}
void gzr_index_copy(void *h, uint32_t **arrays) {
  gzr_index *ix = (gzr_index *)h;
  std::vector<uint32_t> *v[7] = {&ix->hashes, &ix->rep_off, &ix->rep_pos,
                                 &ix->mem_off, &ix->mem_pos, &ix->val_off,
                                 &ix->val_pos};
  for (int i = 0; i < 7; ++i)
    if (!v[i]->empty()) memcpy(arrays[i], &(*v[i])[0], v[i]->size() * 4);
}
void gzr_index_free(void *h) { delete (gzr_index *)h; }

}  // extern "C"
