// alphabet.cpp -- protein alphabet, BLOSUM62 and reduced alphabets.
//
// AlphabetCoder(ProteinType) (src/alphabet_coder.cpp:25-50,
// src/protein_type.cpp:12-14): A R N D C Q E G H I L K M F P S T W Y V = 0..19,
// B J Z * = 20..23, every other byte = X = 24; lower case maps like upper.
// ScoreMatrixReader (src/score_matrix_reader.cpp:57-96): a whitespace table
// whose row/column letters are placed by their codes; absent letters (J)
// score 0.  ReducedAlphabetCoder (src/reduced_alphabet_coder.cpp:32-98):
// each letter maps to the first letter of its set; codes are handed out in
// alphabet order to letters that represent themselves.
#include "alphabet.h"

#include <cctype>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

namespace gz {

namespace {
const char kRegular[] = "ARNDCQEGHILKMFPSTWYV";
const char kAmbiguous[] = "BJZ*";
const char kUnknown = 'X';

// NCBI BLOSUM62 in half bits (the standard public matrix).
const char kBlosum62[] =
    "   A  R  N  D  C  Q  E  G  H  I  L  K  M  F  P  S  T  W  Y  V  B  Z  X  *\n"
    "A  4 -1 -2 -2  0 -1 -1  0 -2 -1 -1 -1 -1 -2 -1  1  0 -3 -2  0 -2 -1  0 -4\n"
    "R -1  5  0 -2 -3  1  0 -2  0 -3 -2  2 -1 -3 -2 -1 -1 -3 -2 -3 -1  0 -1 -4\n"
    "N -2  0  6  1 -3  0  0  0  1 -3 -3  0 -2 -3 -2  1  0 -4 -2 -3  3  0 -1 -4\n"
    "D -2 -2  1  6 -3  0  2 -1 -1 -3 -4 -1 -3 -3 -1  0 -1 -4 -3 -3  4  1 -1 -4\n"
    "C  0 -3 -3 -3  9 -3 -4 -3 -3 -1 -1 -3 -1 -2 -3 -1 -1 -2 -2 -1 -3 -3 -2 -4\n"
    "Q -1  1  0  0 -3  5  2 -2  0 -3 -2  1  0 -3 -1  0 -1 -2 -1 -2  0  3 -1 -4\n"
    "E -1  0  0  2 -4  2  5 -2  0 -3 -3  1 -2 -3 -1  0 -1 -3 -2 -2  1  4 -1 -4\n"
    "G  0 -2  0 -1 -3 -2 -2  6 -2 -4 -4 -2 -3 -3 -2  0 -2 -2 -3 -3 -1 -2 -1 -4\n"
    "H -2  0  1 -1 -3  0  0 -2  8 -3 -3 -1 -2 -1 -2 -1 -2 -2  2 -3  0  0 -1 -4\n"
    "I -1 -3 -3 -3 -1 -3 -3 -4 -3  4  2 -3  1  0 -3 -2 -1 -3 -1  3 -3 -3 -1 -4\n"
    "L -1 -2 -3 -4 -1 -2 -3 -4 -3  2  4 -2  2  0 -3 -2 -1 -2 -1  1 -4 -3 -1 -4\n"
    "K -1  2  0 -1 -3  1  1 -2 -1 -3 -2  5 -1 -3 -1  0 -1 -3 -2 -2  0  1 -1 -4\n"
    "M -1 -1 -2 -3 -1  0 -2 -3 -2  1  2 -1  5  0 -2 -1 -1 -1 -1  1 -3 -1 -1 -4\n"
    "F -2 -3 -3 -3 -2 -3 -3 -3 -1  0  0 -3  0  6 -4 -2 -2  1  3 -1 -3 -3 -1 -4\n"
    "P -1 -2 -2 -1 -3 -1 -1 -2 -2 -3 -3 -1 -2 -4  7 -1 -1 -4 -3 -2 -2 -1 -2 -4\n"
    "S  1 -1  1  0 -1  0  0  0 -1 -2 -2  0 -1 -2 -1  4  1 -3 -2 -2  0  0  0 -4\n"
    "T  0 -1  0 -1 -1 -1 -1 -2 -2 -1 -1 -1 -1 -2 -1  1  5 -2 -2  0 -1 -1  0 -4\n"
    "W -3 -3 -4 -4 -2 -2 -3 -2 -2 -3 -2 -3 -1  1 -4 -3 -2 11  2 -3 -4 -3 -2 -4\n"
    "Y -2 -2 -2 -3 -2 -1 -2 -3  2 -1 -1 -2 -1  3 -3 -2 -2  2  7 -1 -3 -2 -1 -4\n"
    "V  0 -3 -3 -3 -1 -2 -2 -3 -3  3  1 -2  1 -1 -2 -2  0 -3 -1  4 -3 -2 -1 -4\n"
    "B -2 -1  3  4 -3  0  1 -1  0 -3 -4  0 -3 -3 -2  0 -1 -4 -3 -3  4  1 -1 -4\n"
    "Z -1  0  0  1 -3  3  4 -2  0 -3 -3  1 -1 -3 -1  0 -1 -3 -2 -2  1  4 -1 -4\n"
    "X  0 -1 -1 -1 -2 -1 -1 -1 -1 -1 -1 -1 -1 -1 -2  0  0 -2 -1 -1 -1 -1 -1 -4\n"
    "* -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4 -4  1\n";

struct EncodeTable {
  uint8_t t[256];
  EncodeTable() {
    for (int i = 0; i < 256; ++i) t[i] = kUnknownCode;
    std::string letters = std::string(kRegular) + kAmbiguous;
    for (size_t i = 0; i < letters.size(); ++i) {
      t[(uint8_t)letters[i]] = (uint8_t)i;
      t[(uint8_t)std::tolower(letters[i])] = (uint8_t)i;
    }
  }
};
const EncodeTable &Table() {
  static EncodeTable t;
  return t;
}
}  // namespace

const uint8_t *ProteinEncodeTable() { return Table().t; }

char ProteinDecode(uint8_t code) {
  static const std::string letters = std::string(kRegular) + kAmbiguous + kUnknown;
  return code < letters.size() ? letters[code] : kUnknown;
}

void EncodeProtein(const char *s, size_t n, uint8_t *out) {
  const uint8_t *t = ProteinEncodeTable();
  for (size_t i = 0; i < n; ++i) out[i] = t[(uint8_t)s[i]];
}

const char *Blosum62Text() { return kBlosum62; }

bool ParseScoreMatrix(const char *text, std::vector<int> &matrix) {
  matrix.assign(kNumberLetters * kNumberLetters, 0);
  std::istringstream in(text);
  std::string line;
  std::vector<char> columns;
  bool header = true;
  // a letter is placed when it is a known letter or the unknown letter X
  auto known = [](char c) {
    char u = (char)std::toupper((unsigned char)c);
    return u == kUnknown || Table().t[(uint8_t)u] != kUnknownCode;
  };
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream ls(line);
    std::vector<std::string> tok;
    std::string t;
    while (ls >> t) tok.push_back(t);
    if (tok.empty()) continue;
    if (header) {
      for (size_t i = 0; i < tok.size(); ++i) columns.push_back(tok[i][0]);
      header = false;
      continue;
    }
    const char row = tok[0][0];
    for (size_t i = 1; i < tok.size() && i - 1 < columns.size(); ++i) {
      const char col = columns[i - 1];
      if (known(row) && known(col)) {
        const uint8_t r = Table().t[(uint8_t)row], c = Table().t[(uint8_t)col];
        matrix[r * kNumberLetters + c] = std::atoi(tok[i].c_str());
      }
    }
  }
  return !header;
}

bool ReducedCodeMap(const char *sets_text, uint8_t *map25,
                    uint32_t *max_regular) {
  std::vector<std::string> sets;
  std::istringstream in(sets_text);
  std::string s;
  while (in >> s) {
    for (size_t i = 0; i < s.size(); ++i) s[i] = (char)std::toupper((unsigned char)s[i]);
    sets.push_back(s);
  }
  auto rep = [&](char c) {
    for (size_t i = 0; i < sets.size(); ++i)
      if (sets[i].find(c) != std::string::npos) return sets[i][0];
    return c;
  };
  const std::string letters = std::string(kRegular) + kAmbiguous + kUnknown;
  const size_t n_regular = sizeof(kRegular) - 1;
  uint8_t code_of[256];
  memset(code_of, 0xff, sizeof(code_of));
  uint32_t regular_codes = 0;
  uint8_t next = 0;
  for (size_t i = 0; i < letters.size(); ++i) {
    if (rep(letters[i]) == letters[i]) {
      code_of[(uint8_t)letters[i]] = next++;
      if (i < n_regular) ++regular_codes;
    }
  }
  for (size_t i = 0; i < letters.size(); ++i) {
    if (rep(letters[i]) != letters[i])
      code_of[(uint8_t)letters[i]] = code_of[(uint8_t)rep(letters[i])];
  }
  for (uint32_t c = 0; c < kNumberLetters; ++c) {
    map25[c] = code_of[(uint8_t)ProteinDecode((uint8_t)c)];
  }
  *max_regular = regular_codes - 1;
  return true;
}

}  // namespace gz

extern "C" {

void gz_protein_encode_table(uint8_t *table256) {
  memcpy(table256, gz::ProteinEncodeTable(), 256);
}

gz_status gz_score_matrix_blosum62(int *matrix, uint32_t *number_letters) {
  std::vector<int> m;
  if (!gz::ParseScoreMatrix(gz::Blosum62Text(), m)) {
    gz::SetError("BLOSUM62 parse failed");
    return GZ_ERROR_INVALID_ARGUMENT;
  }
  memcpy(matrix, m.data(), m.size() * sizeof(int));
  *number_letters = gz::kNumberLetters;
  return GZ_OK;
}

const char *gz_score_matrix_blosum62_text(void) { return gz::Blosum62Text(); }

gz_status gz_score_matrix_parse(const char *text, int *matrix,
                                uint32_t *number_letters) {
  std::vector<int> m;
  if (!text || !gz::ParseScoreMatrix(text, m)) {
    gz::SetError("score matrix parse failed");
    return GZ_ERROR_INVALID_ARGUMENT;
  }
  memcpy(matrix, m.data(), m.size() * sizeof(int));
  *number_letters = gz::kNumberLetters;
  return GZ_OK;
}

gz_status gz_reduced_code_map(const char *alphabet_sets, uint8_t *map25,
                              uint32_t *max_regular_code) {
  if (!alphabet_sets || !map25 || !max_regular_code) {
    gz::SetError("gz_reduced_code_map: null argument");
    return GZ_ERROR_INVALID_ARGUMENT;
  }
  gz::ReducedCodeMap(alphabet_sets, map25, max_regular_code);
  return GZ_OK;
}

}  // extern "C"
