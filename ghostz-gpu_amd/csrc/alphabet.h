// alphabet.h -- protein alphabet, score matrices and reduced alphabets.
#pragma once

#include <stdint.h>
#include <stddef.h>
#include <vector>

#include "gz_common.h"

namespace gz {

const uint8_t *ProteinEncodeTable();
char ProteinDecode(uint8_t code);
void EncodeProtein(const char *s, size_t n, uint8_t *out);
const char *Blosum62Text();
// 25 x 25 matrix indexed [row code][column code].
bool ParseScoreMatrix(const char *text, std::vector<int> &matrix);
// map25[protein code] = reduced code; *max_regular = largest regular code.
bool ReducedCodeMap(const char *sets_text, uint8_t *map25, uint32_t *max_regular);

// Default alphabet sets of `db` (src/database_build_main.cpp:60-62).
constexpr const char *kDefaultAlphabetSets = "A KR EDNQ C G H ILVM FYW P ST";

}  // namespace gz
