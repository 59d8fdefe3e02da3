// gz_common.h -- constants and error plumbing shared by host and device code.
#pragma once

#include <stdint.h>
#include <stddef.h>
#include <string>

#include "ghostz_hip.h"

namespace gz {

// Protein alphabet: ARNDCQEGHILKMFPSTWYV = 0..19, BJZ* = 20..23, X = 24
// (src/protein_type.cpp:12-14, src/alphabet_coder.cpp:25-50); the sequence
// delimiter is max code + 1 (src/aligner_common.cpp:70-73).
constexpr uint32_t kNumberLetters = 25;
constexpr uint8_t kDelimiter = 25;
constexpr uint8_t kUnknownCode = 24;

// Packed alphabet: 5-bit codes, 12 per 64-bit word, code i of a word at bits
// [5i, 5i+5) (src/packed_alphabet_code.h:15-21).  Packed buffers carry
// kPackPadCodes delimiter codes in front so that kernels may read one word
// before any position; a position p lives at packed code index
// p + kPackPadCodes.
constexpr uint32_t kCodeBits = 5;
constexpr uint32_t kCodesPerWord = 12;
constexpr uint32_t kPackPadCodes = 48;
constexpr uint32_t kPackTailWords = 4;

// DP sentinel (src/gapped_extender.h:40).
constexpr int kInitScore = -(1 << 10);

// Score tables staged in LDS are 32 x 32 int8 (codes are 5-bit).
constexpr uint32_t kTableDim = 32;

inline size_t PackedWords(size_t codes) {
  return (codes + kPackPadCodes + kCodesPerWord - 1) / kCodesPerWord +
         kPackTailWords;
}

// Thread-local last error message; every C-ABI entry point sets it on
// failure and returns a non-zero gz_status.
void SetError(const std::string &msg);
const char *LastError();

}  // namespace gz

#define GZ_HIP_CHECK(call)                                                  \
  do {                                                                      \
    hipError_t _e = (call);                                                 \
    if (_e != hipSuccess) {                                                 \
      ::gz::SetError(std::string(#call) + ": " + hipGetErrorString(_e));    \
      return GZ_ERROR_HIP;                                                  \
    }                                                                       \
  } while (0)

#define GZ_REQUIRE(cond, msg)                                               \
  do {                                                                      \
    if (!(cond)) {                                                          \
      ::gz::SetError(msg);                                                  \
      return GZ_ERROR_INVALID_ARGUMENT;                                     \
    }                                                                       \
  } while (0)
