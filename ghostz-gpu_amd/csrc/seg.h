// seg.h -- SEG low-complexity filter for query preparation.
#pragma once

#include <string>
#include <vector>

namespace gz {

class Seg {
 public:
  explicit Seg(int window = 12);
  // Returns 0 and the masked sequence ('x' over low-complexity segments, the
  // rest upper-cased); returns 1 when the sequence is shorter than a window
  // (SequenceSegFilter::Filter, src/sequence_seg_filter.cpp:16-22).
  int Mask(const std::string &seq, std::string *masked) const;

 private:
  struct Window;
  struct Segment {
    int begin;
    int end;
  };
  static constexpr int kNonAminoAcid = 20;
  static constexpr double kLowCut = 2.2;
  static constexpr double kHighCut = 2.5;
  static constexpr int kMaxTrim = 100;

  double Entropy(const std::vector<int> &state) const;
  void SetComposition(Window *w) const;
  void SetState(Window *w) const;
  bool Build(Window &parent, int start, int length, Window *w) const;
  bool Shift(Window *w) const;
  double LnProbability(const int *state, int total) const;
  void Trim(Window &seq, int *left_end, int *right_end) const;
  void Segments(Window &seq, int offset, std::vector<Segment> &out) const;

  int window_;
  std::vector<int> aa_code_;
  std::vector<double> entropies_;
};

double SegLnFactorial(int n);

}  // namespace gz
