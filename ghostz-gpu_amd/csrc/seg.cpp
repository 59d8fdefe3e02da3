// seg.cpp -- SEG low-complexity masking of query sequences.
//
// Wootton & Federhen's SEG as the reference runs it (ext/seg/src/seg.cpp,
// window 12, K1 = 2.2 bits, K2 = 2.5 bits, at most 100 residues trimmed),
// restated: per-position window entropies, low-entropy seeds grown to the
// K2 boundary, each segment trimmed to its least probable sub-window, left
// remainders re-segmented recursively, masked residues replaced by 'x'.
// The log-factorial table is ln(n!) rounded to 6 decimals, as the
// reference's table holds it.
#include "seg.h"

#include <algorithm>
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>

namespace gz {

namespace {

const char kAminoAcids[] = "ACDEFGHIKLMNPQRSTVWY";
const double kLn2 = 0.69314718055994530;
const double kLn20 = 2.9957322735539909;
constexpr int kFactorialTable = 10240;

struct LnFactorial {
  std::vector<double> v;
  LnFactorial() : v(kFactorialTable) {
    char buf[64];
    for (int n = 0; n < kFactorialTable; ++n) {
      snprintf(buf, sizeof(buf), "%.6f", std::lgamma((double)n + 1.0));
      v[n] = strtod(buf, nullptr);
    }
  }
};
const std::vector<double> &LnFact() {
  static LnFactorial t;
  return t.v;
}

}  // namespace

double SegLnFactorial(int n) { return LnFact()[n]; }

struct Seg::Window {
  char *seq;
  int start;
  int length;
  int parent_length;
  std::vector<int> state;        // descending non-zero counts, 0-terminated
  std::vector<int> composition;  // per amino acid
  double entropy;
};

Seg::Seg(int window) : window_(window) {
  aa_code_.assign(CHAR_MAX, kNonAminoAcid);
  for (int i = 0; kAminoAcids[i]; ++i) aa_code_[(unsigned char)kAminoAcids[i]] = i;
  entropies_.resize(window_ + 1);
  const double w = window_;
  for (int i = 1; i <= window_; ++i) {
    const double x = i / w;
    entropies_[i] = -x * std::log(x) / kLn2;
  }
}

double Seg::Entropy(const std::vector<int> &state) const {
  int total = 0, n = 0;
  for (n = 0; state[n] != 0; ++n) total += state[n];
  double ent = 0.0;
  if (total == window_) {
    for (int i = 0; i < n; ++i) ent += entropies_[state[i]];
    return ent;
  }
  if (total == 0) return 0.0;
  const double r = 1. / (double)total;
  for (int i = 0; i < n; ++i) {
    const double v = state[i];
    ent += v * std::log(v * r);
  }
  return -ent * r / kLn2;
}

void Seg::SetComposition(Window *w) const {
  w->composition.resize(20, 0);
  for (int i = 0; i < w->length; ++i) {
    const int aa = w->seq[i];
    if (aa_code_[aa] != kNonAminoAcid) ++w->composition[aa_code_[aa]];
  }
}

void Seg::SetState(Window *w) const {
  if (w->composition.empty()) SetComposition(w);
  w->state.resize(21, 0);
  int n = 0;
  for (int aa = 0; aa < 20; ++aa)
    if (w->composition[aa] != 0) w->state[n++] = w->composition[aa];
  w->state[n] = 0;
  std::sort(w->state.begin(), w->state.begin() + n, std::greater<int>());
  w->entropy = Entropy(w->state);
}

bool Seg::Build(Window &parent, int start, int length, Window *w) const {
  if (start < 0 || length < 0 || start + length > parent.length) return false;
  w->start = start;
  w->length = length;
  w->seq = parent.seq + start;
  w->parent_length = parent.length;
  w->entropy = -2.0;
  w->composition.clear();
  w->state.clear();
  SetState(w);
  return true;
}

bool Seg::Shift(Window *w) const {
  const int length = w->length;
  if ((++w->start + length) > w->parent_length) {
    --w->start;
    return false;
  }
  int *comp = &w->composition[0];
  int *st = &w->state[0];
  const int first = w->seq[0];
  if (aa_code_[first] != kNonAminoAcid) {
    const int v = comp[aa_code_[first]];
    for (int i = 0; st[i] != 0; ++i)
      if (st[i] == v && st[i + 1] < v) {
        --st[i];
        break;
      }
    --comp[aa_code_[first]];
  }
  const int last = w->seq[length];
  ++w->seq;
  if (aa_code_[last] != kNonAminoAcid) {
    const int v = comp[aa_code_[last]];
    for (int i = 0;; ++i)
      if (st[i] == v) {
        ++st[i];
        break;
      }
    ++comp[aa_code_[last]];
  }
  if (w->entropy > -2.0) w->entropy = Entropy(w->state);
  return true;
}

double Seg::LnProbability(const int *state, int total) const {
  const std::vector<double> &f = LnFact();
  // number of ways to assign the classes of equal counts to 20 letters
  double ass = f[20];
  if (state[0] != 0) {
    int left = 20, cls = 1, prev = state[0];
    for (int i = 1; i < 20; ++i) {
      const int v = state[i];
      if (v == prev) {
        ++cls;
      } else {
        left -= cls;
        ass -= f[cls];
        if (v == 0) {
          cls = left;
          break;
        }
        cls = 1;
      }
      prev = v;
    }
    ass -= f[cls];
  }
  double perm = f[total];
  for (int i = 0; state[i] != 0; ++i) perm -= f[state[i]];
  return ass + perm - ((double)total) * kLn20;
}

void Seg::Trim(Window &seq, int *left_end, int *right_end) const {
  int best_left = 0, best_right = seq.length - 1;
  int min_length = 1;
  if ((seq.length - kMaxTrim) > min_length) min_length = seq.length - kMaxTrim;
  double min_prob = 1.0;
  for (int len = seq.length; len > min_length; --len) {
    Window w;
    Build(seq, 0, len, &w);
    int i = 0;
    bool shift = true;
    while (shift) {
      const double p = LnProbability(&w.state[0], len);
      if (p < min_prob) {
        min_prob = p;
        best_left = i;
        best_right = len + i - 1;
      }
      shift = Shift(&w);
      ++i;
    }
  }
  *left_end += best_left;
  *right_end -= (seq.length - best_right - 1);
}

void Seg::Segments(Window &seq, int offset, std::vector<Segment> &out) const {
  if (window_ > seq.length) return;
  // entropy of the window starting at each position (the last window is
  // repeated where no full window starts)
  std::vector<double> H(seq.length, -1);
  {
    Window w;
    Build(seq, 0, window_, &w);
    w.entropy = Entropy(w.state);
    for (int i = 0; i <= seq.length - 1; ++i) {
      H[i] = w.entropy;
      Shift(&w);
    }
  }
  const int last = seq.length - 1;
  int left_limit = 0;
  for (int i = 0; i <= last; ++i) {
    if (H[i] <= kLowCut && H[i] != -1) {
      int left_i, right_i;
      for (left_i = i; left_i >= left_limit; --left_i)
        if (H[left_i] == -1 || H[left_i] > kHighCut) break;
      ++left_i;
      for (right_i = i; right_i <= last; ++right_i)
        if (H[right_i] == -1 || H[right_i] > kHighCut) break;
      --right_i;
      int left_end = left_i, right_end = right_i;
      Window w;
      Build(seq, left_end, right_end - left_end + 1, &w);
      Trim(w, &left_end, &right_end);
      if (i < left_end) {
        const int lend = left_i, rend = left_end - 1;
        Window lw;
        Build(seq, lend, rend - lend + 1, &lw);
        Segments(lw, offset + lend, out);
      }
      out.push_back(Segment{left_end + offset, right_end + offset});
      i = std::min(right_i, right_end);
      left_limit = i + 1;
    }
  }
}

int Seg::Mask(const std::string &seq, std::string *masked) const {
  if (seq.size() < (size_t)window_) return 1;
  std::vector<char> copy(seq.begin(), seq.end());
  Window base;
  base.length = (int)seq.size();
  base.seq = copy.data();
  base.parent_length = base.length;
  base.start = 0;
  SetState(&base);
  std::vector<Segment> segs;
  Segments(base, 0, segs);
  // merge overlapping neighbours (minimum gap 0)
  if (!segs.empty()) {
    if (segs[0].begin < 0) segs[0].begin = 0;
    std::vector<Segment> merged;
    merged.push_back(segs[0]);
    for (size_t k = 1; k < segs.size(); ++k) {
      if (segs[k].begin - merged.back().end - 1 < 0) {
        merged.back().end = segs[k].end;
      } else {
        merged.push_back(segs[k]);
      }
    }
    if (base.length - merged.back().end - 1 < 0) merged.back().end = base.length - 1;
    segs.swap(merged);
  }
  for (int i = 0; i < base.length; ++i) copy[i] = (char)std::toupper((unsigned char)copy[i]);
  for (size_t k = 0; k < segs.size(); ++k)
    memset(copy.data() + segs[k].begin, 'x', segs[k].end - segs[k].begin + 1);
  masked->assign(copy.begin(), copy.end());
  return 0;
}

}  // namespace gz
