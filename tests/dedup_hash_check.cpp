// Uniqueness of the duplicate-removal group hashes (sbrsolver_amd/csrc/dedup_hash.h) over a
// structured key set: every (transmitter < NT, depth <= 2, coplanar ids < NC, receiver < NR).
// Structured keys are where a weak combination of the fields collides (a transmitter change
// cancelled by a coplanar-id change); random keys would not show it.  Prints
// "<keys> <slot-hash duplicates> <check-word duplicates>" and exits 1 on any duplicate.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "dedup_hash.h"

int main(int argc, char** argv) {
  const int NT = argc > 1 ? atoi(argv[1]) : 48, NC = argc > 2 ? atoi(argv[2]) : 40, NR = argc > 3 ? atoi(argv[3]) : 12;
  std::vector<unsigned long long> ha, hb;
  for (int t = 0; t < NT; t++)
    for (int depth = 0; depth <= 2; depth++) {
      int combos = 1;
      for (int k = 0; k < depth; k++) combos *= NC;
      for (int c = 0; c < combos; c++) {
        unsigned long long a, b;
        sbr::seq_seed(t, depth, &a, &b);
        int rest = c;
        for (int k = 0; k < depth; k++) {
          sbr::seq_fold(rest % NC, k, &a, &b);
          rest /= NC;
        }
        for (int r = 0; r < NR; r++) {
          unsigned long long x, y;
          sbr::group_hash(a, b, r, &x, &y);
          ha.push_back(x);
          hb.push_back(y);
        }
      }
    }
  auto dups = [](std::vector<unsigned long long>& v) {
    std::sort(v.begin(), v.end());
    return (long)(v.size() - (std::unique(v.begin(), v.end()) - v.begin()));
  };
  const size_t n = ha.size();
  const long da = dups(ha), db = dups(hb);
  printf("%zu %ld %ld\n", n, da, db);
  return (da || db) ? 1 : 0;
}
