// Duplicate-removal keys (sbrsolver_amd/csrc/dedup_keys.h), host build: over a structured set --
// every sequence of length <= 3 over NC coplanar ids (and both interaction types) for NT
// transmitters, interned into a trie the way the device does (a sequence's id is the index of the
// slot holding its exact (parent id, element) key), then paired with NR receivers -- trie keys and
// group keys are unique, group keys are never 0 and trie keys never the empty marker.
// Prints "<groups> <trie-key duplicates> <group-key duplicates> <marker hits>".
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <vector>

#include "dedup_keys.h"

int main(int argc, char** argv) {
  const int NT = argc > 1 ? atoi(argv[1]) : 48, NC = argc > 2 ? atoi(argv[2]) : 40, NR = argc > 3 ? atoi(argv[3]) : 12;
  std::map<unsigned long long, int64_t> trie;   // key -> slot (insertion index)
  std::vector<unsigned long long> trieKeys, groupKeys;
  long marker = 0;
  std::vector<std::pair<int64_t, int>> frontier;   // (sequence id, length)
  for (int t = 0; t < NT; t++) frontier.push_back({sbr::seq_root(t), 0});
  std::vector<int64_t> all;
  for (auto& f : frontier) all.push_back(f.first);
  for (size_t i = 0; i < frontier.size(); i++) {
    const auto [id, len] = frontier[i];
    if (len == 3) continue;
    for (int c = 0; c < NC; c++)
      for (int type = 0; type <= 2; type += 2) {
        const unsigned long long k = sbr::seq_key(id, sbr::seq_elem(c, type == 0 ? 0 : (c % 2) + 1));
        if (k == sbr::kSeqEmpty) marker++;
        auto it = trie.find(k);
        if (it == trie.end()) {
          it = trie.emplace(k, (int64_t)trie.size()).first;
          trieKeys.push_back(k);
          frontier.push_back({it->second, len + 1});
          all.push_back(it->second);
        }
      }
  }
  for (int64_t s : all)
    for (int r = 0; r < NR; r++) {
      const unsigned long long g = sbr::group_key(s, r);
      if (g == 0) marker++;
      groupKeys.push_back(g);
    }
  auto dups = [](std::vector<unsigned long long>& v) {
    std::sort(v.begin(), v.end());
    return (long)(v.size() - (std::unique(v.begin(), v.end()) - v.begin()));
  };
  const size_t n = groupKeys.size();
  const long dt = dups(trieKeys), dg = dups(groupKeys);
  printf("%zu %ld %ld %ld\n", n, dt, dg, marker);
  return (dt || dg || marker) ? 1 : 0;
}
