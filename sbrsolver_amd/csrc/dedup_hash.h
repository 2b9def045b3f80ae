// Hash of a duplicate-removal group: (transmitter, receiver, coplanar-id sequence of the path).
// RemoveDuplicatePath (RTSolver.cpp:435-459) keys a receiver's paths by their hitCoplanarId
// sequence; the device table keys by two independent 64-bit hashes of it (slot and check word),
// so that a collision of the slot hash is detected instead of merging two groups.  Host-callable
// for tests.
#pragma once

#include <cstdint>

#include "sbrsolver/sbr_math.h"

namespace sbr {

SBR_HD unsigned long long mix64(unsigned long long x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

// Seed per (transmitter, depth).  It is mixed before the ids are folded in: with a plain xor, a
// change of transmitter in the low bits could be cancelled by a change of the first coplanar id.
SBR_HD void seq_seed(int txi, int depth, unsigned long long* a, unsigned long long* b) {
  *a = mix64(0x243f6a8885a308d3ull + ((unsigned long long)depth << 48) + (unsigned long long)txi * 0xa0761d6478bd642full);
  *b = mix64(0x13198a2e03707344ull ^
             ((unsigned long long)depth * 0x51ull + ((unsigned long long)txi << 20)) * 0xe7037ed1a0b428dbull);
}

// Folds the coplanar id of the k-th interaction (root first).
SBR_HD void seq_fold(int32_t cop, int k, unsigned long long* a, unsigned long long* b) {
  *a = mix64(*a + (unsigned long long)(uint32_t)cop + 0x9e3779b97f4a7c15ull);
  *b = mix64(*b ^ ((unsigned long long)(uint32_t)cop * 0xd6e8feb86659fd93ull + (unsigned long long)k));
}

// Slot hash and check word of the group for receiver r (never zero: zero marks an empty slot).
SBR_HD void group_hash(unsigned long long a, unsigned long long b, int32_t r, unsigned long long* hA,
                       unsigned long long* hB) {
  *hA = mix64(a ^ ((unsigned long long)(uint32_t)r * 0x9e3779b97f4a7c15ull));
  *hB = mix64(b + ((unsigned long long)(uint32_t)r << 7) + 0x632be59bd9b4e019ull);
  if (*hA == 0) *hA = 1;
  if (*hB == 0) *hB = 1;
}

}  // namespace sbr
