// Exact keys of duplicate removal.  RemoveDuplicatePath (RTSolver.cpp:435-459) keys a receiver's
// paths by their hitCoplanarId sequence.  The device gives every distinct (transmitter, sequence)
// prefix an id through a trie stored in a hash table: the id of a sequence is the table slot that
// holds the exact pair (id of the sequence without its last element, last element), so two different
// sequences never share an id and no hash collision can merge them.  A duplicate-removal group is
// then the exact pair (sequence id, receiver).  The hash below only spreads keys over table slots;
// the slots store the keys themselves.  Host-callable for tests.
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

// Marks an empty slot of the sequence table (no key has it: an element is never 0xffffffff).
constexpr unsigned long long kSeqEmpty = ~0ull;

// Id of the empty sequence of transmitter txi (negative; sequence ids of longer ones are slots >= 0).
SBR_HD int64_t seq_root(int txi) { return -1 - (int64_t)txi; }

// No id: the sequence table was full (the trace is redone with a larger one).
constexpr int64_t kSeqNone = INT64_MIN;

// Element of the sequence at interaction k: the coplanar id of the surface hit before it, and with
// typed keys (an extension) the interaction's type (1 reflection, 2 transmission) in bits 30-31.
// Coplanar ids are below 2^30 (checked when the trace starts).
SBR_HD uint32_t seq_elem(int32_t cop, int type) { return (uint32_t)cop | ((uint32_t)type << 30); }

// Key of a trie edge: (parent sequence id, element), exact.
SBR_HD unsigned long long seq_key(int64_t parent, uint32_t elem) {
  return ((unsigned long long)(uint32_t)(int32_t)parent << 32) | (unsigned long long)elem;
}

// Key of a duplicate-removal group: (sequence id, receiver), exact and never 0 (0 marks an empty slot
// of the group table).  Sequence ids >= 0 map to 1 .. 2^31, transmitters' empty sequences to 2^31 and
// above.
SBR_HD unsigned long long group_key(int64_t seq, int32_t r) {
  const uint32_t hi = seq >= 0 ? (uint32_t)seq + 1u : (uint32_t)(int32_t)seq;
  return ((unsigned long long)hi << 32) | (unsigned long long)(uint32_t)r;
}

}  // namespace sbr
