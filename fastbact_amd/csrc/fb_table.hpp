// fb_table.hpp -- the global open-addressing solid table (device side).
#pragma once
#include "fb_device.hpp"
#include "fb_internal.hpp"

namespace fb {

// Finds or claims the slot of `key`.  Key words are write-once (empty ->
// value), claimed word by word with compare-and-swap in order; a slot whose
// prefix matches but whose next word holds another value belongs to another
// key.  A plain load that returns a stale empty word only sends the thread to
// the compare-and-swap, which returns the stored word.
template <int KW>
__device__ __forceinline__ uint32_t* table_find_global(const DevSolidTable& t, const uint64_t (&key)[KW], uint64_t h) {
  const uint64_t cmask = t.cap - 1ull;
  uint64_t slot = h & cmask;
  for (uint32_t probe = 0; probe < t.max_probe; ++probe) {
    uint64_t* s = t.slots + slot * t.stride;
    bool match = true;
#pragma unroll
    for (int j = 0; j < KW; ++j) {
      if (!match) break;
      uint64_t v = s[j];
      if (v == kEmpty) {
        uint64_t old = atomicCAS((unsigned long long*)(s + j), (unsigned long long)kEmpty,
                                 (unsigned long long)key[j]);
        if (old == kEmpty) {
          v = key[j];
          if (j == 0) atomicAdd(t.status + 1, 1u);
        } else {
          v = old;
        }
      }
      match = (v == key[j]);
    }
    if (match) return reinterpret_cast<uint32_t*>(s + KW);
    slot = (slot + 1ull) & cmask;
  }
  atomicOr(t.status, 1u);
  return nullptr;
}

// Slot hash of a key given as packed 62-bit words.
template <int KW>
__device__ __forceinline__ uint64_t slot_hash_words(const uint64_t (&w)[KW]) {
  uint64_t h = w[0];
#pragma unroll
  for (int j = 1; j < KW; ++j) h = fbk::splitmix64(h) ^ w[j];
  return fbk::splitmix64(h ^ 0x2545f4914f6cdd1dull);
}

}  // namespace fb
