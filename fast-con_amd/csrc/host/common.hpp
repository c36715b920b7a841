// Host-side helpers shared by every translation unit of libfastcon:
// the per-thread error text behind fc_last_error(), and the counter-based
// random stream used by the generators and samplers.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>

#include "fastcon.h"

namespace fc {

// Records the message for fc_last_error() and returns the status, so call
// sites read `return fail(FC_ERR_INVALID, "...")`.
fc_status fail(fc_status st, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// SplitMix64 finaliser: a bijective 64-bit mixer (Steele, Lea, Flood 2014).
inline uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Random 64-bit word number `i` of stream (seed, lane): a pure function, so
// generation is reproducible at any thread count.
inline uint64_t stream_word(uint64_t seed, uint64_t lane, uint64_t i) {
  return mix64(mix64(seed ^ (lane * 0xD1B54A32D192ED03ull)) + i * 0x9E3779B97F4A7C15ull);
}

// Uniform double in [0,1) from the top 53 bits.
inline double unit_double(uint64_t w) { return (double)(w >> 11) * (1.0 / 9007199254740992.0); }

// Bijection of [0, 2^bits): alternating odd multiplications and xor-shifts,
// each invertible modulo 2^bits. Used to scramble R-MAT vertex ids.
struct Scrambler {
  uint64_t mask, m1, m2, a1, a2;
  int shift;
  Scrambler(int bits, uint64_t seed) {
    mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
    m1 = mix64(seed + 1) | 1ull;
    m2 = mix64(seed + 2) | 1ull;
    a1 = mix64(seed + 3);
    a2 = mix64(seed + 4);
    shift = bits / 2 > 0 ? bits / 2 : 1;
  }
  uint64_t operator()(uint64_t x) const {
    x = (x * m1 + a1) & mask;
    x ^= x >> shift;
    x = (x * m2 + a2) & mask;
    x ^= x >> shift;
    x = (x * m1 + a2) & mask;
    return x;
  }
};

}  // namespace fc
