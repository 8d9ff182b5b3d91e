/*
 * rng.c -- the reference's deterministic input generators, restated for the
 * oracle and the seeded test/benchmark inputs.  Test infrastructure only.
 */
#include <stdlib.h>
#include <string.h>

#include "rs_oracle.h"

/* qc_mix: SplitMix64 step (src/tests/mod.rs:37-43). */
uint64_t ora_qc_mix(uint64_t* seed) {
  *seed += 0x9e3779b97f4a7c15ull;
  uint64_t z = *seed;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* qc_seed (src/tests/mod.rs:45-52). */
uint64_t ora_qc_seed(const uint64_t* parts, size_t n) {
  uint64_t seed = 0x6a09e667f3bcc909ull;
  for (size_t i = 0; i < n; ++i) {
    seed ^= parts[i] + 0x9e3779b97f4a7c15ull;
    (void)ora_qc_mix(&seed);
  }
  return seed;
}

/* deterministic_fill_u8 (src/tests/mod.rs:80-85). */
void ora_qc_fill(uint8_t* out, size_t n, uint64_t seed) {
  for (size_t i = 0; i < n; ++i) out[i] = (uint8_t)ora_qc_mix(&seed);
}

/* deterministic_corrupt_positions: partial Fisher-Yates (src/tests/mod.rs:67-78). */
void ora_corrupt_positions(size_t n, size_t total, uint64_t seed, size_t* out) {
  if (n > total) n = total;
  size_t* p = (size_t*)malloc((total ? total : 1) * sizeof(size_t));
  for (size_t i = 0; i < total; ++i) p[i] = i;
  for (size_t i = 0; i < n; ++i) {
    size_t remaining = total - i;
    size_t j = i + (size_t)(ora_qc_mix(&seed) % remaining);
    size_t t = p[i];
    p[i] = p[j];
    p[j] = t;
  }
  memcpy(out, p, n * sizeof(size_t));
  free(p);
}

/* rand 0.10 SmallRng = Xoshiro256++; seed_from_u64 fills the 256-bit state with
 * SplitMix64 outputs; random::<u8>() takes the low byte of next_u32(), which is
 * the high half of next_u64().  Used by benches/common/mod.rs:224-236 and
 * tests/golden_vectors.rs:31-34. */
static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void ora_smallrng_fill(uint64_t seed, uint8_t* out, size_t n) {
  uint64_t s[4];
  uint64_t st = seed;
  for (int i = 0; i < 4; ++i) {
    st += 0x9e3779b97f4a7c15ull;
    uint64_t z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    s[i] = z ^ (z >> 31);
  }
  for (size_t i = 0; i < n; ++i) {
    uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    out[i] = (uint8_t)(result >> 32);
  }
}

/* derived_seed (benches/common/mod.rs:10,205-222). op tags: encode 0x11,
 * verify 0x22, reconstruct 0x33, reconstruct_data 0x44, update 0x55. */
uint64_t ora_derived_seed(int op_tag, size_t k, size_t m, size_t shard_size) {
  const uint64_t base = 0x00EC5EED20260524ull;
  return base ^ (uint64_t)op_tag ^ ((uint64_t)k << 48) ^ ((uint64_t)m << 32) ^ (uint64_t)shard_size;
}

/* checksum_bytes / checksum_shards (tests/golden_vectors.rs:5-18). */
uint64_t ora_checksum_shards(const uint8_t* const* shards, const size_t* lens, size_t n) {
  uint64_t acc = 0xcbf29ce484222325ull;
  for (size_t s = 0; s < n; ++s) {
    uint64_t h = 0xcbf29ce484222325ull;
    for (size_t i = 0; i < lens[s]; ++i) h = (h ^ shards[s][i]) * 0x100000001b3ull;
    acc ^= h;
    acc *= 0x100000001b3ull;
  }
  return acc;
}
