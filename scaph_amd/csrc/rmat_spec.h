// rmat_spec.h — the synthetic-graph specification shared by the HIP
// generator kernels and the C++ host generator (both must produce the same
// edges bit for bit).  The reference ships no generator (its graphs come
// from an external toolchain), so this is the project's own definition:
//
//   word(key, i)  = i-th output (0-based) of SplitMix64 seeded with `key`
//                   (Steele, Lea, Flood, "Fast splittable pseudorandom number
//                   generators", OOPSLA 2014; finaliser constants as in
//                   Vigna's splitmix64.c).
//   R-MAT edge e of a scale-S graph uses words e*ceil(S/2) + j; each word
//   gives two u32 draws (low half first).  Draw l picks the quadrant of bit
//   S-1-l by integer compares against thr_a <= thr_ab <= thr_abc.
//   Vertex IDs are then mapped through perm(), a keyed bijection on S bits
//   (4 rounds of odd-multiply-add and xorshift, all invertible mod 2^S).
#ifndef SCAPH_RMAT_SPEC_H
#define SCAPH_RMAT_SPEC_H

#include <stdint.h>

#if defined(__HIPCC__)
#define SCAPH_HD __host__ __device__ __forceinline__
#else
#define SCAPH_HD static inline
#endif

namespace scaph_spec {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kPermSalt = 0x7065726D7574650Aull;  // "permute\n"

SCAPH_HD uint64_t fmix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// i-th SplitMix64 output for state `key` (state advances by kGolden first).
SCAPH_HD uint64_t word(uint64_t key, uint64_t i) { return fmix64(key + (i + 1) * kGolden); }

struct PermKey {
  uint64_t mul[4];
  uint64_t add[4];
  uint32_t bits;
  uint32_t shift;
  uint64_t mask;
};

SCAPH_HD PermKey make_perm_key(uint64_t seed, uint32_t bits) {
  PermKey k;
  const uint64_t pk = fmix64(seed ^ kPermSalt);
  k.bits = bits;
  k.mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
  k.shift = (bits + 1) / 2;
  if (k.shift == 0) k.shift = 1;
  for (int r = 0; r < 4; ++r) {
    k.mul[r] = (word(pk, 2 * r) | 1ull) & k.mask;
    k.add[r] = word(pk, 2 * r + 1) & k.mask;
    if (k.mul[r] == 0) k.mul[r] = 1;  // bits == 0 corner
  }
  return k;
}

SCAPH_HD uint64_t perm_bits(const PermKey& k, uint64_t x) {
  for (int r = 0; r < 4; ++r) {
    x = (x * k.mul[r] + k.add[r]) & k.mask;
    x ^= x >> k.shift;
  }
  return x;
}

// Bijection on [0, n): cycle-walk the bijection on the enclosing power of 2.
SCAPH_HD uint64_t perm_range(const PermKey& k, uint64_t x, uint64_t n) {
  do {
    x = perm_bits(k, x);
  } while (x >= n);
  return x;
}

// One R-MAT edge (before permutation) of a scale-`scale` graph.
SCAPH_HD void rmat_edge(uint64_t seed, uint64_t e, uint32_t scale, uint32_t thr_a,
                        uint32_t thr_ab, uint32_t thr_abc, uint32_t* u_out, uint32_t* v_out) {
  const uint32_t words = (scale + 1) / 2;
  uint32_t u = 0, v = 0;
  uint64_t base = e * words;
  for (uint32_t j = 0; j < words; ++j) {
    const uint64_t w = word(seed, base + j);
    for (uint32_t h = 0; h < 2; ++h) {
      const uint32_t l = 2 * j + h;
      if (l >= scale) break;
      const uint32_t r = h == 0 ? (uint32_t)w : (uint32_t)(w >> 32);
      const uint32_t bit = 1u << (scale - 1 - l);
      if (r < thr_a) {
      } else if (r < thr_ab) {
        v |= bit;
      } else if (r < thr_abc) {
        u |= bit;
      } else {
        u |= bit;
        v |= bit;
      }
    }
  }
  *u_out = u;
  *v_out = v;
}

}  // namespace scaph_spec

#endif
