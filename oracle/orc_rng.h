// ORACLE -- test infrastructure only.
//
// The counter-based sample stream shared (by specification, not by code)
// with the device kernels.  The reference draws its samples from a
// per-thread std::mt19937 (util/random-std.h:22-40) and shuffles jittered
// grids per channel (render/sample-set.cc:36-56, render/grid.cc:23-93); that
// stream is implementation-defined, so parity with the reference is
// statistical.  This stream keeps the same structure -- a jittered grid
// per (pixel, channel), randomly permuted over the pixel's samples, and
// plain uniform numbers past min_path_len -- keyed by counters instead of a
// sequential generator so every path can be evaluated independently.
//
// mix64 is the splitmix64 output finaliser (Steele, Lea, Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014).
#pragma once
#include <cstdint>

namespace orc {

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static const uint64_t GOLD = 0x9e3779b97f4a7c15ULL;

// Key of one pixel of one job.
static inline uint64_t pixel_key(uint64_t seed, int px, int py) {
  uint64_t p = ((uint64_t)(uint32_t)px << 32) | (uint64_t)(uint32_t)py;
  return mix64(seed ^ mix64(p + GOLD));
}
// Key of one path (pixel, sample).
static inline uint64_t path_key(uint64_t pk, uint32_t sample) {
  return mix64(pk ^ (((uint64_t)sample + 1) * 0xd1b54a32d192ed03ULL));
}
// Key of one stratified channel of a pixel.
static inline uint64_t chan_key(uint64_t pk, uint32_t chan) {
  return mix64(pk + ((uint64_t)chan + 1) * 0xa0761d6478bd642fULL);
}
// 32 random bits number D of a key.
static inline uint32_t draw32(uint64_t key, uint32_t d) {
  return (uint32_t)(mix64(key + ((uint64_t)d + 1) * GOLD) >> 32);
}
// Uniform float in [0, 1) with 24 random bits.
static inline float u01(uint32_t r) {
  return (float)(r >> 8) * (1.0f / 16777216.0f);
}
static inline float draw(uint64_t key, uint32_t d) { return u01(draw32(key, d)); }

// Random permutation of [0, n) by cycle walking a bijection of
// [0, 2^k) (Black & Rogaway, "Ciphers with arbitrary finite domains",
// CT-RSA 2002).  The bijection is a chain of xor-with-key, odd multiply
// and xorshift steps, each invertible modulo 2^k.
static inline uint32_t permute(uint32_t i, uint32_t n, uint64_t key) {
  if (n <= 1) return 0;
  uint32_t w = n - 1;
  w |= w >> 1; w |= w >> 2; w |= w >> 4; w |= w >> 8; w |= w >> 16;
  uint32_t mask = w;
  unsigned bits = 0;
  while (w) { bits++; w >>= 1; }
  unsigned hs = (bits + 1) / 2;
  uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
  do {
    i ^= k0 & mask;
    i = (i * 0x9e3779b1u) & mask;
    i ^= i >> hs;
    i = (i * (k1 | 1u)) & mask;
    i ^= i >> hs;
    i = (i + (k1 >> 16)) & mask;
  } while (i >= n);
  return i;
}

// Grid shape of a UV channel with N total samples (render/grid.cc:26-38,
// :75-82): ceil(sqrt N) x round(sqrt N) cells; N is first rounded up to
// that product (render/sample-set.h:330-345).
struct GridShape {
  uint32_t u_steps, v_steps, total;
};
static inline GridShape grid_shape(uint32_t n) {
  uint32_t up = 1, down = 1;
  // ceil(sqrt(n)) and floor(sqrt(n) + 0.5) in integers.
  while (up * up < n) up++;
  down = up;
  while (down > 1 && (2 * down - 1) * (2 * down - 1) > 4 * n) down--;
  GridShape g{up, down, up * down};
  return g;
}

}  // namespace orc
