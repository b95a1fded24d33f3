// XXH3_64bits_withSeed for an 8-byte key, the hash the reference binds for vec_hash / col_hash
// (include/types.h:8-9) and restates for CUDA (include/cuda_xxhash64.cuh:425-469,587-591).
// Restated here from the published XXH3 short-input path (length 4..8, then rrmxmx). The same
// inline code serves the HIP kernels (host+device) and the host C++ classes.
//
// Split form used by the update kernels: mix(x) = x ^ rotl(x,49) ^ rotl(x,24) is linear over
// GF(2), so mix(swap(key) ^ bitflip(seed)) = mix(swap(key)) ^ mix(bitflip(seed)); the first
// term is computed once per update and the second once per column seed.
#pragma once
#include <cstdint>

#if defined(__HIP__)
#define GS_HD __host__ __device__ __forceinline__
#else
#define GS_HD inline
#endif

namespace gs {

constexpr uint64_t kSecret8 = 0x1cad21f72c81017cULL;   // LE64 of secret bytes 8..15
constexpr uint64_t kSecret16 = 0xdb979083e96dd4deULL;  // LE64 of secret bytes 16..23
constexpr uint64_t kRrmxmxMul = 0x9FB21C651E98DF25ULL;

GS_HD uint64_t rotl64(uint64_t x, unsigned r) { return (x << r) | (x >> (64u - r)); }
GS_HD uint32_t bswap32(uint32_t x) {
  return (x >> 24) | ((x >> 8) & 0xff00u) | ((x << 8) & 0xff0000u) | (x << 24);
}
GS_HD uint64_t xxh3_bitflip(uint64_t seed) {
  uint64_t s = seed ^ ((uint64_t)bswap32((uint32_t)seed) << 32);
  return (kSecret8 ^ kSecret16) - s;
}
GS_HD uint64_t xxh3_mix(uint64_t x) { return x ^ rotl64(x, 49) ^ rotl64(x, 24); }
GS_HD uint64_t xxh3_swap(uint64_t key) { return (key >> 32) | (key << 32); }
GS_HD uint64_t xxh3_finish(uint64_t x) {
  x *= kRrmxmxMul;
  x ^= (x >> 35) + 8u;
  x *= kRrmxmxMul;
  x ^= x >> 28;
  return x;
}
GS_HD uint64_t xxh3_len8(uint64_t key, uint64_t seed) {
  return xxh3_finish(xxh3_mix(xxh3_swap(key) ^ xxh3_bitflip(seed)));
}
// per-column constant of the split form
GS_HD uint64_t xxh3_seed_key(uint64_t seed) { return xxh3_mix(xxh3_bitflip(seed)); }
// ctz(hash | 1<<max_depth): Bucket_Boruvka::get_index_depth (include/bucket.h:54-59)
GS_HD uint32_t depth_of(uint64_t h, uint32_t max_depth) {
  return (uint32_t)__builtin_ctzll(h | (1ULL << max_depth));
}
// 32x32 -> low 32 bits of the product of the low 24 bits (v_mul_u32_u24 on gfx950)
// The device form is written as the instruction itself: with only the low bits of the product in
// use, the compiler otherwise turns __umul24 back into v_mul_lo_u32, which cannot fuse an add.
GS_HD uint32_t mul24_lo(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "s"(b), "v"(a));
  return r;
#else
  return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
#endif
}
// (low 24 bits of a) * (low 24 bits of b) + c, low 32 bits (v_mad_u32_u24, one full-rate
// instruction; the same asm form as mul24_lo keeps the compiler from widening it)
GS_HD uint32_t mad24_lo(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t r;
  asm("v_mad_u32_u24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(b), "v"(c));
  return r;
#else
  return (a & 0xFFFFFFu) * (b & 0xFFFFFFu) + c;
#endif
}
// low 32 bits of (hi:lo) >> s, 0 < s < 32 (v_alignbit_b32)
GS_HD uint32_t align_bits(uint32_t hi, uint32_t lo, uint32_t s) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_alignbit(hi, lo, s);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> s);
#endif
}
// The low 16 bits of xxh3_finish(x), computed with a cheaper second multiply (bits 16..31 of
// the result are not those of the hash). Bits 0..15 of h = x3 ^ (x3 >> 28) come from bits 0..43
// of x3 = x2 * M, whose high word is needed mod 2^12 only, so its two cross products run as
// 24-bit multiply-adds. The first product x1 = x * M is needed in full (x2 = x1 ^ ((x1 >> 35) +
// 8) feeds x1's high word into x2's low word; the added term is below 2^30, so x2's high word
// equals x1's). 13 instructions on gfx950 for the XOR of x with a key and this.
GS_HD uint32_t hash_low16(uint64_t x) {
  constexpr uint32_t kMlo = (uint32_t)kRrmxmxMul, kMhi = (uint32_t)(kRrmxmxMul >> 32);
  const uint32_t xlo = (uint32_t)x, xhi = (uint32_t)(x >> 32);
  const uint64_t p = (uint64_t)xlo * kMlo;
  const uint32_t x1hi = (uint32_t)(p >> 32) + xlo * kMhi + xhi * kMlo;
  const uint32_t x2lo = (uint32_t)p ^ ((x1hi >> 3) + 8u);
  const uint64_t q = (uint64_t)x2lo * kMlo;
  const uint32_t x3hi = mad24_lo(x2lo, kMhi, mad24_lo(x1hi, kMlo, (uint32_t)(q >> 32)));
  return (uint32_t)q ^ align_bits(x3hi, (uint32_t)q, 28);
}
// depth_of(xxh3_finish(x), max_depth) capped at 16: ctz of the low 16 bits of the hash, capped
// at min(max_depth, 16). When those 16 bits are all zero (probability 2^-16) and max_depth > 16
// the full hash decides (depth_fast).
GS_HD uint32_t depth_low16(uint64_t x, uint32_t max_depth) {
  const uint32_t cap = max_depth < 16u ? max_depth : 16u;
  return (uint32_t)__builtin_ctz(hash_low16(x) | (1u << cap));
}
GS_HD uint32_t depth_fast(uint64_t x, uint32_t max_depth) {
  const uint32_t d = depth_low16(x, max_depth);
  if (d < 16u || max_depth <= 16u) return d;
  return depth_of(xxh3_finish(x), max_depth);
}

// concat_pairing_fn (src/util.cpp:45-51)
GS_HD uint64_t concat_pair(uint32_t i, uint32_t j) {
  return i < j ? (((uint64_t)i << 32) | j) : (((uint64_t)j << 32) | i);
}

}  // namespace gs
