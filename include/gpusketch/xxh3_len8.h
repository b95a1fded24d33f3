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
// concat_pairing_fn (src/util.cpp:45-51)
GS_HD uint64_t concat_pair(uint32_t i, uint32_t j) {
  return i < j ? (((uint64_t)i << 32) | j) : (((uint64_t)j << 32) | i);
}

}  // namespace gs
