// Bucket and its update/test functions (the reference's include/bucket.h:1-73). The bucket is
// 16 bytes, {alpha = XOR of indices, gamma = XOR of their 64-bit checksums}, the same byte image
// the device store exports.
#pragma once
#include <cstdint>

#include "types.h"
#include "xxh3_len8.h"

#pragma pack(push, 1)
struct Bucket {
  vec_t alpha;
  vec_hash_t gamma;
};
#pragma pack(pop)
static_assert(sizeof(Bucket) == 16, "Bucket must be 16 bytes");

namespace Bucket_Boruvka {
static constexpr size_t col_hash_bits = sizeof(col_hash_t) * 8;

// depth of an update in a column: trailing zeros of XXH3(idx, column seed), capped at max_depth
inline col_hash_t get_index_depth(const vec_t update_idx, const long seed_and_col,
                                  const vec_hash_t max_depth) {
  return gs::depth_of(gs::xxh3_len8(update_idx, (uint64_t)seed_and_col), (uint32_t)max_depth);
}

// checksum contribution of an index
inline vec_hash_t get_index_hash(const vec_t update_idx, const long sketch_seed) {
  return gs::xxh3_len8(update_idx, (uint64_t)sketch_seed);
}

// a bucket holding exactly one index has gamma == hash(alpha)
inline bool is_good(const Bucket& bucket, const long sketch_seed) {
  return bucket.gamma == get_index_hash(bucket.alpha, sketch_seed);
}

inline void update(Bucket& bucket, const vec_t update_idx, const vec_hash_t update_hash) {
  bucket.alpha ^= update_idx;
  bucket.gamma ^= update_hash;
}
}  // namespace Bucket_Boruvka
