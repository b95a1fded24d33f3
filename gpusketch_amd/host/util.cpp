// Pairing helpers (src/util.cpp:1-57 of the reference).
#include "gpusketch/util.h"
#include "gpusketch/xxh3_len8.h"

#include <algorithm>
#include <cmath>

unsigned long long int double_to_ull(double d, double epsilon) {
  return (unsigned long long)(d + epsilon);
}

edge_id_t nondirectional_non_self_edge_pairing_fn(node_id_t i, node_id_t j) {
  uint64_t lo = std::min(i, j), hi = std::max(i, j);
  // hi * (hi - 1) / 2 without overflowing the intermediate product
  uint64_t tri = (hi % 2 == 0) ? (hi / 2) * (hi - 1) : hi * ((hi - 1) / 2);
  return lo + tri;
}

Edge inv_nondir_non_self_edge_pairing_fn(edge_id_t idx) {
  // largest j with j(j-1)/2 <= idx, corrected after the floating-point estimate
  uint64_t j = (uint64_t)((std::sqrt(8.0 * (double)idx + 1.0) + 1.0) / 2.0);
  auto tri = [](uint64_t x) { return x % 2 == 0 ? (x / 2) * (x - 1) : x * ((x - 1) / 2); };
  while (j > 1 && tri(j) > idx) j--;
  while (tri(j + 1) <= idx) j++;
  return {(node_id_t)(idx - tri(j)), (node_id_t)j};
}

edge_id_t concat_pairing_fn(node_id_t i, node_id_t j) { return gs::concat_pair(i, j); }

Edge inv_concat_pairing_fn(edge_id_t idx) {
  return {(node_id_t)(idx >> 32), (node_id_t)(idx & 0xFFFFFFFFu)};
}

std::pair<size_t, size_t> get_ith_partition(size_t len, size_t i, size_t parts) {
  const size_t base = len / parts, extra = len % parts;
  const size_t begin = i * base + std::min(i, extra);
  return {begin, begin + base + (i < extra ? 1 : 0)};
}
