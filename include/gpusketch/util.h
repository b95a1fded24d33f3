// Pairing helpers (the reference's include/util.h:1-44, src/util.cpp:1-57).
#pragma once
#include <utility>

#include "types.h"

unsigned long long int double_to_ull(double d, double epsilon = 0.00000001);

// i + j(j-1)/2 for i < j, order-insensitive; (i, i) is not a valid input
edge_id_t nondirectional_non_self_edge_pairing_fn(node_id_t i, node_id_t j);
Edge inv_nondir_non_self_edge_pairing_fn(edge_id_t idx);

// (min << 32) | max: the edge index every sketch is keyed on
edge_id_t concat_pairing_fn(node_id_t i, node_id_t j);
Edge inv_concat_pairing_fn(edge_id_t idx);

// [begin, end) of part i when len items are split into parts near-equal contiguous parts
std::pair<size_t, size_t> get_ith_partition(size_t len, size_t i, size_t parts);

#define likely_if(x) if (__builtin_expect((bool)(x), true))
#define unlikely_if(x) if (__builtin_expect((bool)(x), false))
