// Core value types of the C++ API. Mirrors the reference's include/types.h:7-41 and the
// StreamingUtilities types it includes (node_id_t, Edge, UpdateType, GraphStreamUpdate), which are
// not vendored in the reference; the widths follow its use (node_id_t must be 32-bit because the
// concat pairing shifts by 32, src/util.cpp:12,50; vec_hash_t is 64-bit, the device kernels do
// 8-byte atomics on it, src/cuda_kernel.cu:49,192,351).
#pragma once
#include <cstddef>
#include <cstdint>
#include <tuple>
#include <vector>

typedef uint32_t node_id_t;
typedef uint64_t vec_t;
typedef uint64_t vec_hash_t;
typedef uint64_t col_hash_t;
typedef uint64_t edge_id_t;

struct Edge {
  node_id_t src = 0;
  node_id_t dst = 0;
  bool operator<(const Edge& o) const { return std::tie(src, dst) < std::tie(o.src, o.dst); }
  bool operator==(const Edge& o) const { return src == o.src && dst == o.dst; }
  bool operator!=(const Edge& o) const { return !(*this == o); }
};

enum UpdateType { INSERT = 0, DELETE = 1, BREAKPOINT = 2 };

// an update as the driver hands it to the algorithm (include/types.h:12-15)
struct GraphUpdate {
  Edge edge;
  UpdateType type;
};

// an update as a GraphStream returns it
struct GraphStreamUpdate {
  Edge edge;
  UpdateType type;
};

// min-cut path: destination tagged with the highest subgraph its edge maps to
// (include/types.h:17-34)
struct SubgraphTaggedUpdate {
  node_id_t subgraph;
  node_id_t dst;
  bool operator<(const SubgraphTaggedUpdate& o) const {
    return subgraph != o.subgraph ? subgraph < o.subgraph : dst < o.dst;
  }
  bool operator>(const SubgraphTaggedUpdate& o) const { return o < *this; }
  bool operator==(const SubgraphTaggedUpdate& o) const {
    return subgraph == o.subgraph && dst == o.dst;
  }
};

struct TaggedUpdateBatch {
  node_id_t src;
  node_id_t min_subgraph;
  node_id_t first_es_subgraph;
  std::vector<SubgraphTaggedUpdate> dsts_data;
};
