// Query answers (the reference's include/return_types.h:1-40, src/return_types.cpp:1-62).
#pragma once
#include <set>
#include <unordered_set>
#include <vector>

#include "dsu.h"
#include "types.h"

class ConnectedComponents {
  std::vector<node_id_t> parent_arr;  // a representative per vertex
  node_id_t num_vertices;
  node_id_t num_cc;

 public:
  ConnectedComponents(node_id_t num_vertices, DisjointSetUnion_MT<node_id_t>& dsu);
  // from a component label per vertex (labels need not be vertex ids of the component)
  ConnectedComponents(node_id_t num_vertices, const node_id_t* labels);

  std::vector<std::set<node_id_t>> get_component_sets();
  bool is_connected(node_id_t a, node_id_t b) const { return parent_arr[a] == parent_arr[b]; }
  node_id_t size() const { return num_cc; }
  const std::vector<node_id_t>& labels() const { return parent_arr; }
};

class SpanningForest {
  node_id_t num_vertices;
  std::vector<Edge> edges;
  std::vector<Edge> sorted_adjacency;
  bool has_adjacency = false;

 public:
  SpanningForest(node_id_t num_vertices, const std::unordered_set<node_id_t>* spanning_forest);
  SpanningForest(node_id_t num_vertices, std::vector<Edge> edges);

  const std::vector<Edge>& get_edges() const { return edges; }
  // both directions of every edge, sorted
  const std::vector<Edge>& get_sorted_adjacency();
};
