// Query answers (src/return_types.cpp:1-62 of the reference).
#include "gpusketch/return_types.h"

#include <algorithm>
#include <map>
#include <unordered_map>

ConnectedComponents::ConnectedComponents(node_id_t n, DisjointSetUnion_MT<node_id_t>& dsu)
    : parent_arr(n), num_vertices(n), num_cc(0) {
  for (node_id_t i = 0; i < n; i++) {
    parent_arr[i] = dsu.find_root(i);
    if (parent_arr[i] == i) num_cc++;
  }
}

ConnectedComponents::ConnectedComponents(node_id_t n, const node_id_t* labels)
    : parent_arr(labels, labels + n), num_vertices(n), num_cc(0) {
  std::vector<node_id_t> seen(parent_arr);
  std::sort(seen.begin(), seen.end());
  num_cc = (node_id_t)(std::unique(seen.begin(), seen.end()) - seen.begin());
}

std::vector<std::set<node_id_t>> ConnectedComponents::get_component_sets() {
  std::map<node_id_t, std::set<node_id_t>> by_label;
  for (node_id_t i = 0; i < num_vertices; ++i) by_label[parent_arr[i]].insert(i);
  std::vector<std::set<node_id_t>> out;
  out.reserve(by_label.size());
  for (auto& kv : by_label) out.push_back(std::move(kv.second));
  return out;
}

SpanningForest::SpanningForest(node_id_t n, const std::unordered_set<node_id_t>* sf)
    : num_vertices(n) {
  for (node_id_t src = 0; src < n; src++)
    for (node_id_t dst : sf[src]) edges.push_back({src, dst});
}

SpanningForest::SpanningForest(node_id_t n, std::vector<Edge> e)
    : num_vertices(n), edges(std::move(e)) {}

const std::vector<Edge>& SpanningForest::get_sorted_adjacency() {
  if (has_adjacency) return sorted_adjacency;
  sorted_adjacency.clear();
  sorted_adjacency.reserve(2 * edges.size());
  for (const Edge& e : edges) {
    sorted_adjacency.push_back(e);
    sorted_adjacency.push_back({e.dst, e.src});
  }
  std::sort(sorted_adjacency.begin(), sorted_adjacency.end());
  has_adjacency = true;
  return sorted_adjacency;
}
