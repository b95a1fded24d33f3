// GraphVerifier (see graph_verifier.h): the checks of the reference's
// test/util/graph_verifier.cpp over an edge set.
#include "gpusketch/graph_verifier.h"

#include <fstream>
#include <sstream>

namespace {
uint32_t find(std::vector<uint32_t>& p, uint32_t x) {
  while (p[x] != x) x = p[x] = p[p[x]];
  return x;
}
// the smaller root survives, so every label is the smallest vertex of its component
bool unite(std::vector<uint32_t>& p, uint32_t a, uint32_t b) {
  a = find(p, a);
  b = find(p, b);
  if (a == b) return false;
  if (a < b) p[b] = a;
  else p[a] = b;
  return true;
}
}  // namespace

GraphVerifier::GraphVerifier(node_id_t n) : num_vertices(n) {}

GraphVerifier::GraphVerifier(node_id_t n, const std::string& cumul_file_name) : num_vertices(n) {
  std::ifstream in(cumul_file_name);
  if (!in) throw std::runtime_error("cannot open " + cumul_file_name);
  uint64_t file_n = 0, m = 0;
  in >> file_n >> m;
  if (!in || file_n > n) throw std::runtime_error("bad cumulative file header: " + cumul_file_name);
  for (uint64_t i = 0; i < m; i++) {
    uint64_t a, b;
    if (!(in >> a >> b)) throw std::runtime_error("truncated cumulative file: " + cumul_file_name);
    edge_update({(node_id_t)a, (node_id_t)b});
  }
  kruskal();
}

void GraphVerifier::edge_update(Edge e) {
  if (e.src >= num_vertices || e.dst >= num_vertices)
    throw BadEdgeException("Source " + std::to_string(e.src) + " or Destination " +
                           std::to_string(e.dst) + " out of bounds!");
  const uint64_t k = key(e.src, e.dst);
  if (!edges.erase(k)) edges.insert(k);
  need_query_compute = true;
}

void GraphVerifier::kruskal() {
  if (!need_query_compute) return;
  std::vector<uint32_t> p(num_vertices);
  for (node_id_t v = 0; v < num_vertices; v++) p[v] = v;
  kruskal_ccs = num_vertices;
  for (uint64_t k : edges)
    if (unite(p, (uint32_t)(k >> 32), (uint32_t)k)) kruskal_ccs--;
  labels.resize(num_vertices);
  for (node_id_t v = 0; v < num_vertices; v++) labels[v] = find(p, v);
  need_query_compute = false;
}

void GraphVerifier::verify_edge(Edge e) {
  if (e.src >= num_vertices || e.dst >= num_vertices || !edges.count(key(e.src, e.dst)))
    throw BadEdgeException("The edge (" + std::to_string(e.src) + ", " + std::to_string(e.dst) +
                           ") is not in the graph!");
}

void GraphVerifier::verify_connected_components(const ConnectedComponents& cc) {
  kruskal();
  if (cc.size() != kruskal_ccs) throw IncorrectCCException("Incorrect number of components!");
  // same partition: every vertex is connected to its component's smallest vertex in cc, and the
  // counts agree, so no two of the graph's components are joined in cc
  for (node_id_t v = 0; v < num_vertices; v++)
    if (!cc.is_connected(labels[v], v)) throw IncorrectCCException("Incorrect Connectivity!");
}

void GraphVerifier::verify_spanning_forests(std::vector<SpanningForest> SFs) {
  const auto backup = edges;
  try {
    for (SpanningForest& forest : SFs) {
      kruskal();  // the components of the graph minus the earlier forests
      const std::vector<node_id_t> want = labels;
      std::vector<uint32_t> p(num_vertices);
      for (node_id_t v = 0; v < num_vertices; v++) p[v] = v;
      for (const Edge& e : forest.get_edges()) {
        verify_edge(e);
        if (!unite(p, e.src, e.dst))
          throw IncorrectForestException("Found an edge: (" + std::to_string(e.src) + ", " +
                                         std::to_string(e.dst) +
                                         ") that is redundant within a single spanning forest!");
        edge_update(e);  // gone for the later forests: they must be edge disjoint
      }
      for (node_id_t v = 0; v < num_vertices; v++)
        if (find(p, v) != want[v])
          throw IncorrectForestException("Forest does not match expected component sets!");
    }
  } catch (...) {
    edges = backup;
    need_query_compute = true;
    throw;
  }
  edges = backup;
  need_query_compute = true;
}

void GraphVerifier::combine(const GraphVerifier& oth) {
  for (uint64_t k : oth.edges)
    if (!edges.erase(k)) edges.insert(k);
  need_query_compute = true;
}

std::vector<std::vector<bool>> GraphVerifier::extract_adj_matrix() {
  std::vector<std::vector<bool>> m(num_vertices);
  for (node_id_t i = 0; i < num_vertices; i++) m[i].assign(num_vertices - i, false);
  for (uint64_t k : edges) {
    const node_id_t a = (node_id_t)(k >> 32), b = (node_id_t)k;
    m[a][b - a] = true;
  }
  return m;
}

node_id_t GraphVerifier::get_num_kruskal_ccs() {
  kruskal();
  return kruskal_ccs;
}
