// GraphVerifier: the exact graph a stream produced, used to check query answers (the interface of
// the reference's include/test/graph_verifier.h, test/util/graph_verifier.cpp). The reference's
// driver builds one while it reads a stream when VERIFY_SAMPLES_F is defined and hands a copy to
// the algorithm (include/graph_sketch_driver.h:64,150-189), whose queries then check their
// answers against it (src/cc_sketch_alg.cpp:179,550,580,592,620). Here the graph is a set of
// edges (toggle semantics) rather than an n x n bit matrix, so it fits any vertex count the
// sketches do.
#pragma once
#include <cstdint>
#include <exception>
#include <string>
#include <unordered_set>
#include <vector>

#include "return_types.h"
#include "types.h"

class BadEdgeException : public std::exception {
  std::string err_msg;

 public:
  explicit BadEdgeException(std::string err) : err_msg(std::move(err)) {}
  const char* what() const noexcept override { return err_msg.c_str(); }
};

class IncorrectCCException : public std::exception {
  std::string err_msg;

 public:
  explicit IncorrectCCException(std::string err) : err_msg(std::move(err)) {}
  const char* what() const noexcept override { return err_msg.c_str(); }
};

class IncorrectForestException : public std::exception {
  std::string err_msg;

 public:
  explicit IncorrectForestException(std::string err) : err_msg(std::move(err)) {}
  const char* what() const noexcept override { return err_msg.c_str(); }
};

class GraphVerifier {
 public:
  explicit GraphVerifier(node_id_t num_vertices);
  // the graph of a cumulative text file: a "num_vertices num_edges" line, then one "src dst" line
  // per edge (an edge listed twice cancels), the format of the reference's test/res/*.txt
  GraphVerifier(node_id_t num_vertices, const std::string& cumul_file_name);
  GraphVerifier(const GraphVerifier& oth) = default;

  // toggle an edge (throws BadEdgeException for a vertex out of range)
  void edge_update(Edge edge);
  // throws BadEdgeException unless the edge is in the graph
  void verify_edge(Edge edge);
  // throws IncorrectCCException unless cc has the graph's components exactly
  void verify_connected_components(const ConnectedComponents& cc);
  // forest i must be a spanning forest of the graph minus forests 0..i-1: its edges in that
  // graph, none redundant within the forest, and its components that graph's; throws
  // IncorrectForestException (or BadEdgeException) otherwise; the graph is left unchanged
  void verify_spanning_forests(std::vector<SpanningForest> SFs);
  // the graph of both streams (edge sets XORed)
  void combine(const GraphVerifier& oth);

  // row i holds edges (i, i + j) at index j, as the reference's matrix (O(n^2) bits: small n)
  std::vector<std::vector<bool>> extract_adj_matrix();
  node_id_t get_num_kruskal_ccs();
  size_t num_edges() const { return edges.size(); }

  bool operator==(const GraphVerifier& oth) const {
    return num_vertices == oth.num_vertices && edges == oth.edges;
  }
  bool operator!=(const GraphVerifier& oth) const { return !(*this == oth); }

 private:
  static uint64_t key(node_id_t a, node_id_t b) {
    return a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
  }
  // the graph's components: labels[v] = smallest vertex of v's component
  void kruskal();

  node_id_t num_vertices;
  std::unordered_set<uint64_t> edges;
  std::vector<node_id_t> labels;
  node_id_t kruskal_ccs = 0;
  bool need_query_compute = true;
};
