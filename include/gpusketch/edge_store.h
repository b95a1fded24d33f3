// EdgeStore: exact adjacency lists for the sparse subgraphs of the min-cut path (the reference's
// include/edge_store.h:1-74, src/edge_store.cpp:1-303). Each entry of adjlist[src] is
// {subgraph tag, dst}, kept sorted; inserting an entry that is present removes it (an update
// toggles the edge). Edges whose tag is below the first stored subgraph are handed back to the
// caller for the sketch subgraphs. When the store outgrows its budget it contracts: the first
// stored subgraph moves up by one and every vertex's entries are handed back (once) so the caller
// can build that subgraph's sketches.
#pragma once
#include <atomic>
#include <memory>
#include <mutex>
#include <vector>

#include "types.h"

class EdgeStore {
 public:
  EdgeStore(size_t seed, node_id_t num_vertices, size_t sketch_bytes, size_t num_subgraphs,
            size_t start_subgraph = 0);
  ~EdgeStore();

  // tags each destination with get_index_depth(concat(src, dst), seed, num_subgraphs - 1); for
  // use while no sketch subgraph exists
  TaggedUpdateBatch insert_adj_edges(node_id_t src, const std::vector<node_id_t>& dst_vertices);
  // pre-tagged destinations; caller_first_es_subgraph is the first stored subgraph the caller saw
  TaggedUpdateBatch insert_adj_edges(node_id_t src, node_id_t caller_first_es_subgraph,
                                     SubgraphTaggedUpdate* dst_data, size_t dst_data_size);
  // hand back the next uncontracted vertex's entries during a contraction
  TaggedUpdateBatch vertex_advance_subgraph(node_id_t cur_first_es_subgraph);

  size_t get_num_edges() { return num_edges; }
  size_t get_footprint() { return num_edges * store_edge_bytes; }
  size_t get_first_store_subgraph() { return cur_subgraph; }
  std::vector<Edge> get_edges();
  bool contract_in_progress() { return true_min_subgraph < cur_subgraph; }

 private:
  static constexpr size_t store_edge_bytes = sizeof(SubgraphTaggedUpdate);

  std::vector<SubgraphTaggedUpdate> vertex_contract(node_id_t src);  // adj_mutex[src] held
  void check_if_too_big();

  size_t seed;
  node_id_t num_vertices;
  size_t num_subgraphs;
  std::atomic<size_t> cur_subgraph;       // subgraph at which edges enter the store
  std::atomic<size_t> true_min_subgraph;  // smallest subgraph still present in the store
  std::atomic<edge_id_t> num_edges{0};
  std::atomic<node_id_t> needs_contraction{0};  // next vertex to claim in a contraction
  std::atomic<node_id_t> vertices_handled{0};   // vertices of the contraction handled so far
  std::vector<std::vector<SubgraphTaggedUpdate>> adjlist;
  std::vector<char> vertex_contracted;
  size_t max_edges;
  std::unique_ptr<std::mutex[]> adj_mutex;
  std::mutex contract_lock;
};
