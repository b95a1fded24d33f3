// MCGPUSketchAlg: k-connectivity and minimum cut of a graph stream (the reference's
// include/mc_gpu_sketch_alg.h:13-273, src/mc_gpu_sketch_alg.cpp:1-209, and the MCSketchAlg
// queries it inherits, src/mc_sketch_alg.cpp:738-1003). Every edge has a subgraph depth
// t = get_index_depth(concat(u, v), seed, num_subgraphs - 1) and belongs to subgraphs 0..t.
// Subgraphs below the EdgeStore's first stored subgraph are sketch subgraphs, each a device
// sketch store; the rest are kept exactly in the EdgeStore. For subgraph g, k edge-disjoint
// spanning forests (on the device) form a certificate whose minimum cut equals the subgraph's
// when that is below k; the first subgraph whose certificate cut is below k gives the answer,
// scaled by 2^g.
#pragma once
#include <atomic>
#include <memory>
#include <mutex>
#include <set>
#include <vector>

#include "cc_alg_configuration.h"
#include "cc_gpu_sketch_alg.h"
#include "device_sketch_store.h"
#include "edge_store.h"
#include "eager_dsu.h"
#include "gpu_sketch_params.h"
#include "return_types.h"

struct MinCut {
  std::set<node_id_t> left_vertices;
  std::set<node_id_t> right_vertices;
  size_t value;
};

class MCGPUSketchAlg {
 public:
  MCGPUSketchAlg(node_id_t num_vertices, int num_threads, int num_reader_threads,
                 SketchParams sketch_params, int num_subgraphs, int max_sketch_graphs, int k,
                 size_t sketch_bytes, int initial_sketch_graphs,
                 CCAlgConfiguration config = CCAlgConfiguration());
  ~MCGPUSketchAlg();

  // driver contract
  size_t get_desired_updates_per_batch();
  void pre_insert(GraphUpdate upd, int thr_id = 0) { eager.pre_insert(upd); }
  void allocate_worker_memory(size_t num_workers);
  void apply_update_batch(int thr_id, node_id_t src_vertex,
                          const std::vector<node_id_t>& dst_vertices);
  bool has_cached_query(int query_code) { return query_code == CONNECTIVITY && eager.valid(); }
  void print_configuration() { std::cout << config << std::endl; }
  node_id_t get_num_vertices() { return num_vertices; }

  // finish pending EdgeStore contractions and apply every buffered batch
  void apply_flush_updates();

  ConnectedComponents connected_components();
  SpanningForest calc_spanning_forest(size_t graph_id = 0);
  std::vector<SpanningForest> calc_disjoint_spanning_forests(size_t graph_id, size_t k);
  // exact minimum cut of the multigraph given by an edge list
  MinCut calc_minimum_cut(const std::vector<Edge>& edges);

  std::vector<Edge> get_adjlist_spanning_forests() { return edge_store.get_edges(); }
  int get_num_sketch_graphs() { return (int)cur_subgraphs.load(); }
  size_t get_num_adjlist_edges() { return edge_store.get_num_edges(); }
  void print_subgraph_edges();
  void display_time();
  DeviceSketchStore& subgraph_store(int graph_id) { return *subgraphs.at((size_t)graph_id); }

  inline size_t get_seed() { return seed; }
  inline bool get_update_locked() { return update_locked; }

 private:
  void complete_update_batch(int thr_id, const TaggedUpdateBatch& updates);
  void ensure_subgraphs(size_t count);

  node_id_t num_vertices;
  size_t seed;
  size_t num_samples;
  size_t num_subgraphs;
  int max_sketch_graphs;
  int k;
  int num_host_threads;
  CCAlgConfiguration config;
  EdgeStore edge_store;
  std::vector<std::unique_ptr<DeviceSketchStore>> subgraphs;  // max_sketch_graphs slots
  std::vector<std::atomic<uint64_t>> subgraph_updates;
  std::atomic<size_t> cur_subgraphs{0};
  std::mutex sketch_creation_lock;
  std::atomic<bool> update_locked{false};
  EagerDSU eager;
};

// The certificate loop of the reference's min-cut tool (tools/min_cut.cu:205-252): returns the
// cut of the first subgraph whose k-forest certificate has a cut below k (or of the first
// EdgeStore subgraph), with value scaled by 2^graph_id.
struct MinCutAnswer {
  size_t value;    // scaled estimate of the graph's minimum cut
  int graph_id;    // subgraph that produced it
  MinCut cut;      // the subgraph's cut
};
MinCutAnswer mc_minimum_cut(MCGPUSketchAlg& alg, int num_graphs, int k);
