// MCSketchAlg: k-connectivity and minimum cut from sketch subgraphs, the base of MCGPUSketchAlg
// (the reference's include/mc_sketch_alg.h:77-304, src/mc_sketch_alg.cpp). Sketch graph g holds
// one l0 sketch per vertex of subgraph g; here each is a device sketch store, and the queries
// (Boruvka, k disjoint spanning forests) run on the device. MCGPUSketchAlg adds the routing of
// each edge to the subgraphs its depth selects and the exact EdgeStore for the sparse ones. Code
// that holds an MCSketchAlg* works on either. With a GraphVerifier set, answers about sketch
// graph 0 (the whole graph) are checked against the exact graph.
#pragma once
#include <atomic>
#include <chrono>
#include <iostream>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "cc_alg_configuration.h"
#include "cc_sketch_alg.h"
#include "device_sketch_store.h"
#include "eager_dsu.h"
#include "gpu_sketch_params.h"
#include "graph_verifier.h"
#include "return_types.h"

struct MinCut {
  std::set<node_id_t> left_vertices;
  std::set<node_id_t> right_vertices;
  size_t value;
};

class MCSketchAlg {
 public:
  // include/mc_sketch_alg.h: no sketch graph exists until create_sketch_graph(0, ...)
  MCSketchAlg(node_id_t num_vertices, size_t seed, int max_sketch_graphs,
              CCAlgConfiguration config = CCAlgConfiguration());
  virtual ~MCSketchAlg();

  // sketch graph graph_id (the next one, below max_sketch_graphs); the parameters' geometry
  // fields, when set, must match the vertex count's (src/mc_sketch_alg.cpp:93-135)
  void create_sketch_graph(int graph_id, const SketchParams sketchParams);
  // a write_binary file as sketch graph 0 (src/mc_sketch_alg.cpp:43-56)
  static MCSketchAlg* construct_from_serialized_data(const std::string& input_file,
                                                     CCAlgConfiguration config = CCAlgConfiguration());

  // driver contract
  size_t get_desired_updates_per_batch();
  void pre_insert(GraphUpdate upd, int thr_id = 0) { eager.pre_insert(upd); }
  void allocate_worker_memory(size_t num_workers);
  // into sketch graph 0 (src/mc_sketch_alg.cpp:151-163)
  void apply_update_batch(int thr_id, node_id_t src_vertex,
                          const std::vector<node_id_t>& dst_vertices);
  void apply_update_batch_single_graph(int thr_id, int graph_id, node_id_t src_vertex,
                                       const std::vector<node_id_t>& dst_vertices);
  bool has_cached_query(int query_code) { return query_code == CONNECTIVITY && eager.valid(); }
  void print_configuration() { std::cout << config << std::endl; }
  void set_verifier(std::unique_ptr<GraphVerifier> verifier) { this->verifier = std::move(verifier); }

  void apply_raw_buckets_update(node_id_t src_vertex, Bucket* raw_buckets);
  // one undirected update of sketch graph graph_id, applied at once (src/mc_sketch_alg.cpp:186-194)
  void update_subgraph(int graph_id, GraphUpdate upd);

  ConnectedComponents connected_components();
  bool point_query(node_id_t a, node_id_t b);
  SpanningForest calc_spanning_forest(size_t graph_id = 0);
  std::vector<SpanningForest> calc_disjoint_spanning_forests(size_t graph_id, size_t k);
  // exact minimum cut of the multigraph given by an edge list (the reference calls VieCut)
  MinCut calc_minimum_cut(const std::vector<Edge>& edges);
  // sketch graph 0 in the CCSketchAlg file format (src/mc_sketch_alg.cpp:1004-1013)
  void write_binary(const std::string& filename);

  DeviceSketchStore& subgraph_store(int graph_id) { return *subgraphs.at((size_t)graph_id); }

  std::chrono::steady_clock::time_point cc_alg_start;
  std::chrono::steady_clock::time_point cc_alg_end;
  std::chrono::steady_clock::time_point sf_query_start;
  std::chrono::steady_clock::time_point sf_query_end;
  std::chrono::duration<double> query_time{0};
  size_t last_query_rounds = 0;

  inline node_id_t get_num_vertices() { return num_vertices; }
  inline size_t get_seed() { return seed; }
  inline size_t max_rounds() { return num_samples; }
  inline bool get_update_locked() { return update_locked; }

 protected:
  MCSketchAlg(node_id_t num_vertices, size_t seed, size_t num_samples, int max_sketch_graphs,
              int num_threads, CCAlgConfiguration config);
  // sketch graphs 0..count-1 exist afterwards (thread-safe)
  void ensure_subgraphs(size_t count);

  node_id_t num_vertices;
  size_t seed;
  size_t num_samples;
  int max_sketch_graphs;
  int num_host_threads;
  CCAlgConfiguration config;
  std::vector<std::unique_ptr<DeviceSketchStore>> subgraphs;  // max_sketch_graphs slots
  std::vector<std::atomic<uint64_t>> subgraph_updates;
  std::atomic<size_t> cur_subgraphs{0};
  std::mutex sketch_creation_lock;
  std::atomic<bool> update_locked{false};
  EagerDSU eager;
  std::unique_ptr<GraphVerifier> verifier;
};
