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
#include <vector>

#include "edge_store.h"
#include "mc_sketch_alg.h"

class MCGPUSketchAlg : public MCSketchAlg {
 public:
  MCGPUSketchAlg(node_id_t num_vertices, int num_threads, int num_reader_threads,
                 SketchParams sketch_params, int num_subgraphs, int max_sketch_graphs, int k,
                 size_t sketch_bytes, int initial_sketch_graphs,
                 CCAlgConfiguration config = CCAlgConfiguration());
  ~MCGPUSketchAlg() override;

  // driver contract (the rest is MCSketchAlg's): each edge goes to the sketch subgraphs its
  // depth selects and, past them, to the EdgeStore (src/mc_gpu_sketch_alg.cpp:139-178)
  void apply_update_batch(int thr_id, node_id_t src_vertex,
                          const std::vector<node_id_t>& dst_vertices);

  // finish pending EdgeStore contractions and apply every buffered batch
  void apply_flush_updates();

  // connectivity of the whole graph: sketch graph 0, or the EdgeStore when no sketch graph holds
  // anything
  ConnectedComponents connected_components();

  std::vector<Edge> get_adjlist_spanning_forests() { return edge_store.get_edges(); }
  int get_num_sketch_graphs() { return (int)cur_subgraphs.load(); }
  size_t get_num_adjlist_edges() { return edge_store.get_num_edges(); }
  void print_subgraph_edges();
  void display_time();

 private:
  void complete_update_batch(int thr_id, const TaggedUpdateBatch& updates);

  size_t num_subgraphs;
  int k;
  EdgeStore edge_store;
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
