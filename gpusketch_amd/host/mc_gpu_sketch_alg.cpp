// MCGPUSketchAlg (src/mc_gpu_sketch_alg.cpp:1-209 and tools/min_cut.cu:205-252 of the
// reference): subgraph routing and the EdgeStore over MCSketchAlg.
#include "gpusketch/mc_gpu_sketch_alg.h"

#include <algorithm>
#include <cmath>
#include <iostream>
#include <stdexcept>
#include <thread>

MCGPUSketchAlg::MCGPUSketchAlg(node_id_t n, int num_threads, int /*num_reader_threads*/,
                               SketchParams p, int num_subgraphs, int max_sketch_graphs, int k,
                               size_t sketch_bytes, int initial_sketch_graphs,
                               CCAlgConfiguration cfg)
    : MCSketchAlg(n, p.seed,
                  p.num_samples ? p.num_samples
                                : Sketch::calc_cc_samples(n, cfg.get_sketches_factor()),
                  max_sketch_graphs, num_threads, cfg),
      num_subgraphs((size_t)num_subgraphs),
      k(k),
      edge_store(p.seed, n, sketch_bytes, (size_t)num_subgraphs, (size_t)initial_sketch_graphs) {
  if (max_sketch_graphs < initial_sketch_graphs)
    throw std::invalid_argument("Cannot have initial sketch graphs > max sketch graphs");
  if (num_subgraphs < 1 || num_subgraphs > 64)
    throw std::invalid_argument("num_subgraphs must be in [1, 64]");
  ensure_subgraphs((size_t)initial_sketch_graphs);
}

MCGPUSketchAlg::~MCGPUSketchAlg() = default;

void MCGPUSketchAlg::complete_update_batch(int thr, const TaggedUpdateBatch& upd) {
  const size_t first_es = upd.first_es_subgraph;
  if (first_es == 0) return;
  ensure_subgraphs(first_es);
  std::vector<std::vector<node_id_t>> per(first_es);
  for (const auto& d : upd.dsts_data) {
    const size_t top = std::min<size_t>(d.subgraph, first_es - 1);
    for (size_t g = upd.min_subgraph; g <= top; g++) per[g].push_back(d.dst);
  }
  for (size_t g = upd.min_subgraph; g < first_es; g++) {
    if (per[g].empty()) continue;
    subgraphs[g]->add_batch(thr, upd.src, per[g].data(), per[g].size());
    subgraph_updates[g] += per[g].size();
  }
}

void MCGPUSketchAlg::apply_update_batch(int thr, node_id_t src,
                                        const std::vector<node_id_t>& dsts) {
  if (update_locked) throw UpdateLockedException();
  const node_id_t first_es = (node_id_t)edge_store.get_first_store_subgraph();
  if (first_es == 0) {
    TaggedUpdateBatch more = edge_store.insert_adj_edges(src, dsts);
    if (!more.dsts_data.empty()) complete_update_batch(thr, more);
    return;
  }
  std::vector<SubgraphTaggedUpdate> store_edges, sketch_edges(dsts.size());
  for (size_t i = 0; i < dsts.size(); i++) {
    const vec_t idx = concat_pairing_fn(src, dsts[i]);
    const node_id_t t =
        (node_id_t)Bucket_Boruvka::get_index_depth(idx, (long)seed, num_subgraphs - 1);
    if (t >= first_es) store_edges.push_back({t, dsts[i]});
    sketch_edges[i] = {t, dsts[i]};
  }
  TaggedUpdateBatch more =
      edge_store.insert_adj_edges(src, first_es, store_edges.data(), store_edges.size());
  if (!sketch_edges.empty()) complete_update_batch(thr, {src, 0, first_es, std::move(sketch_edges)});
  if (!more.dsts_data.empty()) complete_update_batch(thr, more);
}

void MCGPUSketchAlg::apply_flush_updates() {
  std::vector<std::thread> threads;
  for (int t = 0; t < num_host_threads; t++)
    threads.emplace_back([this, t] {
      while (edge_store.contract_in_progress()) {
        TaggedUpdateBatch more = edge_store.vertex_advance_subgraph(
            (node_id_t)edge_store.get_first_store_subgraph());
        if (!more.dsts_data.empty()) complete_update_batch(t, more);
        else std::this_thread::yield();  // the last claimed vertices are being handled
      }
    });
  for (auto& th : threads) th.join();
  for (size_t g = 0; g < cur_subgraphs; g++) subgraphs[g]->flush();
}

ConnectedComponents MCGPUSketchAlg::connected_components() {
  if (cur_subgraphs == 0 && !eager.valid()) {  // every edge is in the EdgeStore
    DisjointSetUnion_MT<node_id_t> d(num_vertices);
    for (const Edge& e : edge_store.get_edges()) d.merge(e.src, e.dst);
    ConnectedComponents cc(num_vertices, d);
    if (verifier) verifier->verify_connected_components(cc);
    return cc;
  }
  return MCSketchAlg::connected_components();
}

void MCGPUSketchAlg::print_subgraph_edges() {
  std::cout << "Number of inserted updates for each subgraph:\n";
  for (size_t g = 0; g < cur_subgraphs; g++)
    std::cout << "  Sub-Graph " << g << "(Sketch): " << subgraph_updates[g].load() << std::endl;
  std::cout << "  Adjacency list:      " << edge_store.get_num_edges() << std::endl;
}

void MCGPUSketchAlg::display_time() {
  for (size_t g = 0; g < cur_subgraphs; g++) {
    double ms = 0;
    uint64_t l = 0, u = 0;
    subgraphs[g]->timing(&ms, &l, &u);
    std::cout << "Subgraph #" << g << ": " << l << " profiled launches, " << ms << " ms\n";
  }
}

MinCutAnswer mc_minimum_cut(MCGPUSketchAlg& alg, int num_graphs, int k) {
  const int sketch_graphs = alg.get_num_sketch_graphs();
  for (int g = 0; g < num_graphs; g++) {
    std::vector<Edge> edges;
    if (g >= sketch_graphs) {
      edges = alg.get_adjlist_spanning_forests();
    } else {
      for (const auto& sf : alg.calc_disjoint_spanning_forests((size_t)g, (size_t)k))
        edges.insert(edges.end(), sf.get_edges().begin(), sf.get_edges().end());
    }
    MinCut mc = alg.calc_minimum_cut(edges);
    if (g >= sketch_graphs || mc.value < (size_t)k) {
      const size_t scaled = (size_t)std::llround((double)mc.value * std::pow(2.0, g));
      return {scaled, g, std::move(mc)};
    }
  }
  throw std::runtime_error("min-cut: no subgraph produced an answer");
}
