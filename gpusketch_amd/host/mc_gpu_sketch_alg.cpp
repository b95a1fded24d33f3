// MCGPUSketchAlg (src/mc_gpu_sketch_alg.cpp:1-209, src/mc_sketch_alg.cpp:738-1003 and
// tools/min_cut.cu:205-252 of the reference).
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
    : num_vertices(n),
      seed(p.seed),
      num_samples(p.num_samples ? p.num_samples
                                : Sketch::calc_cc_samples(n, cfg.get_sketches_factor())),
      num_subgraphs((size_t)num_subgraphs),
      max_sketch_graphs(max_sketch_graphs),
      k(k),
      num_host_threads(std::max(1, num_threads)),
      config(cfg),
      edge_store(p.seed, n, sketch_bytes, (size_t)num_subgraphs, (size_t)initial_sketch_graphs),
      subgraphs((size_t)std::max(0, max_sketch_graphs)),
      subgraph_updates((size_t)std::max(0, max_sketch_graphs)),
      eager(n, true) {
  if (max_sketch_graphs < initial_sketch_graphs)
    throw std::invalid_argument("Cannot have initial sketch graphs > max sketch graphs");
  if (num_subgraphs < 1 || num_subgraphs > 64)
    throw std::invalid_argument("num_subgraphs must be in [1, 64]");
  for (auto& u : subgraph_updates) u = 0;
  ensure_subgraphs((size_t)initial_sketch_graphs);
}

MCGPUSketchAlg::~MCGPUSketchAlg() = default;

void MCGPUSketchAlg::ensure_subgraphs(size_t count) {
  if (cur_subgraphs >= count) return;
  std::lock_guard<std::mutex> lk(sketch_creation_lock);
  while (cur_subgraphs < count) {
    const size_t g = cur_subgraphs;
    if (g >= subgraphs.size())
      throw std::length_error("min-cut: more sketch subgraphs needed than max_sketch_graphs");
    subgraphs[g].reset(new DeviceSketchStore(num_vertices, seed, num_samples,
                                             (size_t)num_host_threads));
    cur_subgraphs++;  // last, so readers only see initialized stores
  }
}

size_t MCGPUSketchAlg::get_desired_updates_per_batch() {
  const size_t nb = Sketch::calc_bkt_per_col(Sketch::calc_vector_length(num_vertices)) *
                        num_samples * Sketch::default_cols_per_sample + 1;
  return (size_t)(nb * sizeof(Bucket) / sizeof(node_id_t) * config.get_batch_factor());
}

void MCGPUSketchAlg::allocate_worker_memory(size_t num_workers) {
  if ((int)num_workers > num_host_threads) {
    num_host_threads = (int)num_workers;
    for (size_t g = 0; g < cur_subgraphs; g++) subgraphs[g]->ensure_streams(num_workers);
  }
}

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
      }
    });
  for (auto& th : threads) th.join();
  for (size_t g = 0; g < cur_subgraphs; g++) subgraphs[g]->flush();
}

ConnectedComponents MCGPUSketchAlg::connected_components() {
  if (eager.valid()) return eager.components();
  if (cur_subgraphs == 0) {  // every edge is in the EdgeStore
    DisjointSetUnion_MT<node_id_t> d(num_vertices);
    for (const Edge& e : edge_store.get_edges()) d.merge(e.src, e.dst);
    return ConnectedComponents(num_vertices, d);
  }
  update_locked = true;
  std::vector<Edge> forest;
  std::vector<node_id_t> labels;
  try {
    labels = subgraphs[0]->connected_components(&forest, nullptr);
  } catch (...) {
    update_locked = false;
    throw;
  }
  eager.adopt(forest);
  update_locked = false;
  return ConnectedComponents(num_vertices, labels.data());
}

SpanningForest MCGPUSketchAlg::calc_spanning_forest(size_t graph_id) {
  std::vector<Edge> forest;
  subgraph_store((int)graph_id).connected_components(&forest, nullptr);
  return SpanningForest(num_vertices, std::move(forest));
}

std::vector<SpanningForest> MCGPUSketchAlg::calc_disjoint_spanning_forests(size_t graph_id,
                                                                           size_t kk) {
  update_locked = true;
  std::vector<std::vector<Edge>> f;
  try {
    f = subgraph_store((int)graph_id).k_spanning_forests(kk);
  } catch (...) {
    update_locked = false;
    throw;
  }
  update_locked = false;
  std::vector<SpanningForest> out;
  for (auto& e : f) out.emplace_back(num_vertices, std::move(e));
  return out;
}

MinCut MCGPUSketchAlg::calc_minimum_cut(const std::vector<Edge>& edges) {
  std::vector<uint32_t> pairs(2 * edges.size());
  for (size_t i = 0; i < edges.size(); i++) {
    pairs[2 * i] = edges[i].src;
    pairs[2 * i + 1] = edges[i].dst;
  }
  uint64_t value = 0;
  std::vector<uint8_t> side(num_vertices);
  gs_throw_if(gs_min_cut(num_vertices, pairs.data(), edges.size(), &value, side.data()));
  MinCut mc;
  mc.value = value;
  for (node_id_t v = 0; v < num_vertices; v++)
    (side[v] ? mc.left_vertices : mc.right_vertices).insert(v);
  return mc;
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
