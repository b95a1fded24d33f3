// MCSketchAlg on device sketch graphs (the reference's src/mc_sketch_alg.cpp: sketch graphs
// :93-135, updates :151-194, queries :738-1003, serialization :43-56,1004-1013, with the
// GraphVerifier checks of its VERIFY_SAMPLES_F builds, :784,814,832,910).
#include "gpusketch/mc_sketch_alg.h"

#include <algorithm>
#include <fstream>
#include <stdexcept>

MCSketchAlg::MCSketchAlg(node_id_t n, size_t seed, size_t samples, int max_graphs, int threads,
                         CCAlgConfiguration cfg)
    : num_vertices(n),
      seed(seed),
      num_samples(samples),
      max_sketch_graphs(max_graphs),
      num_host_threads(std::max(1, threads)),
      config(cfg),
      subgraphs((size_t)std::max(0, max_graphs)),
      subgraph_updates((size_t)std::max(0, max_graphs)),
      eager(n, true) {
  for (auto& u : subgraph_updates) u = 0;
}

MCSketchAlg::MCSketchAlg(node_id_t n, size_t seed, int max_graphs, CCAlgConfiguration cfg)
    : MCSketchAlg(n, seed, Sketch::calc_cc_samples(n, cfg.get_sketches_factor()), max_graphs, 1,
                  cfg) {}

MCSketchAlg::~MCSketchAlg() = default;

void MCSketchAlg::ensure_subgraphs(size_t count) {
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

void MCSketchAlg::create_sketch_graph(int graph_id, const SketchParams p) {
  if (graph_id < 0 || graph_id >= max_sketch_graphs || (size_t)graph_id != cur_subgraphs)
    throw std::invalid_argument("create_sketch_graph: graph " + std::to_string(graph_id) +
                                " is not the next sketch graph below " +
                                std::to_string(max_sketch_graphs));
  const size_t b = Sketch::calc_bkt_per_col(Sketch::calc_vector_length(num_vertices));
  if ((p.bkt_per_col && p.bkt_per_col != b) ||
      (p.num_samples && p.num_samples != num_samples))
    throw std::invalid_argument("create_sketch_graph: SketchParams do not match the geometry of " +
                                std::to_string(num_vertices) + " vertices");
  ensure_subgraphs((size_t)graph_id + 1);
}

MCSketchAlg* MCSketchAlg::construct_from_serialized_data(const std::string& file,
                                                         CCAlgConfiguration cfg) {
  std::ifstream in(file, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + file);
  uint64_t seed = 0;
  node_id_t n = 0;
  double factor = 1;
  read_sketch_file_header(in, file, &seed, &n, &factor);
  cfg.sketches_factor(factor);
  std::unique_ptr<MCSketchAlg> alg(new MCSketchAlg(n, seed, 1, cfg));
  alg->ensure_subgraphs(1);
  DeviceSketchStore& st = *alg->subgraphs[0];
  const size_t nb = st.geometry().num_buckets;
  const node_id_t chunk = (node_id_t)std::max<size_t>(1, (size_t(64) << 20) / (nb * sizeof(Bucket)));
  std::vector<Bucket> buf((size_t)chunk * nb);
  for (node_id_t v0 = 0; v0 < n; v0 += chunk) {
    const node_id_t cnt = std::min<node_id_t>(chunk, n - v0);
    in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)((size_t)cnt * nb * sizeof(Bucket)));
    if (!in) throw std::runtime_error("truncated sketch file: " + file);
    st.import_buckets(v0, cnt, buf.data());
  }
  alg->eager.invalidate();
  return alg.release();
}

size_t MCSketchAlg::get_desired_updates_per_batch() {
  const size_t nb = Sketch::calc_bkt_per_col(Sketch::calc_vector_length(num_vertices)) *
                        num_samples * Sketch::default_cols_per_sample + 1;
  return (size_t)(nb * sizeof(Bucket) / sizeof(node_id_t) * config.get_batch_factor());
}

void MCSketchAlg::allocate_worker_memory(size_t num_workers) {
  if ((int)num_workers > num_host_threads) {
    num_host_threads = (int)num_workers;
    for (size_t g = 0; g < cur_subgraphs; g++) subgraphs[g]->ensure_streams(num_workers);
  }
}

void MCSketchAlg::apply_update_batch(int thr, node_id_t src, const std::vector<node_id_t>& dsts) {
  apply_update_batch_single_graph(thr, 0, src, dsts);
}

void MCSketchAlg::apply_update_batch_single_graph(int thr, int graph_id, node_id_t src,
                                                  const std::vector<node_id_t>& dsts) {
  if (update_locked) throw UpdateLockedException();
  if ((size_t)graph_id >= cur_subgraphs)
    throw std::out_of_range("sketch graph " + std::to_string(graph_id) + " does not exist");
  subgraphs[(size_t)graph_id]->add_batch(thr, src, dsts.data(), dsts.size());
  subgraph_updates[(size_t)graph_id] += dsts.size();
}

void MCSketchAlg::apply_raw_buckets_update(node_id_t src, Bucket* raw) {
  DeviceSketchStore& st = subgraph_store(0);
  const size_t nb = st.geometry().num_buckets;
  std::vector<Bucket> cur(nb);
  st.export_buckets(src, 1, cur.data());
  for (size_t i = 0; i < nb; i++) Bucket_Boruvka::update(cur[i], raw[i].alpha, raw[i].gamma);
  st.import_buckets(src, 1, cur.data());
  eager.invalidate();
}

void MCSketchAlg::update_subgraph(int graph_id, GraphUpdate upd) {
  pre_insert(upd, 0);
  subgraph_store(graph_id).toggle_edges({upd.edge});
}

ConnectedComponents MCSketchAlg::connected_components() {
  cc_alg_start = std::chrono::steady_clock::now();
  if (eager.valid()) {
    ConnectedComponents cc = eager.components();
    if (verifier) verifier->verify_connected_components(cc);
    cc_alg_end = std::chrono::steady_clock::now();
    return cc;
  }
  update_locked = true;
  std::vector<Edge> forest;
  std::vector<node_id_t> labels;
  try {
    labels = subgraph_store(0).connected_components(&forest, &last_query_rounds);
  } catch (...) {
    update_locked = false;
    throw;
  }
  eager.adopt(forest);
  update_locked = false;
  ConnectedComponents cc(num_vertices, labels.data());
  if (verifier) {
    for (const Edge& e : forest) verifier->verify_edge(e);
    verifier->verify_connected_components(cc);
  }
  cc_alg_end = std::chrono::steady_clock::now();
  query_time += cc_alg_end - cc_alg_start;
  return cc;
}

bool MCSketchAlg::point_query(node_id_t a, node_id_t b) {
  return connected_components().is_connected(a, b);
}

SpanningForest MCSketchAlg::calc_spanning_forest(size_t graph_id) {
  std::vector<Edge> forest;
  subgraph_store((int)graph_id).connected_components(&forest, &last_query_rounds);
  SpanningForest sf(num_vertices, std::move(forest));
  if (verifier && graph_id == 0) verifier->verify_spanning_forests(std::vector<SpanningForest>{sf});
  return sf;
}

std::vector<SpanningForest> MCSketchAlg::calc_disjoint_spanning_forests(size_t graph_id,
                                                                        size_t kk) {
  sf_query_start = std::chrono::steady_clock::now();
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
  if (verifier && graph_id == 0) verifier->verify_spanning_forests(out);
  sf_query_end = std::chrono::steady_clock::now();
  query_time += sf_query_end - sf_query_start;
  return out;
}

MinCut MCSketchAlg::calc_minimum_cut(const std::vector<Edge>& edges) {
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

void MCSketchAlg::write_binary(const std::string& filename) {
  std::ofstream out(filename, std::ios::binary | std::ios::trunc);
  if (!out) throw std::runtime_error("cannot write " + filename);
  const uint64_t s = seed;
  const double factor = config.get_sketches_factor();
  out.write(reinterpret_cast<const char*>(&s), sizeof(s));
  out.write(reinterpret_cast<const char*>(&num_vertices), sizeof(num_vertices));
  out.write(reinterpret_cast<const char*>(&factor), sizeof(factor));
  DeviceSketchStore& st = subgraph_store(0);
  const size_t nb = st.geometry().num_buckets;
  const node_id_t chunk = (node_id_t)std::max<size_t>(1, (size_t(64) << 20) / (nb * sizeof(Bucket)));
  std::vector<Bucket> buf((size_t)chunk * nb);
  for (node_id_t v0 = 0; v0 < num_vertices; v0 += chunk) {
    const node_id_t cnt = std::min<node_id_t>(chunk, num_vertices - v0);
    st.export_buckets(v0, cnt, buf.data());
    out.write(reinterpret_cast<const char*>(buf.data()),
              (std::streamsize)((size_t)cnt * nb * sizeof(Bucket)));
  }
  if (!out) throw std::runtime_error("write failed: " + filename);
}
