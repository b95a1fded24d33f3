// CCSketchAlg on device sketches (the reference's src/cc_sketch_alg.cpp:11-160 update and
// serialization paths, :541-683 queries, with the GraphVerifier checks it makes under
// VERIFY_SAMPLES_F, :179,550,580,592,620).
#include "gpusketch/cc_sketch_alg.h"

#include <algorithm>
#include <fstream>
#include <iostream>
#include <stdexcept>

CCSketchAlg::CCSketchAlg(node_id_t n, size_t seed, size_t samples, int num_threads,
                         CCAlgConfiguration cfg, bool valid)
    : num_vertices(n),
      seed(seed),
      num_samples(samples),
      config(cfg),
      eager(n, valid) {
  store.reset(new DeviceSketchStore(n, seed, samples, (size_t)std::max(1, num_threads)));
  const gs_geometry& g = store->geometry();
  sketchParams = SketchParams();
  sketchParams.num_samples = g.num_samples;
  sketchParams.num_columns = g.num_columns;
  sketchParams.bkt_per_col = g.bkt_per_col;
  sketchParams.num_buckets = g.num_buckets;
  sketchParams.seed = seed;
}

CCSketchAlg::CCSketchAlg(node_id_t n, size_t seed, CCAlgConfiguration cfg)
    : CCSketchAlg(n, seed, Sketch::calc_cc_samples(n, cfg.get_sketches_factor()), 1, cfg, true) {}

CCSketchAlg::~CCSketchAlg() = default;

void read_sketch_file_header(std::istream& in, const std::string& file, uint64_t* seed,
                             node_id_t* n, double* factor) {
  in.read(reinterpret_cast<char*>(seed), sizeof(*seed));
  in.read(reinterpret_cast<char*>(n), sizeof(*n));
  in.read(reinterpret_cast<char*>(factor), sizeof(*factor));
  if (!in || *n < 2) throw std::runtime_error("bad sketch file header: " + file);
}

void CCSketchAlg::load_buckets(std::istream& in, const std::string& file) {
  const size_t nb = store->geometry().num_buckets;
  const node_id_t chunk = (node_id_t)std::max<size_t>(1, (size_t(64) << 20) / (nb * sizeof(Bucket)));
  std::vector<Bucket> buf((size_t)chunk * nb);
  for (node_id_t v0 = 0; v0 < num_vertices; v0 += chunk) {
    node_id_t cnt = std::min<node_id_t>(chunk, num_vertices - v0);
    in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)((size_t)cnt * nb * sizeof(Bucket)));
    if (!in) throw std::runtime_error("truncated sketch file: " + file);
    store->import_buckets(v0, cnt, buf.data());
  }
}

CCSketchAlg* CCSketchAlg::construct_from_serialized_data(const std::string& file,
                                                         CCAlgConfiguration cfg) {
  std::ifstream in(file, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + file);
  uint64_t seed = 0;
  node_id_t n = 0;
  double factor = 1;
  read_sketch_file_header(in, file, &seed, &n, &factor);
  cfg.sketches_factor(factor);
  std::unique_ptr<CCSketchAlg> alg(
      new CCSketchAlg(n, seed, Sketch::calc_cc_samples(n, factor), 1, cfg, false));
  alg->load_buckets(in, file);
  return alg.release();
}

size_t CCSketchAlg::get_desired_updates_per_batch() {
  size_t num = sketchParams.num_buckets * sizeof(Bucket) / sizeof(node_id_t);
  return (size_t)(num * config.get_batch_factor());
}

void CCSketchAlg::pre_insert(GraphUpdate upd, int) { eager.pre_insert(upd); }

void CCSketchAlg::allocate_worker_memory(size_t num_workers) {
  store->ensure_streams(num_workers);
}

void CCSketchAlg::apply_update_batch(int thr_id, node_id_t src,
                                        const std::vector<node_id_t>& dsts) {
  if (update_locked) throw UpdateLockedException();
  store->add_batch(thr_id, src, dsts.data(), dsts.size());
}

bool CCSketchAlg::has_cached_query(int query_code) {
  return query_code == CONNECTIVITY && eager.valid();
}

void CCSketchAlg::print_configuration() { std::cout << config << std::endl; }

void CCSketchAlg::apply_raw_buckets_update(node_id_t src, Bucket* raw) {
  const size_t nb = sketchParams.num_buckets;
  std::vector<Bucket> cur(nb);
  store->export_buckets(src, 1, cur.data());
  for (size_t i = 0; i < nb; i++) Bucket_Boruvka::update(cur[i], raw[i].alpha, raw[i].gamma);
  store->import_buckets(src, 1, cur.data());
  eager.invalidate();
}

void CCSketchAlg::update(GraphUpdate upd) {
  pre_insert(upd, 0);
  store->toggle_edges({upd.edge});
}

ConnectedComponents CCSketchAlg::connected_components() {
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
    labels = store->connected_components(&forest, &last_query_rounds);
  } catch (...) {
    update_locked = false;
    throw;
  }
  eager.adopt(forest);
  update_locked = false;
  ConnectedComponents cc(num_vertices, labels.data());
  if (verifier) {  // every sampled edge is a graph edge, and the partition is the graph's
    for (const Edge& e : forest) verifier->verify_edge(e);
    verifier->verify_connected_components(cc);
  }
  cc_alg_end = std::chrono::steady_clock::now();
  query_time += cc_alg_end - cc_alg_start;
  return cc;
}

bool CCSketchAlg::point_query(node_id_t a, node_id_t b) {
  return connected_components().is_connected(a, b);
}

SpanningForest CCSketchAlg::calc_spanning_forest() {
  connected_components();
  SpanningForest sf = eager.spanning_forest();
  if (verifier) verifier->verify_spanning_forests(std::vector<SpanningForest>{sf});
  return sf;
}

std::vector<SpanningForest> CCSketchAlg::calc_disjoint_spanning_forests(size_t k) {
  std::vector<SpanningForest> sfs;
  std::vector<Edge> removed;
  size_t max_rounds = 0;
  // the forests are checked together at the end (forest i spans the graph minus forests
  // 0..i-1), not by each connected_components call on the thinned sketches
  struct Aside {
    std::unique_ptr<GraphVerifier>& slot;
    std::unique_ptr<GraphVerifier> held;
    ~Aside() { slot = std::move(held); }
  } aside{verifier, std::move(verifier)};
  for (size_t i = 0; i < k; i++) {
    connected_components();
    sfs.push_back(eager.spanning_forest());
    max_rounds = std::max(max_rounds, last_query_rounds);
    const auto& e = sfs.back().get_edges();
    store->toggle_edges(e);  // delete this forest before sampling the next one
    removed.insert(removed.end(), e.begin(), e.end());
    eager.invalidate();
  }
  store->toggle_edges(removed);  // and put every forest back
  eager.invalidate();
  last_query_rounds = max_rounds;
  if (aside.held) aside.held->verify_spanning_forests(sfs);
  return sfs;
}

void CCSketchAlg::write_binary(const std::string& filename) {
  std::ofstream out(filename, std::ios::binary | std::ios::trunc);
  if (!out) throw std::runtime_error("cannot write " + filename);
  const uint64_t s = seed;
  const double factor = config.get_sketches_factor();
  out.write(reinterpret_cast<const char*>(&s), sizeof(s));
  out.write(reinterpret_cast<const char*>(&num_vertices), sizeof(num_vertices));
  out.write(reinterpret_cast<const char*>(&factor), sizeof(factor));
  const size_t nb = sketchParams.num_buckets;
  const node_id_t chunk = (node_id_t)std::max<size_t>(1, (size_t(64) << 20) / (nb * sizeof(Bucket)));
  std::vector<Bucket> buf((size_t)chunk * nb);
  for (node_id_t v0 = 0; v0 < num_vertices; v0 += chunk) {
    node_id_t cnt = std::min<node_id_t>(chunk, num_vertices - v0);
    store->export_buckets(v0, cnt, buf.data());
    out.write(reinterpret_cast<const char*>(buf.data()),
              (std::streamsize)((size_t)cnt * nb * sizeof(Bucket)));
  }
  if (!out) throw std::runtime_error("write failed: " + filename);
}

Sketch CCSketchAlg::get_sketch(node_id_t v) {
  Sketch s(Sketch::calc_vector_length(num_vertices), seed, num_samples);
  store->export_buckets(v, 1, s.get_bucket_ptr());
  return s;
}
