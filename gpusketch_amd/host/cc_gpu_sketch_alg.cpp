// CCGPUSketchAlg (the reference's src/cc_gpu_sketch_alg.cpp:1-80 plus the CCSketchAlg query and
// serialization paths it inherits, src/cc_sketch_alg.cpp:57-160,541-683).
#include "gpusketch/cc_gpu_sketch_alg.h"

#include <algorithm>
#include <fstream>
#include <iostream>
#include <stdexcept>

namespace {
size_t samples_for(node_id_t n, const SketchParams& p, const CCAlgConfiguration& c) {
  return p.num_samples ? p.num_samples : Sketch::calc_cc_samples(n, c.get_sketches_factor());
}
}  // namespace

CCGPUSketchAlg::CCGPUSketchAlg(node_id_t n, size_t seed, size_t samples, int num_threads,
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

CCGPUSketchAlg::CCGPUSketchAlg(node_id_t n, size_t /*num_updates*/, int num_threads,
                               SketchParams p, CCAlgConfiguration cfg)
    : CCGPUSketchAlg(n, p.seed, samples_for(n, p, cfg), num_threads, cfg, true) {
  const size_t b = Sketch::calc_bkt_per_col(Sketch::calc_vector_length(n));
  if ((p.bkt_per_col && p.bkt_per_col != b) ||
      (p.num_columns && p.num_columns != num_samples * Sketch::default_cols_per_sample) ||
      (p.num_buckets && p.num_buckets != sketchParams.num_buckets))
    throw std::invalid_argument("SketchParams do not match the sketch geometry of " +
                                std::to_string(n) + " vertices");
  sketchParams.cudaUVM_enabled = p.cudaUVM_enabled;
}

CCGPUSketchAlg::~CCGPUSketchAlg() = default;

CCGPUSketchAlg* CCGPUSketchAlg::construct_from_serialized_data(const std::string& file,
                                                               int num_threads,
                                                               CCAlgConfiguration cfg) {
  std::ifstream in(file, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open " + file);
  uint64_t seed = 0;
  node_id_t n = 0;
  double factor = 1;
  in.read(reinterpret_cast<char*>(&seed), sizeof(seed));
  in.read(reinterpret_cast<char*>(&n), sizeof(n));
  in.read(reinterpret_cast<char*>(&factor), sizeof(factor));
  if (!in || n < 2) throw std::runtime_error("bad sketch file header: " + file);
  cfg.sketches_factor(factor);
  auto* alg = new CCGPUSketchAlg(n, seed, Sketch::calc_cc_samples(n, factor), num_threads, cfg,
                                 false);
  const size_t nb = alg->store->geometry().num_buckets;
  const node_id_t chunk = (node_id_t)std::max<size_t>(1, (size_t(64) << 20) / (nb * sizeof(Bucket)));
  std::vector<Bucket> buf((size_t)chunk * nb);
  for (node_id_t v0 = 0; v0 < n; v0 += chunk) {
    node_id_t cnt = std::min<node_id_t>(chunk, n - v0);
    in.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)((size_t)cnt * nb * sizeof(Bucket)));
    if (!in) {
      delete alg;
      throw std::runtime_error("truncated sketch file: " + file);
    }
    alg->store->import_buckets(v0, cnt, buf.data());
  }
  return alg;
}

size_t CCGPUSketchAlg::get_desired_updates_per_batch() {
  size_t num = sketchParams.num_buckets * sizeof(Bucket) / sizeof(node_id_t);
  return (size_t)(num * config.get_batch_factor());
}

void CCGPUSketchAlg::pre_insert(GraphUpdate upd, int) { eager.pre_insert(upd); }

void CCGPUSketchAlg::allocate_worker_memory(size_t num_workers) {
  store->ensure_streams(num_workers);
}

void CCGPUSketchAlg::apply_update_batch(int thr_id, node_id_t src,
                                        const std::vector<node_id_t>& dsts) {
  if (update_locked) throw UpdateLockedException();
  store->add_batch(thr_id, src, dsts.data(), dsts.size());
}

bool CCGPUSketchAlg::has_cached_query(int query_code) {
  return query_code == CONNECTIVITY && eager.valid();
}

void CCGPUSketchAlg::print_configuration() { std::cout << config << std::endl; }

void CCGPUSketchAlg::apply_raw_buckets_update(node_id_t src, Bucket* raw) {
  const size_t nb = sketchParams.num_buckets;
  std::vector<Bucket> cur(nb);
  store->export_buckets(src, 1, cur.data());
  for (size_t i = 0; i < nb; i++) Bucket_Boruvka::update(cur[i], raw[i].alpha, raw[i].gamma);
  store->import_buckets(src, 1, cur.data());
  eager.invalidate();
}

void CCGPUSketchAlg::update(GraphUpdate upd) {
  pre_insert(upd, 0);
  store->toggle_edges({upd.edge});
}

ConnectedComponents CCGPUSketchAlg::connected_components() {
  cc_alg_start = std::chrono::steady_clock::now();
  if (eager.valid()) {
    ConnectedComponents cc = eager.components();
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
  cc_alg_end = std::chrono::steady_clock::now();
  query_time += cc_alg_end - cc_alg_start;
  return cc;
}

bool CCGPUSketchAlg::point_query(node_id_t a, node_id_t b) {
  return connected_components().is_connected(a, b);
}

SpanningForest CCGPUSketchAlg::calc_spanning_forest() {
  connected_components();
  return eager.spanning_forest();
}

std::vector<SpanningForest> CCGPUSketchAlg::calc_disjoint_spanning_forests(size_t k) {
  std::vector<SpanningForest> sfs;
  std::vector<Edge> removed;
  size_t max_rounds = 0;
  for (size_t i = 0; i < k; i++) {
    sfs.push_back(calc_spanning_forest());
    max_rounds = std::max(max_rounds, last_query_rounds);
    const auto& e = sfs.back().get_edges();
    store->toggle_edges(e);  // delete this forest before sampling the next one
    removed.insert(removed.end(), e.begin(), e.end());
    eager.invalidate();
  }
  store->toggle_edges(removed);  // and put every forest back
  eager.invalidate();
  last_query_rounds = max_rounds;
  return sfs;
}

void CCGPUSketchAlg::write_binary(const std::string& filename) {
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

void CCGPUSketchAlg::flush_buffers() { store->flush(); }

void CCGPUSketchAlg::display_time() {
  double ms = 0;
  uint64_t launches = 0, upd = 0;
  store->timing(&ms, &launches, &upd);
  std::cout << "Device update kernels: " << launches << " launches, " << ms << " ms (profiled)"
            << std::endl;
  std::cout << "Directed updates submitted: " << store->submitted_updates() << std::endl;
  std::cout << "Device query time (s): " << query_time.count() << std::endl;
}

Sketch CCGPUSketchAlg::get_sketch(node_id_t v) {
  Sketch s(Sketch::calc_vector_length(num_vertices), seed, num_samples);
  store->export_buckets(v, 1, s.get_bucket_ptr());
  return s;
}
