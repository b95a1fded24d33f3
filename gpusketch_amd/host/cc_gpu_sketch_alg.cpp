// CCGPUSketchAlg (the reference's src/cc_gpu_sketch_alg.cpp:1-80): its constructor, buffering and
// timing over CCSketchAlg.
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
  read_sketch_file_header(in, file, &seed, &n, &factor);
  cfg.sketches_factor(factor);
  std::unique_ptr<CCGPUSketchAlg> alg(new CCGPUSketchAlg(
      n, seed, Sketch::calc_cc_samples(n, factor), num_threads, cfg, false));
  alg->load_buckets(in, file);
  return alg.release();
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

