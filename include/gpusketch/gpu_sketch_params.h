// SketchParams (the reference's include/cuda_kernel.cuh:23-37): the geometry a caller computes
// with the Sketch sizing functions and hands to the GPU algorithms. The UVM fields are kept so
// callers compile unchanged; sketches always live in device memory owned by the library here,
// so cudaUVM_enabled / cudaUVM_buckets / d_buckets are ignored.
#pragma once
#include <cstddef>
#include <cstdint>

#include "bucket.h"
#include "sketch.h"

struct SketchParams {
  size_t num_samples = 0;
  size_t num_buckets = 0;
  size_t num_columns = 0;
  size_t bkt_per_col = 0;
  size_t seed = 0;
  bool cudaUVM_enabled = false;
  Bucket* cudaUVM_buckets = nullptr;
  Bucket* d_buckets = nullptr;
};

// The computation the reference tools repeat before constructing an algorithm
// (tools/cuda_process_stream.cu:97-116, tools/min_cut.cu:110-119).
inline SketchParams make_sketch_params(node_id_t num_vertices, uint64_t seed,
                                       double sketches_factor = 1) {
  SketchParams p;
  p.num_samples = Sketch::calc_cc_samples(num_vertices, sketches_factor);
  p.num_columns = p.num_samples * Sketch::default_cols_per_sample;
  p.bkt_per_col = Sketch::calc_bkt_per_col(Sketch::calc_vector_length(num_vertices));
  p.num_buckets = p.num_columns * p.bkt_per_col + 1;
  p.seed = seed;
  return p;
}
