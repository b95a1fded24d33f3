// Sketch: one vertex's l0-sampling sketch as a host object (the reference's include/sketch.h:1-230,
// src/sketch.cpp). Bucket order is the reference's: bucket [column * bkt_per_col + depth], the
// deterministic bucket last. The device store exports and imports exactly this image, so a Sketch
// can wrap exported data (the borrowed-buffer constructor), serialize it, merge it and sample it.
// Streams of updates belong on the device (CCGPUSketchAlg and friends); Sketch::update is the
// single-index update of the reference for callers that work on one sketch at a time.
#pragma once
#include <algorithm>
#include <cmath>
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_set>

#include "bucket.h"
#include "util.h"

enum SampleResult {
  GOOD,  // sampling returned a single non-zero index
  ZERO,  // the sketched vector is zero
  FAIL   // no single non-zero index could be isolated
};

struct SketchSample {
  vec_t idx;
  SampleResult result;
};

struct ExhaustiveSketchSample {
  std::unordered_set<vec_t> idxs;
  SampleResult result;
};

class OutOfSamplesException : public std::exception {
  std::string err_msg;

 public:
  OutOfSamplesException(size_t seed, size_t num_samples, size_t sample_idx)
      : err_msg("This sketch (seed=" + std::to_string(seed) + ", max samples=" +
                std::to_string(num_samples) + ") cannot be sampled more times (cur idx=" +
                std::to_string(sample_idx) + ")!") {}
  // the same condition reported by the device query, with its message
  explicit OutOfSamplesException(std::string msg) : err_msg(std::move(msg)) {}
  const char* what() const noexcept override { return err_msg.c_str(); }
};

class Sketch {
  const uint64_t seed;
  size_t num_samples;
  size_t cols_per_sample;
  size_t num_columns;
  size_t bkt_per_col;
  size_t num_buckets;
  size_t sample_idx = 0;
  Bucket* buckets;
  bool borrowed = false;  // buckets live in a caller-owned array

  void init_geometry(vec_t vector_len, size_t samples, size_t cols);

 public:
  // n(n-1)/2: the length of the edge-incidence vector of an n-vertex graph
  static vec_t calc_vector_length(node_id_t num_vertices) {
    return (vec_t)std::ceil(double(num_vertices) * (num_vertices - 1) / 2);
  }
  // samples needed for connected components, scaled by a factor f
  static size_t calc_cc_samples(node_id_t num_vertices, double f) {
    return std::max(size_t(18), (size_t)std::ceil(f * std::log2(num_vertices) / num_samples_div));
  }
  // ceil(log2 n) + 1 depths, skipping multiples of 16
  static size_t calc_bkt_per_col(size_t n) {
    size_t b = (size_t)std::ceil(std::log2(n)) + 1;
    return b % 16 == 0 ? b + 1 : b;
  }

  // zeroed sketch over a borrowed bucket array: sketch_id's buckets start at
  // _buckets[sketch_id * num_buckets]
  Sketch(vec_t vector_len, uint64_t seed, node_id_t sketch_id, Bucket* _buckets,
         size_t num_samples = 1, size_t cols_per_sample = default_cols_per_sample);
  Sketch(vec_t vector_len, uint64_t seed, size_t num_samples = 1,
         size_t cols_per_sample = default_cols_per_sample);
  // read the serialized buckets from a stream
  Sketch(vec_t vector_len, uint64_t seed, std::istream& binary_in, size_t num_samples = 1,
         size_t cols_per_sample = default_cols_per_sample);
  Sketch(const Sketch& s);
  ~Sketch();
  Sketch& operator=(const Sketch&) = delete;

  void update(const vec_t update_idx);
  SketchSample sample();
  ExhaustiveSketchSample exhaustive_sample();

  std::mutex mutex;  // held by callers that merge into the sketch from several threads

  void merge(const Sketch& other);
  // merge only samples [start_sample, start_sample + n_samples) and the deterministic bucket
  void range_merge(const Sketch& other, size_t start_sample, size_t n_samples);
  void merge_raw_bucket_buffer(const Bucket* raw_buckets);
  void zero_contents();

  friend bool operator==(const Sketch& a, const Sketch& b);
  friend std::ostream& operator<<(std::ostream& os, const Sketch& sketch);

  void serialize(std::ostream& binary_out) const;

  inline void reset_sample_state() { sample_idx = 0; }
  inline size_t bucket_array_bytes() const { return num_buckets * sizeof(Bucket); }
  inline const Bucket* get_readonly_bucket_ptr() const { return buckets; }
  inline Bucket* get_bucket_ptr() { return buckets; }
  inline uint64_t get_seed() const { return seed; }
  inline size_t column_seed(size_t column_idx) const { return seed + column_idx * 5; }
  inline size_t checksum_seed() const { return seed; }
  inline size_t get_columns() const { return num_columns; }
  inline size_t get_buckets() const { return num_buckets; }
  inline size_t get_num_samples() const { return num_samples; }
  inline size_t get_bkt_per_col() const { return bkt_per_col; }

#ifdef L0_SAMPLING
  static constexpr size_t default_cols_per_sample = 7;
  static constexpr double num_samples_div = 0.584962500721156;  // log2(3) - 1
#else
  static constexpr size_t default_cols_per_sample = 1;
  static constexpr double num_samples_div = 1 - 0.263034406;  // 1 - log2(2 - 0.8)
#endif
};
