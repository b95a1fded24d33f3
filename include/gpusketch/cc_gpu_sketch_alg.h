// CCGPUSketchAlg: connected components of a graph stream with every vertex sketch resident in
// device memory. Keeps the reference's API (include/cc_gpu_sketch_alg.h:8-110 and the CCSketchAlg
// members it inherits, include/cc_sketch_alg.h:62-261) so it drops into GraphSketchDriver and the
// reference's tools. Updates: worker thread thr_id buffers vertex batches and submits them on its
// own device stream. Queries: Boruvka runs on the device; an eager union-find on the host answers
// connectivity without a device query while the stream only adds connectivity
// (CCSketchAlg::pre_insert, src/cc_sketch_alg.cpp:105-130).
#pragma once
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>

#include "cc_alg_configuration.h"
#include "device_sketch_store.h"
#include "dsu.h"
#include "eager_dsu.h"
#include "gpu_sketch_params.h"
#include "return_types.h"
#include "sketch.h"

class UpdateLockedException : public std::exception {
  const char* what() const noexcept override {
    return "Cannot update the algorithm: Connected components currently running";
  }
};

// query kinds for has_cached_query (include/cc_sketch_alg.h:59-63, include/mc_sketch_alg.h:60-65)
enum QueryCode { CONNECTIVITY, KSPANNINGFORESTS, MINIMUMCUT };

class CCGPUSketchAlg {
 public:
  CCGPUSketchAlg(node_id_t num_vertices, size_t num_updates, int num_threads,
                 SketchParams sketch_params, CCAlgConfiguration config = CCAlgConfiguration());
  virtual ~CCGPUSketchAlg();

  // read a file written by write_binary: u64 seed, u32 num_vertices, f64 sketches_factor, then
  // every vertex's buckets in reference order (src/cc_sketch_alg.cpp:57-91,674-683)
  static CCGPUSketchAlg* construct_from_serialized_data(const std::string& input_file,
                                                        int num_threads = 1,
                                                        CCAlgConfiguration config = {});

  // driver contract
  size_t get_desired_updates_per_batch();
  void pre_insert(GraphUpdate upd, int thr_id = 0);
  void allocate_worker_memory(size_t num_workers);
  void apply_update_batch(int thr_id, node_id_t src_vertex,
                          const std::vector<node_id_t>& dst_vertices);
  bool has_cached_query(int query_code);
  void print_configuration();

  // XOR raw reference-ordered buckets into a vertex sketch (src/cc_sketch_alg.cpp:146-149)
  void apply_raw_buckets_update(node_id_t src_vertex, Bucket* raw_buckets);
  // one undirected update applied at once; not thread-safe (src/cc_sketch_alg.cpp:152-160)
  void update(GraphUpdate upd);

  ConnectedComponents connected_components();
  bool point_query(node_id_t a, node_id_t b);
  SpanningForest calc_spanning_forest();
  std::vector<SpanningForest> calc_disjoint_spanning_forests(size_t k);
  void write_binary(const std::string& filename);

  // submit every buffered batch and wait for the device (CCGPUSketchAlg::flush_buffers)
  void flush_buffers();
  void display_time();

  // the sketch of one vertex as a host Sketch (a copy of the device buckets)
  Sketch get_sketch(node_id_t v);
  DeviceSketchStore& device_store() { return *store; }

  std::chrono::steady_clock::time_point cc_alg_start;
  std::chrono::steady_clock::time_point cc_alg_end;
  size_t last_query_rounds = 0;

  inline node_id_t get_num_vertices() { return num_vertices; }
  inline size_t get_seed() { return seed; }
  inline size_t max_rounds() { return num_samples; }
  inline bool get_update_locked() { return update_locked; }

 protected:
  CCGPUSketchAlg(node_id_t num_vertices, size_t seed, size_t num_samples, int num_threads,
                 CCAlgConfiguration config, bool dsu_valid);
  node_id_t num_vertices;
  size_t seed;
  size_t num_samples;
  CCAlgConfiguration config;
  SketchParams sketchParams;
  std::unique_ptr<DeviceSketchStore> store;
  std::atomic<bool> update_locked{false};
  EagerDSU eager;
  std::chrono::duration<double> query_time{0};
};
