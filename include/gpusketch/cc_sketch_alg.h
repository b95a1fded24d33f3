// CCSketchAlg: connected components of a graph stream from one l0 sketch per vertex, the base of
// the GPU algorithms (the reference's include/cc_sketch_alg.h:62-261, src/cc_sketch_alg.cpp). The
// reference's base class keeps its sketches on the host and updates them on the CPU; here every
// vertex sketch is resident in device memory (DeviceSketchStore) and every path -- updates,
// Boruvka queries, serialization -- runs through the device, so CCGPUSketchAlg and SKGPUSketchAlg
// add only their constructors, buffering and timing. Code that holds a CCSketchAlg* (or
// instantiates GraphSketchDriver<CCSketchAlg>) works on either.
//
// Updates: worker thread thr_id buffers vertex batches and submits them on its own device
// stream. Queries: Boruvka on the device; an eager union-find on the host answers connectivity
// without a device query while the stream only adds connectivity (CCSketchAlg::pre_insert,
// src/cc_sketch_alg.cpp:105-130). With a GraphVerifier set (the reference's VERIFY_SAMPLES_F
// builds set one through the driver), every answer is checked against the exact graph.
#pragma once
#include <atomic>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cc_alg_configuration.h"
#include "device_sketch_store.h"
#include "dsu.h"
#include "eager_dsu.h"
#include "gpu_sketch_params.h"
#include "graph_verifier.h"
#include "return_types.h"
#include "sketch.h"

class UpdateLockedException : public std::exception {
  const char* what() const noexcept override {
    return "Cannot update the algorithm: Connected components currently running";
  }
};

// query kinds for has_cached_query (include/cc_sketch_alg.h:59-63, include/mc_sketch_alg.h:60-65)
enum QueryCode { CONNECTIVITY, KSPANNINGFORESTS, MINIMUMCUT };

class CCSketchAlg {
 public:
  // include/cc_sketch_alg.h: CCSketchAlg(num_vertices, seed, config); one device stream until
  // allocate_worker_memory asks for more
  CCSketchAlg(node_id_t num_vertices, size_t seed, CCAlgConfiguration config = CCAlgConfiguration());
  virtual ~CCSketchAlg();

  // read a file written by write_binary: u64 seed, u32 num_vertices, f64 sketches_factor, then
  // every vertex's buckets in reference order (src/cc_sketch_alg.cpp:57-91,674-683)
  static CCSketchAlg* construct_from_serialized_data(const std::string& input_file,
                                                     CCAlgConfiguration config = CCAlgConfiguration());

  // driver contract
  size_t get_desired_updates_per_batch();
  void pre_insert(GraphUpdate upd, int thr_id = 0);
  void allocate_worker_memory(size_t num_workers);
  void apply_update_batch(int thr_id, node_id_t src_vertex,
                          const std::vector<node_id_t>& dst_vertices);
  bool has_cached_query(int query_code);
  void print_configuration();
  // the exact graph the driver saw (include/graph_sketch_driver.h:64,189); answers are checked
  // against it while it is set
  void set_verifier(std::unique_ptr<GraphVerifier> verifier) { this->verifier = std::move(verifier); }

  // XOR raw reference-ordered buckets into a vertex sketch (src/cc_sketch_alg.cpp:146-149)
  void apply_raw_buckets_update(node_id_t src_vertex, Bucket* raw_buckets);
  // one undirected update applied at once; not thread-safe (src/cc_sketch_alg.cpp:152-160)
  void update(GraphUpdate upd);

  ConnectedComponents connected_components();
  bool point_query(node_id_t a, node_id_t b);
  SpanningForest calc_spanning_forest();
  std::vector<SpanningForest> calc_disjoint_spanning_forests(size_t k);
  void write_binary(const std::string& filename);

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
  CCSketchAlg(node_id_t num_vertices, size_t seed, size_t num_samples, int num_threads,
              CCAlgConfiguration config, bool dsu_valid);
  // overwrite the sketches with the buckets of a write_binary file positioned after its header
  void load_buckets(std::istream& in, const std::string& file);

  node_id_t num_vertices;
  size_t seed;
  size_t num_samples;
  CCAlgConfiguration config;
  SketchParams sketchParams;
  std::unique_ptr<DeviceSketchStore> store;
  std::atomic<bool> update_locked{false};
  EagerDSU eager;
  std::unique_ptr<GraphVerifier> verifier;
  std::chrono::duration<double> query_time{0};
};

// the header of a write_binary file: seed, vertex count, sketches factor
void read_sketch_file_header(std::istream& in, const std::string& file, uint64_t* seed,
                             node_id_t* n, double* factor);
