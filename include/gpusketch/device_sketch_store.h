// C++ owner of one device sketch store and its submission streams, over the C ABI of
// gpusketch.h. Each worker thread id has its own stream and its own host batch buffer, so
// add_batch needs no lock across distinct thread ids (the reference gives every thread its own
// CudaStream, include/cuda_stream.h:13-292). Buffered vertex batches go to the device in one
// submission once a buffer holds flush_updates destinations.
#pragma once
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../gpusketch.h"
#include "bucket.h"
#include "types.h"

// Turns a gs_status into the reference's exception types: OutOfSamplesException for
// GS_ERR_OUT_OF_SAMPLES, std::runtime_error otherwise (the reference exits on device errors).
void gs_throw_if(int status);

class DeviceSketchStore {
 public:
  DeviceSketchStore(node_id_t num_vertices, uint64_t seed, size_t num_samples,
                    size_t num_streams, size_t flush_updates = size_t(1) << 20);
  ~DeviceSketchStore();
  DeviceSketchStore(const DeviceSketchStore&) = delete;
  DeviceSketchStore& operator=(const DeviceSketchStore&) = delete;

  // grow to at least n streams (call before concurrent use)
  void ensure_streams(size_t n);
  size_t num_streams() const { return streams.size(); }

  // buffer concat(src, d) toggles for the sketch of src (thread-safe across distinct thr)
  void add_batch(int thr, node_id_t src, const node_id_t* dsts, size_t count);
  // submit thread thr's buffered batches (asynchronous)
  void submit(int thr);
  // submit every buffer and wait for all streams
  void flush();
  // toggle whole undirected edges in both endpoints (flushes first, waits at the end)
  void toggle_edges(const std::vector<Edge>& edges);
  void toggle_pairs(const uint32_t* pairs, uint64_t nedges);

  // device Boruvka: labels (smallest vertex id per component), optional spanning forest
  std::vector<node_id_t> connected_components(std::vector<Edge>* forest, size_t* rounds);
  // k edge-disjoint spanning forests (forest i samples column i in round 0)
  std::vector<std::vector<Edge>> k_spanning_forests(size_t k);

  // reference-ordered buckets (num_buckets per vertex) of vertices [v0, v0 + count)
  void export_buckets(node_id_t v0, node_id_t count, Bucket* out);
  void import_buckets(node_id_t v0, node_id_t count, const Bucket* in);
  void zero();

  void set_profiling(bool on);
  // summed device time of the update kernels on every stream, launches, directed updates
  void timing(double* ms, uint64_t* launches, uint64_t* updates);

  const gs_geometry& geometry() const { return geo; }
  gs_store* handle() { return store; }
  uint64_t submitted_updates() const;

 private:
  struct Buffer {
    std::vector<uint32_t> src;
    std::vector<uint64_t> sizes, offsets;
    std::vector<uint32_t> dsts;
    uint64_t submitted = 0;
  };
  gs_store* store = nullptr;
  std::vector<gs_stream*> streams;
  std::vector<std::unique_ptr<Buffer>> buffers;
  gs_geometry geo{};
  size_t flush_updates;
};
