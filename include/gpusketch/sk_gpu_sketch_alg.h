// SKGPUSketchAlg: the whole stream is buffered in pinned host memory during ingestion and then
// applied on the device, either in one submission (launch_gpu_kernel) or in buffer-sized
// submissions over the per-thread streams (launch_gpu_update). Keeps the reference's API
// (include/sk_gpu_sketch_alg.h:11-138, src/sk_gpu_sketch_alg.cpp:7-257); it derives from
// CCGPUSketchAlg (for its SketchParams constructor and timing) and so from CCSketchAlg, whose
// queries it inherits.
#pragma once
#include <atomic>
#include <mutex>
#include <vector>

#include "cc_gpu_sketch_alg.h"

class SKGPUSketchAlg : public CCGPUSketchAlg {
 public:
  // using_gpu = false asks for the reference's host update path, which this library does not
  // provide (the device is the update path); it throws std::invalid_argument.
  SKGPUSketchAlg(bool using_gpu, node_id_t num_nodes, size_t num_updates, int num_threads,
                 SketchParams sketch_params, CCAlgConfiguration config = CCAlgConfiguration());
  ~SKGPUSketchAlg() override;

  // copies the batch into the pinned stream buffer (thread-safe)
  void apply_update_batch(int thr_id, node_id_t src_vertex,
                          const std::vector<node_id_t>& dst_vertices);

  void launch_cpu_update();
  // buffered batches in submissions of num_batch_per_buffer batches over the thread streams
  void launch_gpu_update();
  // every buffered batch in one submission (one aggregation launch)
  void launch_gpu_kernel();
  // time a host-to-device copy of the whole pinned buffer (the reference's transfer probe)
  void buffer_transfer();
  void flush_buffers();
  void display_time();
  uint64_t get_batch_count() { return batch_count.load(); }

 private:
  size_t num_updates;
  size_t capacity;                 // node ids the pinned buffer holds (2 per stream update)
  node_id_t* h_edgeUpdates = nullptr;
  std::atomic<uint64_t> edgeUpdate_offset{0};
  std::atomic<uint64_t> batch_count{0};
  std::mutex batch_mutex;
  std::vector<uint32_t> batch_src;
  std::vector<uint64_t> batch_size, batch_start;
  size_t applied = 0;              // batches already submitted
  int num_host_threads;
  int num_batch_per_buffer = 540;
  std::chrono::duration<double> kernel_time{0};
};
