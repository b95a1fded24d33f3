// CCGPUSketchAlg: the reference's GPU connected-components algorithm
// (include/cc_gpu_sketch_alg.h:8-110, src/cc_gpu_sketch_alg.cpp:1-80) over CCSketchAlg, whose
// sketches, queries and serialization it inherits (all on the device). It keeps the reference's
// constructor (num_updates, num_threads, SketchParams), flush_buffers and display_time, so it
// drops into GraphSketchDriver and the reference's tools.
#pragma once
#include "cc_sketch_alg.h"

class CCGPUSketchAlg : public CCSketchAlg {
 public:
  CCGPUSketchAlg(node_id_t num_vertices, size_t num_updates, int num_threads,
                 SketchParams sketch_params, CCAlgConfiguration config = CCAlgConfiguration());
  ~CCGPUSketchAlg() override;

  static CCGPUSketchAlg* construct_from_serialized_data(const std::string& input_file,
                                                        int num_threads = 1,
                                                        CCAlgConfiguration config = {});

  // submit every buffered batch and wait for the device (CCGPUSketchAlg::flush_buffers)
  void flush_buffers();
  void display_time();

 protected:
  CCGPUSketchAlg(node_id_t num_vertices, size_t seed, size_t num_samples, int num_threads,
                 CCAlgConfiguration config, bool dsu_valid)
      : CCSketchAlg(num_vertices, seed, num_samples, num_threads, config, dsu_valid) {}
};
