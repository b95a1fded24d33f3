// GraphSketchDriver<Alg> (the reference's include/graph_sketch_driver.h:23-232): reads a
// GraphStream with one or more stream threads, calls Alg::pre_insert on every update, inserts
// both directions into the gutters, and lets worker threads hand full vertex batches to
// Alg::apply_update_batch. The Alg contract is the reference's: allocate_worker_memory,
// get_desired_updates_per_batch, get_num_vertices, pre_insert, apply_update_batch (thread-safe
// across distinct thread ids), has_cached_query, print_configuration, and, when VERIFY_SAMPLES_F
// is defined, set_verifier: the driver then records the exact graph of the stream (a
// GraphVerifier per stream thread, combined at the break point) and hands a copy to the algorithm
// after each process_stream_until (include/graph_sketch_driver.h:64,119-189); check_verifier
// compares it with an expected one.
#pragma once
#include <atomic>
#include <chrono>
#include <iostream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include <memory>
#include <mutex>

#include "driver_configuration.h"
#include "graph_stream.h"
#include "guttering.h"
#include "worker_thread_group.h"
#ifdef VERIFY_SAMPLES_F
#include "graph_verifier.h"
#endif

class DriverException : public std::exception {
  std::string err_msg;

 public:
  explicit DriverException(std::string msg) : err_msg(std::move(msg)) {}
  const char* what() const noexcept override { return err_msg.c_str(); }
};

template <class Alg>
class GraphSketchDriver {
  GutteringSystem* gts;
  Alg* sketching_alg;
  GraphStream* stream;
  WorkerThreadGroup<Alg>* worker_threads;
  size_t num_stream_threads;
  static constexpr size_t update_array_size = 4000;
  std::atomic<size_t> total_updates{0};
#ifdef VERIFY_SAMPLES_F
  std::unique_ptr<GraphVerifier> verifier;
  std::mutex verifier_mtx;
#endif

 public:
  GraphSketchDriver(Alg* alg, GraphStream* stream, DriverConfiguration config,
                    size_t num_stream_threads = 1)
      : sketching_alg(alg), stream(stream), num_stream_threads(num_stream_threads) {
    sketching_alg->allocate_worker_memory(config.get_worker_threads());
    size_t batch = sketching_alg->get_desired_updates_per_batch();
    if (config.gutter_conf().get_gutter_bytes() != GutteringConfiguration::uninit_param)
      batch = config.gutter_conf().get_gutter_bytes() / sizeof(node_id_t);
    gts = new GutteringSystem(sketching_alg->get_num_vertices(), config.get_worker_threads(),
                              batch, config.gutter_conf());
    worker_threads = new WorkerThreadGroup<Alg>(config.get_worker_threads(), this, gts);
#ifdef VERIFY_SAMPLES_F
    verifier.reset(new GraphVerifier(sketching_alg->get_num_vertices()));
#endif
    if (this->num_stream_threads > 1 && !stream->get_update_is_thread_safe()) {
      std::cerr << "WARNING: stream get_update is not thread safe. Setting number of stream "
                   "threads to 1"
                << std::endl;
      this->num_stream_threads = 1;
    }
  }

  ~GraphSketchDriver() {
    delete worker_threads;
    delete gts;
  }

  // Process updates up to (not including) break_edge_idx.
  void process_stream_until(edge_id_t break_edge_idx) {
    if (!stream->set_break_point(break_edge_idx))
      throw DriverException("Could not correctly set breakpoint: " +
                            std::to_string(break_edge_idx));
    worker_threads->resume_workers();
    auto task = [&](int thr_id) {
      std::vector<GraphStreamUpdate> buf(update_array_size);
#ifdef VERIFY_SAMPLES_F
      GraphVerifier local_verifier(sketching_alg->get_num_vertices());
#endif
      for (;;) {
        size_t got = stream->get_update_buffer(buf.data(), buf.size());
        for (size_t i = 0; i < got; i++) {
          if (buf[i].type == BREAKPOINT) {
#ifdef VERIFY_SAMPLES_F
            std::lock_guard<std::mutex> lk(verifier_mtx);
            verifier->combine(local_verifier);
#endif
            return;
          }
          GraphUpdate upd{buf[i].edge, buf[i].type};
          sketching_alg->pre_insert(upd, thr_id);
          gts->insert({upd.edge.src, upd.edge.dst}, thr_id);
          gts->insert({upd.edge.dst, upd.edge.src}, thr_id);
#ifdef VERIFY_SAMPLES_F
          local_verifier.edge_update(upd.edge);
#endif
        }
      }
    };
    std::vector<std::thread> threads;
    for (size_t i = 0; i < num_stream_threads; i++) threads.emplace_back(task, (int)i);
    for (auto& t : threads) t.join();
#ifdef VERIFY_SAMPLES_F
    sketching_alg->set_verifier(std::make_unique<GraphVerifier>(*verifier));
#endif
  }

#ifdef VERIFY_SAMPLES_F
  // the graph the driver recorded must be `expected` (throws DriverException otherwise)
  void check_verifier(const GraphVerifier& expected) {
    if (*verifier != expected)
      throw DriverException("Mismatch between driver verifier and expected verifier");
  }
#endif

  // Make the algorithm's state reflect every update processed so far.
  void prep_query(int query_code) {
    flush_start = std::chrono::steady_clock::now();
    if (!sketching_alg->has_cached_query(query_code)) {
      gts->force_flush();
      worker_threads->flush_workers();
    }
    flush_end = std::chrono::steady_clock::now();
  }

  inline void batch_callback(int thr_id, node_id_t src_vertex,
                             const std::vector<node_id_t>& dst_vertices) {
    total_updates += dst_vertices.size();
    sketching_alg->apply_update_batch(thr_id, src_vertex, dst_vertices);
  }

  size_t get_total_updates() { return total_updates.load(); }

  std::chrono::steady_clock::time_point flush_start;
  std::chrono::steady_clock::time_point flush_end;

  inline GutteringSystem* get_gts() { return gts; }
  inline WorkerThreadGroup<Alg>* get_worker_threads() { return worker_threads; }
};
