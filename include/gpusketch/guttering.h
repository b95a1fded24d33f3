// Update buffering between the stream readers and the worker threads: every directed update
// (src, dst) is appended to src's gutter; a full gutter becomes a vertex batch on a bounded work
// queue that the workers drain. This stands in for the GutteringSystem the reference driver
// builds (include/graph_sketch_driver.h:88-110; StandAloneGutters / GutterTree / CacheGuttering of
// its guttering dependency), all of whose variants hand workers the same thing: batches of
// destinations of one source vertex of at most the algorithm's desired batch size.
#pragma once
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <utility>
#include <vector>

#include "driver_configuration.h"
#include "types.h"

struct update_batch {
  node_id_t node_idx;
  std::vector<node_id_t> upd_vec;
};

class GutteringSystem {
 public:
  GutteringSystem(node_id_t num_nodes, size_t num_workers, size_t batch_updates,
                  const GutteringConfiguration& conf);

  // append upd.second to the gutter of upd.first (thread-safe)
  void insert(const std::pair<node_id_t, node_id_t>& upd, int thr_id = 0);
  // move every non-empty gutter to the work queue
  void force_flush();
  // Wait for work. Returns false once nothing is queued and the queue is released (see
  // set_non_block) or shut down.
  bool get_data(std::vector<update_batch>& out);
  // false: get_data waits for work; true: get_data returns false when the queue is empty
  void set_non_block(bool non_block);
  void shutdown();
  size_t batch_size() const { return batch_updates; }
  node_id_t num_nodes() const { return (node_id_t)gutters.size(); }

 private:
  struct Gutter {
    std::mutex mtx;
    std::vector<node_id_t> dsts;
  };
  void enqueue(update_batch&& b);

  std::vector<Gutter> gutters;
  size_t batch_updates;
  size_t batches_per_elm;
  size_t queue_cap;
  std::mutex q_mtx;
  std::condition_variable q_not_empty, q_not_full;
  std::deque<std::vector<update_batch>> queue;
  std::vector<update_batch> pending;  // batches collecting into the next queue element
  bool non_block = false;
  bool stop = false;
};
