// In-memory vertex gutters and the bounded work queue (see guttering.h).
#include "gpusketch/guttering.h"

GutteringSystem::GutteringSystem(node_id_t num_nodes, size_t num_workers, size_t batch,
                                 const GutteringConfiguration& conf)
    : gutters(num_nodes),
      batch_updates(batch ? batch : 1),
      batches_per_elm(conf.get_wq_batch_per_elm()),
      queue_cap(conf.get_queue_factor() * (num_workers ? num_workers : 1)) {}

void GutteringSystem::enqueue(update_batch&& b) {
  std::unique_lock<std::mutex> lk(q_mtx);
  pending.push_back(std::move(b));
  if (pending.size() < batches_per_elm) return;
  q_not_full.wait(lk, [&] { return queue.size() < queue_cap || stop; });
  queue.push_back(std::move(pending));
  pending.clear();
  q_not_empty.notify_one();
}

void GutteringSystem::insert(const std::pair<node_id_t, node_id_t>& upd, int) {
  Gutter& g = gutters[upd.first];
  std::vector<node_id_t> full;
  {
    std::lock_guard<std::mutex> lk(g.mtx);
    if (g.dsts.capacity() < batch_updates) g.dsts.reserve(batch_updates);
    g.dsts.push_back(upd.second);
    if (g.dsts.size() < batch_updates) return;
    full.swap(g.dsts);
  }
  enqueue({upd.first, std::move(full)});
}

void GutteringSystem::force_flush() {
  for (node_id_t v = 0; v < (node_id_t)gutters.size(); v++) {
    std::vector<node_id_t> part;
    {
      std::lock_guard<std::mutex> lk(gutters[v].mtx);
      if (gutters[v].dsts.empty()) continue;
      part.swap(gutters[v].dsts);
    }
    enqueue({v, std::move(part)});
  }
  std::lock_guard<std::mutex> lk(q_mtx);
  if (!pending.empty()) {
    queue.push_back(std::move(pending));
    pending.clear();
    q_not_empty.notify_one();
  }
}

bool GutteringSystem::get_data(std::vector<update_batch>& out) {
  std::unique_lock<std::mutex> lk(q_mtx);
  q_not_empty.wait(lk, [&] { return !queue.empty() || non_block || stop; });
  if (queue.empty()) return false;
  out = std::move(queue.front());
  queue.pop_front();
  q_not_full.notify_one();
  return true;
}

void GutteringSystem::set_non_block(bool nb) {
  std::lock_guard<std::mutex> lk(q_mtx);
  non_block = nb;
  q_not_empty.notify_all();
}

void GutteringSystem::shutdown() {
  std::lock_guard<std::mutex> lk(q_mtx);
  stop = true;
  q_not_empty.notify_all();
  q_not_full.notify_all();
}
