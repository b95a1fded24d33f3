// Worker threads that drain the gutters into the algorithm (the reference's
// include/worker_thread_group.h:16-169): each worker has a thread id, and every vertex batch it
// takes goes to driver->batch_callback(id, src, dsts). flush_workers() returns once every queued
// batch has been applied.
#pragma once
#include <atomic>
#include <condition_variable>
#include <mutex>
#include <thread>
#include <vector>

#include "guttering.h"

template <class Alg>
class GraphSketchDriver;

template <class Alg>
class WorkerThreadGroup {
 public:
  WorkerThreadGroup(size_t num_workers, GraphSketchDriver<Alg>* driver, GutteringSystem* gts)
      : driver(driver), gts(gts) {
    for (size_t i = 0; i < num_workers; i++) threads.emplace_back([this, i] { work((int)i); });
  }
  ~WorkerThreadGroup() {
    {
      std::lock_guard<std::mutex> lk(mtx);
      shutdown_flag = true;
      resume_cv.notify_all();
    }
    gts->shutdown();
    for (auto& t : threads) t.join();
  }
  WorkerThreadGroup(const WorkerThreadGroup&) = delete;
  WorkerThreadGroup& operator=(const WorkerThreadGroup&) = delete;

  // Apply everything queued so far; returns with the workers waiting for new work.
  void flush_workers() {
    gts->set_non_block(true);
    {
      std::unique_lock<std::mutex> lk(mtx);
      idle_cv.wait(lk, [&] { return idle == threads.size(); });
    }
    gts->set_non_block(false);
    resume_workers();
  }

  // Wake idle workers and wait until every one of them is back waiting on the queue.
  void resume_workers() {
    std::unique_lock<std::mutex> lk(mtx);
    if (idle == 0) return;
    epoch++;
    resume_cv.notify_all();
    idle_cv.wait(lk, [&] { return idle == 0; });
  }

  size_t num_workers() const { return threads.size(); }

 private:
  void work(int id) {
    std::vector<update_batch> elm;
    for (;;) {
      if (gts->get_data(elm)) {
        for (auto& b : elm)
          if (!b.upd_vec.empty()) driver->batch_callback(id, b.node_idx, b.upd_vec);
        continue;
      }
      // the queue was released empty (a flush) or shut down: report idle, wait for a resume
      std::unique_lock<std::mutex> lk(mtx);
      if (shutdown_flag) return;
      idle++;
      idle_cv.notify_all();
      const size_t seen = epoch;
      resume_cv.wait(lk, [&] { return epoch != seen || shutdown_flag; });
      idle--;
      idle_cv.notify_all();
      if (shutdown_flag) return;
    }
  }

  GraphSketchDriver<Alg>* driver;
  GutteringSystem* gts;
  std::vector<std::thread> threads;
  std::mutex mtx;
  std::condition_variable idle_cv, resume_cv;
  size_t idle = 0;
  size_t epoch = 0;
  bool shutdown_flag = false;
};
