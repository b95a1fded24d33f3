// SKGPUSketchAlg (src/sk_gpu_sketch_alg.cpp:7-257 of the reference).
#include "gpusketch/sk_gpu_sketch_alg.h"

#include <cstring>
#include <iostream>
#include <stdexcept>

SKGPUSketchAlg::SKGPUSketchAlg(bool using_gpu, node_id_t n, size_t num_updates, int num_threads,
                               SketchParams p, CCAlgConfiguration cfg)
    : CCGPUSketchAlg(n, num_updates, num_threads, p, cfg),
      num_updates(num_updates),
      capacity(2 * num_updates),
      num_host_threads(num_threads) {
  if (!using_gpu)
    throw std::invalid_argument("SKGPUSketchAlg: the host update path is not provided; "
                                "construct with using_gpu = true");
  void* h = nullptr;
  gs_throw_if(gs_host_alloc(std::max<size_t>(1, capacity) * sizeof(node_id_t), &h));
  h_edgeUpdates = static_cast<node_id_t*>(h);
}

SKGPUSketchAlg::~SKGPUSketchAlg() {
  if (h_edgeUpdates) gs_host_free(h_edgeUpdates);
}

void SKGPUSketchAlg::apply_update_batch(int, node_id_t src, const std::vector<node_id_t>& dsts) {
  if (update_locked) throw UpdateLockedException();
  if (dsts.empty()) return;
  const uint64_t off = edgeUpdate_offset.fetch_add(dsts.size());
  if (off + dsts.size() > capacity)
    throw std::length_error("SKGPUSketchAlg: more updates than the num_updates given");
  std::memcpy(h_edgeUpdates + off, dsts.data(), dsts.size() * sizeof(node_id_t));
  std::lock_guard<std::mutex> lk(batch_mutex);
  batch_src.push_back(src);
  batch_size.push_back(dsts.size());
  batch_start.push_back(off);
  batch_count++;
}

void SKGPUSketchAlg::launch_cpu_update() {
  throw std::invalid_argument("SKGPUSketchAlg: the host update path is not provided");
}

void SKGPUSketchAlg::launch_gpu_kernel() {
  std::lock_guard<std::mutex> lk(batch_mutex);
  const size_t nb = batch_src.size() - applied;
  if (nb == 0) return;
  auto t0 = std::chrono::steady_clock::now();
  gs_stream* s = nullptr;
  gs_throw_if(gs_stream_create(store->handle(), &s));
  int rc = gs_submit_vertex_batches(s, batch_src.data() + applied, batch_size.data() + applied,
                                    batch_start.data() + applied, (uint32_t)nb, h_edgeUpdates);
  if (rc == GS_OK) rc = gs_flush(s);
  gs_stream_destroy(s);
  gs_throw_if(rc);
  applied = batch_src.size();
  kernel_time += std::chrono::steady_clock::now() - t0;
}

void SKGPUSketchAlg::launch_gpu_update() {
  std::lock_guard<std::mutex> lk(batch_mutex);
  auto t0 = std::chrono::steady_clock::now();
  const size_t ns = store->num_streams();
  size_t turn = 0;
  while (applied < batch_src.size()) {
    const size_t nb = std::min<size_t>(num_batch_per_buffer, batch_src.size() - applied);
    for (size_t b = applied; b < applied + nb; b++)
      store->add_batch((int)(turn % ns), batch_src[b], h_edgeUpdates + batch_start[b],
                       batch_size[b]);
    store->submit((int)(turn % ns));
    applied += nb;
    turn++;
  }
  store->flush();
  kernel_time += std::chrono::steady_clock::now() - t0;
}

void SKGPUSketchAlg::buffer_transfer() {
  const size_t bytes = edgeUpdate_offset.load() * sizeof(node_id_t);
  if (bytes == 0) return;
  void* d = nullptr;
  gs_throw_if(gs_device_alloc(bytes, &d));
  auto t0 = std::chrono::steady_clock::now();
  int rc = gs_memcpy_h2d(d, h_edgeUpdates, bytes);
  std::chrono::duration<double> t = std::chrono::steady_clock::now() - t0;
  gs_device_free(d);
  gs_throw_if(rc);
  std::cout << "Total Bytes Transferred between CPU-GPU (GB):      " << bytes / 1e9 << std::endl;
  std::cout << "Data Transfer Time (s):                            " << t.count() << std::endl;
  std::cout << "Data Transfer Time (GB/s):                         " << bytes / 1e9 / t.count()
            << std::endl;
}

void SKGPUSketchAlg::flush_buffers() {
  launch_gpu_update();
  store->flush();
}

void SKGPUSketchAlg::display_time() {
  std::cout << "Batches: " << batch_count.load() << ", buffered updates: "
            << edgeUpdate_offset.load() << std::endl;
  std::cout << "Device update wall time (s): " << kernel_time.count() << std::endl;
  CCGPUSketchAlg::display_time();
}
