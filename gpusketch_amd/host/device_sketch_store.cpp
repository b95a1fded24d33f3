// DeviceSketchStore over the C ABI (see device_sketch_store.h).
#include "gpusketch/device_sketch_store.h"

#include <algorithm>

#include "gpusketch/sketch.h"

void gs_throw_if(int status) {
  if (status == GS_OK) return;
  const std::string msg = gs_last_error();
  if (status == GS_ERR_OUT_OF_SAMPLES) throw OutOfSamplesException(msg);
  throw std::runtime_error("gpusketch: " + msg);
}

// Built with L0_SAMPLING, the sketches follow the CubeSketch rule with its columns per sample,
// as the reference's Sketch does under that definition (src/sketch.cpp:66-83,
// include/sketch.h:208-211).
#ifdef L0_SAMPLING
static constexpr uint32_t kStoreFlags = GS_STORE_CUBESKETCH;
#else
static constexpr uint32_t kStoreFlags = 0;
#endif

DeviceSketchStore::DeviceSketchStore(node_id_t n, uint64_t seed, size_t num_samples,
                                     size_t nstreams, size_t flush)
    : flush_updates(flush ? flush : 1) {
  gs_throw_if(gs_store_create(n, seed, (uint32_t)num_samples,
                              (uint32_t)Sketch::default_cols_per_sample, kStoreFlags, &store));
  gs_throw_if(gs_store_geometry(store, &geo));
  ensure_streams(std::max<size_t>(1, nstreams));
}

DeviceSketchStore::~DeviceSketchStore() {
  for (gs_stream* s : streams) gs_stream_destroy(s);
  if (store) gs_store_destroy(store);
}

void DeviceSketchStore::ensure_streams(size_t n) {
  while (streams.size() < n) {
    gs_stream* s = nullptr;
    gs_throw_if(gs_stream_create(store, &s));
    streams.push_back(s);
    buffers.emplace_back(new Buffer());
  }
}

void DeviceSketchStore::add_batch(int thr, node_id_t src, const node_id_t* dsts, size_t count) {
  if (count == 0) return;
  Buffer& b = *buffers.at((size_t)thr);
  b.src.push_back(src);
  b.sizes.push_back(count);
  b.offsets.push_back(b.dsts.size());
  b.dsts.insert(b.dsts.end(), dsts, dsts + count);
  if (b.dsts.size() >= flush_updates) submit(thr);
}

void DeviceSketchStore::submit(int thr) {
  Buffer& b = *buffers.at((size_t)thr);
  if (b.src.empty()) return;
  gs_throw_if(gs_submit_vertex_batches(streams[(size_t)thr], b.src.data(), b.sizes.data(),
                                       b.offsets.data(), (uint32_t)b.src.size(), b.dsts.data()));
  b.submitted += b.dsts.size();
  b.src.clear();
  b.sizes.clear();
  b.offsets.clear();
  b.dsts.clear();
}

void DeviceSketchStore::flush() {
  for (size_t t = 0; t < streams.size(); t++) submit((int)t);
  for (gs_stream* s : streams) gs_throw_if(gs_flush(s));
}

void DeviceSketchStore::toggle_pairs(const uint32_t* pairs, uint64_t nedges) {
  flush();
  if (nedges) gs_throw_if(gs_submit_edge_pairs(streams[0], pairs, nedges));
  gs_throw_if(gs_flush(streams[0]));
}

void DeviceSketchStore::toggle_edges(const std::vector<Edge>& edges) {
  std::vector<uint32_t> pairs(2 * edges.size());
  for (size_t i = 0; i < edges.size(); i++) {
    pairs[2 * i] = edges[i].src;
    pairs[2 * i + 1] = edges[i].dst;
  }
  toggle_pairs(pairs.data(), edges.size());
}

std::vector<node_id_t> DeviceSketchStore::connected_components(std::vector<Edge>* forest,
                                                               size_t* rounds) {
  flush();
  const uint32_t n = geo.num_vertices;
  std::vector<node_id_t> labels(n);
  std::vector<uint32_t> f(forest ? 2 * (size_t)n : 0);
  uint32_t ncc = 0, nr = 0;
  uint64_t flen = 0;
  gs_throw_if(gs_query_cc(store, labels.data(), &ncc, &nr, forest ? f.data() : nullptr,
                          forest ? n : 0, forest ? &flen : nullptr));
  if (rounds) *rounds = nr;
  if (forest) {
    forest->resize(flen);
    for (uint64_t i = 0; i < flen; i++) (*forest)[i] = {f[2 * i], f[2 * i + 1]};
  }
  return labels;
}

std::vector<std::vector<Edge>> DeviceSketchStore::k_spanning_forests(size_t k) {
  flush();
  const uint64_t n = geo.num_vertices;
  std::vector<uint32_t> e(2 * k * n);
  std::vector<uint64_t> counts(k);
  gs_throw_if(gs_k_spanning_forests(store, (uint32_t)k, e.data(), k * n, counts.data()));
  std::vector<std::vector<Edge>> out(k);
  uint64_t pos = 0;
  for (size_t i = 0; i < k; i++) {
    out[i].resize(counts[i]);
    for (uint64_t j = 0; j < counts[i]; j++, pos++) out[i][j] = {e[2 * pos], e[2 * pos + 1]};
  }
  return out;
}

void DeviceSketchStore::export_buckets(node_id_t v0, node_id_t count, Bucket* out) {
  flush();
  gs_throw_if(gs_export_buckets(store, v0, count, reinterpret_cast<gs_bucket*>(out)));
}

void DeviceSketchStore::import_buckets(node_id_t v0, node_id_t count, const Bucket* in) {
  flush();
  gs_throw_if(gs_import_buckets(store, v0, count, reinterpret_cast<const gs_bucket*>(in)));
}

void DeviceSketchStore::zero() {
  flush();
  gs_throw_if(gs_store_zero(store));
}

void DeviceSketchStore::set_profiling(bool on) {
  for (gs_stream* s : streams) gs_throw_if(gs_stream_set_profiling(s, on ? 1 : 0));
}

void DeviceSketchStore::timing(double* ms, uint64_t* launches, uint64_t* updates) {
  double tm = 0;
  uint64_t tl = 0, tu = 0;
  for (gs_stream* s : streams) {
    double m = 0;
    uint64_t l = 0, u = 0;
    gs_throw_if(gs_stream_timing(s, &m, &l, &u));
    tm += m;
    tl += l;
    tu += u;
  }
  if (ms) *ms = tm;
  if (launches) *launches = tl;
  if (updates) *updates = tu;
}

uint64_t DeviceSketchStore::submitted_updates() const {
  uint64_t t = 0;
  for (auto& b : buffers) t += b->submitted;
  return t;
}
