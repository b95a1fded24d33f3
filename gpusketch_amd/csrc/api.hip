// C-ABI entry points of libgpusketch: store and stream lifecycle, submissions, export/import,
// timing and device-memory helpers. Kernels live in ingest.hip and boruvka.hip.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>

#include "gs_internal.h"

namespace gs {

static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void geometry(uint32_t n, double factor, uint32_t cps, int l0, uint64_t out[5]) {
  // include/sketch.h:63-77 (vector length, samples) and :198-206 (buckets per column)
  uint64_t len = (uint64_t)std::ceil((double)n * (uint32_t)(n - 1u) / 2);
  uint64_t bpc = (uint64_t)(std::ceil(std::log2((double)len)) + 1);
  if (bpc % 16 == 0) bpc += 1;
  double div = l0 ? (std::log2(3.0) - 1) : (1 - 0.263034406);
  uint64_t s = (uint64_t)std::ceil(factor * std::log2((double)n) / div);
  if (s < 18) s = 18;
  out[0] = len;
  out[1] = bpc;
  out[2] = s;
  out[3] = s * cps;
  out[4] = s * cps * bpc + 1;
}

// ---- staging helpers -------------------------------------------------------------------------
int ensure_stage(Stream* s, size_t bytes, bool device) {
  if (s->stage_pending) {
    GS_HIP(hipEventSynchronize(s->stage_free));
    s->stage_pending = false;
  }
  if (s->h_stage_bytes < bytes) {
    if (s->h_stage) hipHostFree(s->h_stage);
    s->h_stage = nullptr;
    size_t cap = std::max(bytes, (size_t)1 << 20);
    if (hipHostMalloc((void**)&s->h_stage, cap, hipHostMallocDefault) != hipSuccess)
      return fail(GS_ERR_OOM, "pinned staging of %zu bytes failed", cap);
    s->h_stage_bytes = cap;
  }
  if (device && s->d_stage_bytes < bytes) {
    GS_HIP(hipStreamSynchronize(s->stream));
    if (s->d_stage) hipFree(s->d_stage);
    s->d_stage = nullptr;
    size_t cap = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc((void**)&s->d_stage, cap) != hipSuccess)
      return fail(GS_ERR_OOM, "device staging of %zu bytes failed", cap);
    s->d_stage_bytes = cap;
  }
  return GS_OK;
}

int init_stream_objects(Stream* s) {
  if (!s->stream) {
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
      return fail(GS_ERR_HIP, "HIP stream creation failed");
    s->own_stream = true;
  }
  if (hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) != hipSuccess)
    return fail(GS_ERR_HIP, "HIP copy stream creation failed");
  if (hipStreamCreateWithFlags(&s->gstream, hipStreamNonBlocking) != hipSuccess)
    return fail(GS_ERR_HIP, "HIP grouping stream creation failed");
  hipEvent_t* evs[] = {&s->stage_free,    &s->pend_free[0],  &s->pend_free[1], &s->copied,
                       &s->grouped[0],    &s->grouped[1],    &s->aggregated[0],
                       &s->aggregated[1]};
  for (hipEvent_t* e : evs)
    if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess)
      return fail(GS_ERR_HIP, "HIP event creation failed");
  return GS_OK;
}

void destroy_stream_resources(Stream& s) {
  if (s.stream) (void)hipStreamSynchronize(s.stream);
  if (s.gstream) (void)hipStreamSynchronize(s.gstream);
  if (s.copy_stream) (void)hipStreamSynchronize(s.copy_stream);
  free_pair_scratch(s.ws);
  free_pair_scratch(s.ws_alt);
  for (int i = 0; i < 2; i++) {
    if (s.grouped[i]) (void)hipEventDestroy(s.grouped[i]);
    if (s.aggregated[i]) (void)hipEventDestroy(s.aggregated[i]);
    s.grouped[i] = s.aggregated[i] = nullptr;
    s.agg_used[i] = false;
  }
  if (s.h_stage) (void)hipHostFree(s.h_stage);
  if (s.d_stage) (void)hipFree(s.d_stage);
  for (int b = 0; b < 2; b++) {
    if (s.d_pend[b]) (void)hipFree(s.d_pend[b]);
    if (s.pend_free[b]) (void)hipEventDestroy(s.pend_free[b]);
    s.d_pend[b] = nullptr;
    s.pend_free[b] = nullptr;
  }
  if (s.copied) (void)hipEventDestroy(s.copied);
  if (s.d_traffic) (void)hipFree(s.d_traffic);
  s.d_traffic = nullptr;
  for (auto& p : s.ev_pairs) {
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  for (auto& g : s.ev_group) {
    if (g.first) (void)hipEventDestroy(g.first);
    if (g.second) (void)hipEventDestroy(g.second);
  }
  s.ev_group.clear();
  if (s.stage_free) (void)hipEventDestroy(s.stage_free);
  if (s.copy_stream) (void)hipStreamDestroy(s.copy_stream);
  if (s.gstream) (void)hipStreamDestroy(s.gstream);
  s.gstream = nullptr;
  if (s.stream && s.own_stream) (void)hipStreamDestroy(s.stream);
  s.h_stage = nullptr;
  s.d_stage = nullptr;
  s.ev_pairs.clear();
  s.stage_free = nullptr;
  s.copied = nullptr;
  s.copy_stream = nullptr;
  s.stream = nullptr;
  s.pend_cap = s.pend_n = s.fill_n = 0;
  s.zsegs.clear();
}

int sync_store(Store* st) {
  std::vector<Stream*> ss;
  {
    std::lock_guard<std::mutex> lk(st->mu);
    ss = st->streams;
  }
  for (Stream* s : ss) {
    int rc = drain_stream(s);
    if (rc) return rc;
  }
  for (Stream* s : ss) GS_HIP(hipStreamSynchronize(s->stream));
  GS_HIP(hipStreamSynchronize(st->qstream));
  return GS_OK;
}

int release_stage(Stream* s) {
  GS_HIP(hipEventRecord(s->stage_free, s->stream));
  s->stage_pending = true;
  return GS_OK;
}

}  // namespace gs

using namespace gs;

extern "C" {

const char* gs_last_error(void) { return g_last_error.c_str(); }

#ifndef GS_SOURCE_DIGEST
#define GS_SOURCE_DIGEST "unknown"
#endif

int gs_source_digest(char* out, size_t cap) {
  static const char digest[] = GS_SOURCE_DIGEST;
  GS_CHECK(out && cap >= sizeof(digest), GS_ERR_INVALID, "gs_source_digest: buffer too small");
  std::memcpy(out, digest, sizeof(digest));
  return GS_OK;
}

int gs_geometry_for(uint32_t n, double sketches_factor, uint32_t cols_per_sample,
                    gs_geometry* out) {
  return gs_geometry_for_sketch(n, sketches_factor, cols_per_sample, 0, out);
}

int gs_geometry_for_sketch(uint32_t n, double sketches_factor, uint32_t cols_per_sample,
                           uint32_t flags, gs_geometry* out) {
  GS_CHECK(out && n >= 2 && cols_per_sample >= 1 && sketches_factor > 0, GS_ERR_INVALID,
           "gs_geometry_for: bad argument");
  GS_CHECK((flags & ~GS_STORE_CUBESKETCH) == 0, GS_ERR_INVALID, "gs_geometry_for: unknown flags");
  uint64_t g[5];
  geometry(n, sketches_factor, cols_per_sample, (flags & GS_STORE_CUBESKETCH) ? 1 : 0, g);
  std::memset(out, 0, sizeof(*out));
  out->flags = flags;
  out->num_vertices = n;
  out->vector_len = g[0];
  out->bkt_per_col = (uint32_t)g[1];
  out->num_samples = (uint32_t)g[2];
  out->cols_per_sample = cols_per_sample;
  out->num_columns = (uint32_t)g[3];
  out->num_buckets = (uint32_t)g[4];
  uint64_t vstride = ((uint64_t)g[1] * g[3] + 7) & ~7ULL;
  out->device_bytes = (uint64_t)n * (vstride + 1) * 16;
  return GS_OK;
}

int gs_hash_known_answer(const uint64_t* keys, const uint64_t* seeds, uint64_t count,
                         uint64_t* out) {
  GS_CHECK((keys && seeds && out) || count == 0, GS_ERR_INVALID, "null argument");
  return hash_known_answer(keys, seeds, count, out);
}

int gs_store_create(uint32_t n, uint64_t seed, uint32_t num_samples, uint32_t cols_per_sample,
                    uint32_t flags, gs_store** out) {
  GS_CHECK(out, GS_ERR_INVALID, "gs_store_create: null out");
  *out = nullptr;
  GS_CHECK(n >= 2, GS_ERR_INVALID, "gs_store_create: need at least 2 vertices");
  GS_CHECK(num_samples >= 1 && cols_per_sample >= 1, GS_ERR_INVALID, "gs_store_create: bad samples");
  GS_CHECK((flags & ~GS_STORE_CUBESKETCH) == 0, GS_ERR_INVALID, "gs_store_create: unknown flags");
  uint64_t g[5];
  geometry(n, 1.0, 1, 0, g);
  const uint32_t bpc = (uint32_t)g[1];
  const uint64_t cols = (uint64_t)num_samples * cols_per_sample;
  GS_CHECK(bpc < 64 && cols <= 4096, GS_ERR_INVALID, "gs_store_create: geometry out of range");
  GS_CHECK(n <= (1u << 26), GS_ERR_UNSUPPORTED, "gs_store_create: at most 2^26 vertices per store");
  gs_store* h = new gs_store();
  Store& st = h->impl;
  st.num_samples = num_samples;
  st.cols_per_sample = cols_per_sample;
  st.flags = flags;
  hipGetDevice(&st.device);
  st.view.n = n;
  st.view.bpc = bpc;
  st.view.cols = (uint32_t)cols;
  st.view.vstride = (uint32_t)(((uint64_t)bpc * cols + 7) & ~7ULL);
  st.view.seed = seed;
  st.part = make_part_params(n);
  size_t rows_bytes = (size_t)n * st.view.vstride * sizeof(Bucket16);
  if (hipMalloc((void**)&st.view.rows, rows_bytes) != hipSuccess ||
      hipMalloc((void**)&st.view.det, (size_t)n * sizeof(Bucket16)) != hipSuccess) {
    hipFree(st.view.rows);
    hipFree(st.view.det);
    delete h;
    return fail(GS_ERR_OOM, "gs_store_create: %zu bytes of HBM unavailable", rows_bytes);
  }
  if (hipStreamCreateWithFlags(&st.qstream, hipStreamNonBlocking) != hipSuccess) {
    hipFree(st.view.rows);
    hipFree(st.view.det);
    delete h;
    return fail(GS_ERR_HIP, "gs_store_create: stream creation failed");
  }
  int rc = launch_zero_store(&st, st.qstream);
  if (rc == GS_OK && hipStreamSynchronize(st.qstream) != hipSuccess)
    rc = fail(GS_ERR_HIP, "gs_store_create: zeroing failed");
  if (rc) {
    hipStreamDestroy(st.qstream);
    hipFree(st.view.rows);
    hipFree(st.view.det);
    delete h;
    return rc;
  }
  *out = h;
  return GS_OK;
}

int gs_store_destroy(gs_store* h) {
  if (!h) return GS_OK;
  Store& st = h->impl;
  std::vector<Stream*> ss;
  {
    std::lock_guard<std::mutex> lk(st.mu);
    ss = st.streams;
  }
  // a group holds the store's rows and writes them from its own stream: it goes first
  for (Stream* s : ss)
    GS_CHECK(!s->group, GS_ERR_INVALID,
             "gs_store_destroy: the store is a member of a group; destroy the group first");
  hipDeviceSynchronize();
  for (Stream* s : ss) gs_stream_destroy(reinterpret_cast<gs_stream*>(s));
  free_query_scratch(&st);
  if (st.qstream) hipStreamDestroy(st.qstream);
  hipFree(st.view.rows);
  hipFree(st.view.det);
  delete h;
  return GS_OK;
}

int gs_store_geometry(const gs_store* h, gs_geometry* out) {
  GS_CHECK(h && out, GS_ERR_INVALID, "null argument");
  const Store& st = h->impl;
  uint64_t g[5];
  geometry(st.view.n, 1.0, 1, 0, g);
  std::memset(out, 0, sizeof(*out));
  out->num_vertices = st.view.n;
  out->bkt_per_col = st.view.bpc;
  out->num_samples = st.num_samples;
  out->cols_per_sample = st.cols_per_sample;
  out->num_columns = st.view.cols;
  out->num_buckets = st.view.cols * st.view.bpc + 1;
  out->seed = st.view.seed;
  out->vector_len = g[0];
  out->device_bytes = st.bytes();
  out->flags = st.flags;
  return GS_OK;
}

int gs_store_zero(gs_store* h) {
  GS_CHECK(h, GS_ERR_INVALID, "null store");
  int rc = gs_sync(h);
  if (rc) return rc;
  rc = launch_zero_store(&h->impl, h->impl.qstream);
  if (rc) return rc;
  GS_HIP(hipStreamSynchronize(h->impl.qstream));
  h->impl.symmetric = true;
  return GS_OK;
}

int gs_stream_create(gs_store* h, gs_stream** out) {
  GS_CHECK(h && out, GS_ERR_INVALID, "null argument");
  gs_stream* s = new gs_stream();
  Stream& st = s->impl;
  st.store = &h->impl;
  int rc = init_stream_objects(&st);
  if (rc) {
    destroy_stream_resources(st);
    delete s;
    return rc;
  }
  {
    std::lock_guard<std::mutex> lk(h->impl.mu);
    h->impl.streams.push_back(&st);
  }
  *out = s;
  return GS_OK;
}

int gs_stream_destroy(gs_stream* sh) {
  if (!sh) return GS_OK;
  Stream& s = sh->impl;
  GS_CHECK(!s.group, GS_ERR_INVALID, "a group's stream is destroyed with the group");
  (void)drain_stream(&s);  // pending updates are applied before the stream goes away
  hipStreamSynchronize(s.stream);
  {
    std::lock_guard<std::mutex> lk(s.store->mu);
    auto& v = s.store->streams;
    v.erase(std::remove(v.begin(), v.end(), &s), v.end());
  }
  destroy_stream_resources(s);
  delete sh;
  return GS_OK;
}

static int submit_vertex_batches(gs_stream* sh, const uint32_t* src, const uint64_t* sizes,
                                 const uint64_t* offsets, uint32_t nbatches, const uint32_t* dsts,
                                 uint32_t min_depth, uint32_t depth_cap) {
  GS_CHECK(sh, GS_ERR_INVALID, "null stream");
  GS_CHECK(depth_cap < 64 && min_depth <= depth_cap, GS_ERR_INVALID,
           "subgraph filter: need min_depth <= depth_cap < 64");
  if (nbatches == 0) return GS_OK;
  GS_CHECK(src && sizes && offsets && dsts, GS_ERR_INVALID, "null batch array");
  Stream* s = &sh->impl;
  const uint32_t n = s->store->view.n;
  const uint32_t chunk = 256;
  // work items (vertex, len, start) over a compacted destination array
  uint64_t total = 0, nitems = 0;
  for (uint32_t b = 0; b < nbatches; b++) {
    GS_CHECK(src[b] < n, GS_ERR_INVALID, "batch %u: source vertex %u out of range", b, src[b]);
    total += sizes[b];
    nitems += (sizes[b] + chunk - 1) / chunk;
  }
  if (total == 0) return GS_OK;
  s->store->symmetric = false;  // directed updates: the other directions may never come
  // staging: per-class item counts (all items in class 0), the items, the destinations
  const size_t head = 64;
  size_t item_bytes = nitems * sizeof(WorkItem);
  size_t bytes = head + item_bytes + total * sizeof(uint32_t);
  int rc = ensure_stage(s, bytes);
  if (rc) return rc;
  GS_CHECK(total < (1ULL << 32), GS_ERR_INVALID, "too many destinations in one submission");
  std::memset(s->h_stage, 0, head);
  s->h_stage[0] = (uint32_t)nitems;
  WorkItem* hi = reinterpret_cast<WorkItem*>(reinterpret_cast<char*>(s->h_stage) + head);
  uint32_t* hd =
      reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(s->h_stage) + head + item_bytes);
  uint64_t pos = 0, k = 0;
  for (uint32_t b = 0; b < nbatches; b++) {
    for (uint64_t i = 0; i < sizes[b]; i++) {
      uint32_t d = dsts[offsets[b] + i];
      GS_CHECK(d < n, GS_ERR_INVALID, "batch %u: destination %u out of range", b, d);
      hd[pos + i] = d;
    }
    for (uint64_t q = 0; q < sizes[b]; q += chunk) {
      hi[k].vertex = src[b];
      hi[k].len = (uint32_t)std::min<uint64_t>(chunk, sizes[b] - q);  // shared: atomic flush
      hi[k].start = (uint32_t)(pos + q);
      hi[k].graph = 0;
      k++;
    }
    pos += sizes[b];
  }
  GS_HIP(hipMemcpyAsync(s->d_stage, s->h_stage, bytes, hipMemcpyHostToDevice, s->stream));
  const char* base = reinterpret_cast<const char*>(s->d_stage);
  const WorkItem* di = reinterpret_cast<const WorkItem*>(base + head);
  const uint32_t* dd = reinterpret_cast<const uint32_t*>(base + head + item_bytes);
  rc = launch_vertex_batches(s, reinterpret_cast<const uint32_t*>(base), di, (uint32_t)nitems, dd,
                             total, min_depth, depth_cap);
  if (rc) return rc;
  return release_stage(s);
}

int gs_submit_vertex_batches(gs_stream* sh, const uint32_t* src, const uint64_t* sizes,
                             const uint64_t* offsets, uint32_t nbatches, const uint32_t* dsts) {
  return submit_vertex_batches(sh, src, sizes, offsets, nbatches, dsts, 0, 0);
}

int gs_submit_vertex_batches_filtered(gs_stream* sh, const uint32_t* src, const uint64_t* sizes,
                                      const uint64_t* offsets, uint32_t nbatches,
                                      const uint32_t* dsts, uint32_t min_depth,
                                      uint32_t depth_cap) {
  return submit_vertex_batches(sh, src, sizes, offsets, nbatches, dsts, min_depth, depth_cap);
}

int gs_submit_edge_pairs(gs_stream* sh, const uint32_t* pairs, uint64_t nedges) {
  GS_CHECK(sh && (pairs || nedges == 0), GS_ERR_INVALID, "null argument");
  return submit_pairs_host(&sh->impl, pairs, nedges, 0, 0);
}

int gs_submit_edge_pairs_filtered(gs_stream* sh, const uint32_t* pairs, uint64_t nedges,
                                  uint32_t min_depth, uint32_t depth_cap) {
  GS_CHECK(sh && (pairs || nedges == 0), GS_ERR_INVALID, "null argument");
  GS_CHECK(depth_cap < 64 && min_depth <= depth_cap, GS_ERR_INVALID, "bad subgraph filter");
  return submit_pairs_host(&sh->impl, pairs, nedges, min_depth, depth_cap);
}

int gs_submit_edge_pairs_device(gs_stream* sh, const uint32_t* d_pairs, uint64_t nedges) {
  GS_CHECK(sh && (d_pairs || nedges == 0), GS_ERR_INVALID, "null argument");
  return submit_pairs_device(&sh->impl, d_pairs, nedges, 0, 0);
}

int gs_flush(gs_stream* sh) {
  GS_CHECK(sh, GS_ERR_INVALID, "null stream");
  Stream& s = sh->impl;
  int rc = drain_stream(&s);
  if (rc) return rc;
  GS_HIP(hipStreamSynchronize(s.stream));
  uint32_t dropped = 0;
  for (PairScratch* ws : {&s.ws, &s.ws_alt}) {  // (both scratch sets of the pipeline)
    if (!ws->d_counters) continue;
    uint32_t d = 0;
    GS_HIP(hipMemcpy(&d, ws->d_counters + 3, sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (d) GS_HIP(hipMemset(ws->d_counters + 3, 0, sizeof(uint32_t)));
    dropped += d;
  }
  if (dropped)
    return fail(GS_ERR_INVALID,
                "%u submitted updates named a vertex outside the store (or joined two members of "
                "a group) and were dropped", dropped);
  return GS_OK;
}

int gs_flush_async(gs_stream* sh) {
  GS_CHECK(sh, GS_ERR_INVALID, "null stream");
  return drain_stream(&sh->impl);
}

int gs_sync(gs_store* h) {
  GS_CHECK(h, GS_ERR_INVALID, "null store");
  return sync_store(&h->impl);
}

// ---- reference-order transfers ----------------------------------------------------------------
// A vertex's buckets in reference order are [c*B + d] for column c, depth d, then the deterministic
// bucket; the device keeps them as bucket_slot orders them (row stride vstride) with the deterministic
// buckets in their own array. One thread per reference-order bucket of a chunk of vertices: the
// reference-order side is read or written coalesced, the device side gathered within a row.
__global__ void to_reference_order_kernel(const Bucket16* __restrict__ rows,
                                          const Bucket16* __restrict__ det, uint32_t vstride,
                                          uint32_t B, uint32_t C, uint32_t v0, uint64_t total,
                                          Bucket16* __restrict__ out) {
  const uint32_t nb = B * C + 1;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)(i / nb), j = (uint32_t)(i % nb);
    const uint32_t c = j / B, d = j % B;
    out[i] = j == nb - 1 ? det[v0 + v] : rows[(uint64_t)(v0 + v) * vstride + bucket_slot(d, c, C, B)];
  }
}

__global__ void from_reference_order_kernel(const Bucket16* __restrict__ in, uint32_t vstride,
                                            uint32_t B, uint32_t C, uint32_t v0, uint64_t total,
                                            Bucket16* __restrict__ rows, Bucket16* __restrict__ det) {
  const uint32_t nb = B * C + 1;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)(i / nb), j = (uint32_t)(i % nb);
    const uint32_t c = j / B, d = j % B;
    if (j == nb - 1) det[v0 + v] = in[i];
    else rows[(uint64_t)(v0 + v) * vstride + bucket_slot(d, c, C, B)] = in[i];
  }
}

// XOR of bucket slot j over the chunk's vertices into acc[j] (slots across threads: coalesced)
__global__ void slot_xor_kernel(const Bucket16* __restrict__ in, uint32_t nb, uint32_t count,
                                uint32_t per_block, unsigned long long* acc) {
  const uint32_t v_lo = blockIdx.y * per_block;
  const uint32_t v_hi = min(count, v_lo + per_block);
  for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nb; j += gridDim.x * blockDim.x) {
    uint64_t a = 0, g = 0;
    for (uint32_t v = v_lo; v < v_hi; v++) {
      a ^= in[(uint64_t)v * nb + j].alpha;
      g ^= in[(uint64_t)v * nb + j].gamma;
    }
    if (a) atomicXor(&acc[2 * j], (unsigned long long)a);
    if (g) atomicXor(&acc[2 * j + 1], (unsigned long long)g);
  }
}

// vertices per transfer chunk: about 256 MB of reference-order buckets
static uint32_t transfer_chunk(uint32_t nb) {
  return std::max<uint32_t>(1, (uint32_t)((256ull << 20) / ((uint64_t)nb * sizeof(Bucket16))));
}

int gs_export_buckets(gs_store* h, uint32_t v0, uint32_t count, gs_bucket* out) {
  GS_CHECK(h && (out || count == 0), GS_ERR_INVALID, "null argument");
  Store& st = h->impl;
  GS_CHECK((uint64_t)v0 + count <= st.view.n, GS_ERR_INVALID, "vertex range out of bounds");
  int rc = gs_sync(h);
  if (rc || count == 0) return rc;
  const uint32_t B = st.view.bpc, C = st.view.cols, vs = st.view.vstride;
  const uint32_t nb = B * C + 1;
  const uint32_t per = std::min(count, transfer_chunk(nb));
  Bucket16* tmp = nullptr;
  if (hipMalloc((void**)&tmp, (size_t)per * nb * sizeof(Bucket16)) != hipSuccess)
    return fail(GS_ERR_OOM, "export staging of %u vertices unavailable", per);
  hipError_t e = hipSuccess;
  for (uint32_t a = 0; a < count && e == hipSuccess; a += per) {
    const uint32_t m = std::min(per, count - a);
    const uint64_t total = (uint64_t)m * nb;
    hipLaunchKernelGGL(to_reference_order_kernel, dim3((uint32_t)std::min<uint64_t>((total + 255) / 256, 8192)),
                       dim3(256), 0, nullptr, st.view.rows, st.view.det, vs, B, C, v0 + a, total, tmp);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpy(out + (size_t)a * nb, tmp, total * sizeof(Bucket16), hipMemcpyDeviceToHost);
  }
  (void)hipFree(tmp);
  GS_HIP(e);
  return GS_OK;
}

int gs_import_buckets(gs_store* h, uint32_t v0, uint32_t count, const gs_bucket* in) {
  GS_CHECK(h && (in || count == 0), GS_ERR_INVALID, "null argument");
  Store& st = h->impl;
  GS_CHECK((uint64_t)v0 + count <= st.view.n, GS_ERR_INVALID, "vertex range out of bounds");
  int rc = gs_sync(h);
  if (rc) return rc;
  // a whole store is symmetric (its vertex sketches XOR to zero) if the imported ones do; a part
  // of a store cannot be judged from the part alone
  const bool whole = v0 == 0 && count == st.view.n;
  st.symmetric = false;
  if (count == 0) return GS_OK;
  const uint32_t B = st.view.bpc, C = st.view.cols, vs = st.view.vstride;
  const uint32_t nb = B * C + 1;
  const uint32_t per = std::min(count, transfer_chunk(nb));
  Bucket16* tmp = nullptr;
  unsigned long long* acc = nullptr;
  if (hipMalloc((void**)&tmp, (size_t)per * nb * sizeof(Bucket16)) != hipSuccess ||
      hipMalloc((void**)&acc, 2 * (size_t)nb * sizeof(unsigned long long)) != hipSuccess) {
    (void)hipFree(tmp);
    return fail(GS_ERR_OOM, "import staging of %u vertices unavailable", per);
  }
  hipError_t e = hipMemset(acc, 0, 2 * (size_t)nb * sizeof(unsigned long long));
  for (uint32_t a = 0; a < count && e == hipSuccess; a += per) {
    const uint32_t m = std::min(per, count - a);
    const uint64_t total = (uint64_t)m * nb;
    e = hipMemcpy(tmp, in + (size_t)a * nb, total * sizeof(Bucket16), hipMemcpyHostToDevice);
    if (e != hipSuccess) break;
    hipLaunchKernelGGL(from_reference_order_kernel, dim3((uint32_t)std::min<uint64_t>((total + 255) / 256, 8192)),
                       dim3(256), 0, nullptr, tmp, vs, B, C, v0 + a, total, st.view.rows, st.view.det);
    if (whole) {
      const uint32_t pb = 256;  // vertices per block row
      hipLaunchKernelGGL(slot_xor_kernel, dim3((nb + 255) / 256, (m + pb - 1) / pb), dim3(256), 0,
                         nullptr, tmp, nb, m, pb, acc);
    }
    e = hipGetLastError();  // the next chunk's copy into tmp waits for these (same stream)
  }
  if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
  std::vector<unsigned long long> sum(2 * (size_t)nb, 0);
  if (e == hipSuccess && whole)
    e = hipMemcpy(sum.data(), acc, sum.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
  (void)hipFree(tmp);
  (void)hipFree(acc);
  GS_HIP(e);
  bool zero_sum = whole;
  for (unsigned long long x : sum) zero_sum = zero_sum && x == 0;
  st.symmetric = zero_sum;
  return GS_OK;
}

// ---- device-order snapshots --------------------------------------------------------------------
// A 64-byte header (magic, geometry, seed, flags, whether the vertex sketches XOR to zero), then
// the rows and the deterministic buckets exactly as the device holds them: one DMA each way.
namespace {
constexpr uint64_t kSnapMagic = 0x31504e5348534753ULL;  // "GSHSNP1"
struct SnapHeader {
  uint64_t magic;
  uint64_t seed;
  uint32_t n, bpc, cols, vstride;
  uint32_t flags, symmetric, num_samples, cols_per_sample;
  uint64_t pad[2];
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");
uint64_t snapshot_bytes(const Store& st) {
  return sizeof(SnapHeader) + ((uint64_t)st.view.n * st.view.vstride + st.view.n) * sizeof(Bucket16);
}
}  // namespace

int gs_store_snapshot_bytes(gs_store* h, uint64_t* bytes) {
  GS_CHECK(h && bytes, GS_ERR_INVALID, "null argument");
  *bytes = snapshot_bytes(h->impl);
  return GS_OK;
}

int gs_store_snapshot(gs_store* h, void* dst) {
  GS_CHECK(h && dst, GS_ERR_INVALID, "null argument");
  Store& st = h->impl;
  int rc = gs_sync(h);
  if (rc) return rc;
  SnapHeader hd{};
  hd.magic = kSnapMagic;
  hd.seed = st.view.seed;
  hd.n = st.view.n;
  hd.bpc = st.view.bpc;
  hd.cols = st.view.cols;
  hd.vstride = st.view.vstride;
  hd.flags = st.flags;
  hd.symmetric = st.symmetric.load() ? 1u : 0u;
  hd.num_samples = st.num_samples;
  hd.cols_per_sample = st.cols_per_sample;
  unsigned char* p = static_cast<unsigned char*>(dst);
  std::memcpy(p, &hd, sizeof(hd));
  const size_t rows = (size_t)st.view.n * st.view.vstride * sizeof(Bucket16);
  GS_HIP(hipMemcpy(p + sizeof(hd), st.view.rows, rows, hipMemcpyDeviceToHost));
  GS_HIP(hipMemcpy(p + sizeof(hd) + rows, st.view.det, (size_t)st.view.n * sizeof(Bucket16),
                   hipMemcpyDeviceToHost));
  return GS_OK;
}

int gs_store_restore(gs_store* h, const void* src) {
  GS_CHECK(h && src, GS_ERR_INVALID, "null argument");
  Store& st = h->impl;
  SnapHeader hd;
  std::memcpy(&hd, src, sizeof(hd));
  GS_CHECK(hd.magic == kSnapMagic, GS_ERR_INVALID, "not a store snapshot");
  GS_CHECK(hd.seed == st.view.seed && hd.n == st.view.n && hd.bpc == st.view.bpc &&
               hd.cols == st.view.cols && hd.vstride == st.view.vstride && hd.flags == st.flags &&
               hd.num_samples == st.num_samples && hd.cols_per_sample == st.cols_per_sample,
           GS_ERR_INVALID, "snapshot of a store of another geometry or seed");
  int rc = gs_sync(h);
  if (rc) return rc;
  const unsigned char* p = static_cast<const unsigned char*>(src);
  const size_t rows = (size_t)st.view.n * st.view.vstride * sizeof(Bucket16);
  GS_HIP(hipMemcpy(st.view.rows, p + sizeof(hd), rows, hipMemcpyHostToDevice));
  GS_HIP(hipMemcpy(st.view.det, p + sizeof(hd) + rows, (size_t)st.view.n * sizeof(Bucket16),
                   hipMemcpyHostToDevice));
  st.symmetric = hd.symmetric != 0;
  return GS_OK;
}

int gs_stream_set_tuning(gs_stream* sh, const gs_tuning* t) {
  GS_CHECK(sh && t, GS_ERR_INVALID, "null argument");
  GS_CHECK(t->lds_depths <= 16, GS_ERR_INVALID, "lds_depths must be <= 16");
  GS_CHECK(t->item_updates == 0 || (t->item_updates >= 64 && t->item_updates <= 1024),
           GS_ERR_INVALID, "item_updates must be in [64, 1024]");
  GS_CHECK(t->bucket_bits <= 13, GS_ERR_INVALID, "bucket_bits must be <= 13");
  GS_CHECK(t->launch_edges == 0 || (t->launch_edges >= (1u << 12) && t->launch_edges <= (1u << 28)),
           GS_ERR_INVALID, "launch_edges must be in [2^12, 2^28]");
  Stream& s = sh->impl;
  int rc = drain_stream(&s);  // pending edges were counted under the old parameters
  if (rc) return rc;
  GS_HIP(hipStreamSynchronize(s.stream));
  PartParams part{};
  if (t->bucket_bits) {
    part = make_part_params(stream_vertices(&s), (int)t->bucket_bits);
    GS_CHECK(part.nbits - part.local_bits <= 13, GS_ERR_INVALID,
             "bucket_bits too small for %u vertices", stream_vertices(&s));
  }
  s.tune_dl = t->lds_depths;
  s.tune_chunk = t->item_updates;
  s.tune_bucket_bits = t->bucket_bits;
  s.tune_launch = t->launch_edges;
  if (t->bucket_bits) s.tuned_part = part;
  return GS_OK;
}

int gs_stream_get_tuning(gs_stream* sh, gs_tuning* t) {
  GS_CHECK(sh && t, GS_ERR_INVALID, "null argument");
  const Stream& s = sh->impl;
  t->lds_depths = s.tune_dl;
  t->item_updates = s.tune_chunk;
  t->bucket_bits = s.tune_bucket_bits;
  t->launch_edges = s.tune_launch;
  t->kernel_ms = 0;
  return GS_OK;
}

int gs_stream_autotune(gs_stream* sh, const uint32_t* d_pairs, uint64_t nedges, gs_tuning* best) {
  GS_CHECK(sh && d_pairs, GS_ERR_INVALID, "null argument");
  return autotune_stream(&sh->impl, d_pairs, nedges, best);
}

int gs_stream_set_profiling(gs_stream* sh, int enable) {
  GS_CHECK(sh, GS_ERR_INVALID, "null stream");
  sh->impl.profile = enable != 0;
  return GS_OK;
}

int gs_stream_kernel_times(gs_stream* sh, double* grouping_ms, double* agg_ms,
                           uint64_t* launches, uint64_t* updates) {
  GS_CHECK(sh, GS_ERR_INVALID, "null stream");
  Stream& s = sh->impl;
  GS_HIP(hipStreamSynchronize(s.stream));
  for (size_t i = 0; i < s.ev_pairs.size(); i++) {
    const auto& p = s.ev_pairs[i];
    float ms = 0;
    GS_HIP(hipEventElapsedTime(&ms, p.first, p.second));
    s.agg_ms += ms;
    s.agg_launches++;
    if (s.ev_group[i].first) {
      float gms = 0;
      GS_HIP(hipEventElapsedTime(&gms, s.ev_group[i].first, s.ev_group[i].second));
      s.grp_ms += gms;
      hipEventDestroy(s.ev_group[i].first);
      hipEventDestroy(s.ev_group[i].second);
    }
    hipEventDestroy(p.first);
    hipEventDestroy(p.second);
  }
  s.ev_pairs.clear();
  s.ev_group.clear();
  if (grouping_ms) *grouping_ms = s.grp_ms;
  if (agg_ms) *agg_ms = s.agg_ms;
  if (launches) *launches = s.agg_launches;
  if (updates) *updates = s.agg_updates;
  s.grp_ms = 0;
  s.agg_ms = 0;
  s.agg_launches = 0;
  s.agg_updates = 0;
  return GS_OK;
}

int gs_stream_timing(gs_stream* sh, double* agg_ms, uint64_t* launches, uint64_t* updates) {
  return gs_stream_kernel_times(sh, nullptr, agg_ms, launches, updates);
}

int gs_stream_traffic(gs_stream* sh, uint64_t* flushed_buckets, uint64_t* direct_updates) {
  GS_CHECK(sh, GS_ERR_INVALID, "null stream");
  Stream& s = sh->impl;
  unsigned long long h[2] = {0, 0};
  if (s.d_traffic) {
    GS_HIP(hipStreamSynchronize(s.stream));
    GS_HIP(hipMemcpy(h, s.d_traffic, sizeof(h), hipMemcpyDeviceToHost));
    GS_HIP(hipMemset(s.d_traffic, 0, sizeof(h)));
  }
  if (flushed_buckets) *flushed_buckets = h[0];
  if (direct_updates) *direct_updates = h[1];
  return GS_OK;
}

int gs_host_alloc(size_t bytes, void** hptr) {
  GS_CHECK(hptr, GS_ERR_INVALID, "null argument");
  if (hipHostMalloc(hptr, bytes, hipHostMallocDefault) != hipSuccess)
    return fail(GS_ERR_OOM, "hipHostMalloc(%zu) failed", bytes);
  return GS_OK;
}
int gs_host_free(void* hptr) {
  GS_HIP(hipHostFree(hptr));
  return GS_OK;
}

int gs_device_alloc(size_t bytes, void** dptr) {
  GS_CHECK(dptr, GS_ERR_INVALID, "null argument");
  if (hipMalloc(dptr, bytes) != hipSuccess) return fail(GS_ERR_OOM, "hipMalloc(%zu) failed", bytes);
  return GS_OK;
}
int gs_device_free(void* dptr) {
  GS_HIP(hipFree(dptr));
  return GS_OK;
}
int gs_memcpy_h2d(void* dptr, const void* src, size_t bytes) {
  GS_HIP(hipMemcpy(dptr, src, bytes, hipMemcpyHostToDevice));
  return GS_OK;
}
int gs_memcpy_d2h(void* dst, const void* dptr, size_t bytes) {
  GS_HIP(hipMemcpy(dst, dptr, bytes, hipMemcpyDeviceToHost));
  return GS_OK;
}
int gs_device_copy(void* dst, const void* src, size_t bytes) {
  GS_CHECK(dst && src, GS_ERR_INVALID, "null argument");
  return device_copy(dst, src, bytes);
}
int gs_device_synchronize(void) {
  GS_HIP(hipDeviceSynchronize());
  return GS_OK;
}

}  // extern "C"
