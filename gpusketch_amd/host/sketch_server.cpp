// SketchServer (see sketch_server.h): admission, arrival-ordered scheduling, continuous batching
// of many tenants' updates into one store-group submission and of their connectivity queries
// into one batched Boruvka. The worker thread makes the device calls of updates and connectivity
// queries; min-cut queries go to a second thread (the cutter): with batching it computes the k
// forests on the device and the exact cut of their union while the worker serves the other
// tenants, without batching only the exact cut.
#include "gpusketch/sketch_server.h"

#include <algorithm>
#include <iterator>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>

#include "gpusketch/device_sketch_store.h"
#include "gpusketch/sketch.h"


SketchServer::SketchServer(ServerConfig c) : cfg(c), t0(std::chrono::steady_clock::now()) {
  worker = std::thread([this] { loop(); });
  cutter = std::thread([this] { cut_loop(); });
}

SketchServer::~SketchServer() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
  }
  cv.notify_all();
  worker.join();
  {
    std::lock_guard<std::mutex> lk(mu);
    cut_stop = true;
  }
  cut_cv.notify_all();
  cutter.join();
  if (group) gs_group_destroy(group);  // (waits for its copies)
  for (auto& b : payload_inflight) gs_host_free(b.second);
  for (auto& t : tenants) {
    if (t->stream) gs_stream_destroy(t->stream);
    if (t->store) gs_store_destroy(t->store);
    give_back(t->spill);
  }
  for (auto& kv : pinned_free) gs_host_free(kv.second);
  if (arena) gs_host_free(arena);
  for (auto& r : queue)
    if (r.owned) gs_host_free(r.owned);
  for (auto& kv : payload_free) gs_host_free(kv.second);
}

double SketchServer::now_s() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

uint32_t SketchServer::add_tenant(const TenantSpec& spec) {
  gs_geometry g{};
  gs_throw_if(gs_geometry_for(spec.num_vertices, spec.sketches_factor, 1, &g));
  if (g.device_bytes > cfg.memory_budget_bytes)
    throw std::invalid_argument("tenant needs " + std::to_string(g.device_bytes) +
                                " bytes, more than the whole budget");
  std::lock_guard<std::mutex> lk(mu);
  auto t = std::make_unique<Tenant>();
  t->spec = spec;
  t->bytes = g.device_bytes;
  tenants.push_back(std::move(t));
  const uint32_t id = (uint32_t)tenants.size() - 1;
  waiting.push_back(id);
  cv.notify_all();
  return id;
}

void SketchServer::remove_tenant(uint32_t t) {
  std::lock_guard<std::mutex> lk(mu);
  tenants.at(t)->removing = true;
  cv.notify_all();
}

bool SketchServer::is_admitted(uint32_t t) {
  std::lock_guard<std::mutex> lk(mu);
  return tenants.at(t)->admitted;
}

bool SketchServer::has_sketch(uint32_t t) {
  std::lock_guard<std::mutex> lk(mu);
  return tenants.at(t)->admitted || tenants.at(t)->spilled;
}

void SketchServer::wait_admissions() {
  std::unique_lock<std::mutex> lk(mu);
  cv.notify_all();
  idle_cv.wait(lk, [&] {
    return !busy && (waiting.empty() ||
                     admitted + tenants[waiting.front()]->bytes > cfg.memory_budget_bytes);
  });
}

uint64_t SketchServer::admitted_bytes() {
  std::lock_guard<std::mutex> lk(mu);
  return admitted;
}

uint64_t SketchServer::enqueue(Request r) {
  std::lock_guard<std::mutex> lk(mu);
  if (r.tenant >= tenants.size() || tenants[r.tenant]->removing) {
    release_payload_locked(r);
    throw std::invalid_argument("unknown or removed tenant");
  }
  r.id = next_id++;
  tenants[r.tenant]->queued++;
  queue.push_back(r);
  cv.notify_all();
  return next_id - 1;
}

uint64_t SketchServer::submit_update(uint32_t tenant, const uint32_t* pairs, uint64_t nedges) {
  Request r;
  r.tenant = tenant;
  r.arrival_s = now_s();
  r.nedges = nedges;
  const size_t words = 2 * (size_t)nedges;
  if (words) {
    {  // a free pool buffer of at least `words` words (at most twice that)
      std::lock_guard<std::mutex> lk(mu);
      auto it = payload_free.lower_bound(words);
      if (it != payload_free.end() && it->first <= 2 * words) {
        r.owned = it->second;
        r.owned_words = it->first;
        payload_free.erase(it);
      }
    }
    if (!r.owned) {
      size_t cap = 1u << 14;
      while (cap < words) cap *= 2;
      void* p = nullptr;
      gs_throw_if(gs_host_alloc(cap * sizeof(uint32_t), &p));
      r.owned = static_cast<uint32_t*>(p);
      r.owned_words = cap;
    }
    std::memcpy(r.owned, pairs, words * sizeof(uint32_t));
    r.pairs = r.owned;
  }
  return enqueue(r);
}

uint64_t SketchServer::submit_update_pinned(uint32_t tenant, const uint32_t* pairs,
                                            uint64_t nedges) {
  Request r;
  r.tenant = tenant;
  r.arrival_s = now_s();
  r.pairs = pairs;
  r.nedges = nedges;
  return enqueue(r);
}

uint64_t SketchServer::submit_cc_query(uint32_t tenant) {
  Request r;
  r.tenant = tenant;
  r.kind = RequestKind::CC_QUERY;
  r.arrival_s = now_s();
  return enqueue(r);
}

uint64_t SketchServer::submit_mincut_query(uint32_t tenant, uint32_t k) {
  Request r;
  r.tenant = tenant;
  r.kind = RequestKind::MINCUT_QUERY;
  r.k = k;
  r.arrival_s = now_s();
  return enqueue(r);
}

void SketchServer::release_payload_locked(Request& r) {
  if (r.owned) payload_free.emplace(r.owned_words, r.owned);
  r.owned = nullptr;
  r.pairs = nullptr;
}

void SketchServer::fail_locked(Request& r, int status) {
  RequestResult res;
  res.id = r.id;
  res.tenant = r.tenant;
  res.kind = r.kind;
  res.arrival_s = r.arrival_s;
  res.done_s = now_s();
  res.status = status;
  release_payload_locked(r);
  tenants[r.tenant]->queued--;
  answered++;
  done.push_back(std::move(res));
}

void SketchServer::drain() {
  std::unique_lock<std::mutex> lk(mu);
  idle_cv.wait(lk, [&] { return answered == next_id && !busy; });
}

std::vector<RequestResult> SketchServer::results() {
  std::lock_guard<std::mutex> lk(mu);
  return done;
}

std::vector<uint64_t> SketchServer::export_tenant(uint32_t t) {
  drain();
  std::lock_guard<std::mutex> lk(mu);
  Tenant& tn = *tenants.at(t);
  gs_store* st = tn.store;
  if (tn.spilled) {  // the snapshot through a temporary store of the same geometry
    st = nullptr;
    const uint32_t samples =
        (uint32_t)Sketch::calc_cc_samples(tn.spec.num_vertices, tn.spec.sketches_factor);
    gs_throw_if(gs_store_create(tn.spec.num_vertices, tn.spec.seed, samples, 1, 0, &st));
    const int rc = gs_store_restore(st, tn.spill.p);
    if (rc != GS_OK) gs_store_destroy(st);
    gs_throw_if(rc);
  }
  if (!st) throw std::invalid_argument("tenant has no store");
  gs_geometry g{};
  std::vector<uint64_t> out;
  int rc = gs_store_geometry(st, &g);
  if (rc == GS_OK) {
    out.resize((size_t)g.num_vertices * g.num_buckets * 2);
    rc = gs_export_buckets(st, 0, g.num_vertices, reinterpret_cast<gs_bucket*>(out.data()));
  }
  if (tn.spilled) gs_store_destroy(st);
  gs_throw_if(rc);
  return out;
}

// admit waiting tenants in arrival order while their memory fits (mu held, worker thread)
void SketchServer::admit_locked() {
  while (!waiting.empty()) {
    Tenant& t = *tenants[waiting.front()];
    if (admitted + t.bytes > cfg.memory_budget_bytes) break;
    const uint32_t samples =
        (uint32_t)Sketch::calc_cc_samples(t.spec.num_vertices, t.spec.sketches_factor);
    gs_throw_if(gs_store_create(t.spec.num_vertices, t.spec.seed, samples, 1, 0, &t.store));
    if (!cfg.batching) gs_throw_if(gs_stream_create(t.store, &t.stream));
    t.admitted = true;
    admitted += t.bytes;
    waiting.erase(waiting.begin());
  }
}

// The first tenant in queue order that has a pending request but no store: if its memory fits
// now, or once idle resident tenants (no unanswered request) are spilled, least recently served
// first, sets *need and the tenants to spill and returns true. Only the first such tenant is
// considered, so tenants become resident in the order their requests arrived. (mu held)
bool SketchServer::plan_room_locked(std::vector<uint32_t>* evict, uint32_t* need) {
  if (!cfg.evict_idle) return false;
  for (auto& r : queue) {
    const Tenant& t = *tenants[r.tenant];
    if (t.admitted || t.removed) continue;
    uint64_t free = cfg.memory_budget_bytes - admitted;
    std::vector<uint32_t> idle;
    if (t.bytes > free) {
      for (uint32_t i = 0; i < tenants.size(); i++)
        if (tenants[i]->admitted && !tenants[i]->removed && tenants[i]->queued == 0)
          idle.push_back(i);
      std::sort(idle.begin(), idle.end(), [&](uint32_t a, uint32_t b) {
        return tenants[a]->last_step != tenants[b]->last_step
                   ? tenants[a]->last_step < tenants[b]->last_step
                   : a < b;
      });
      size_t k = 0;
      while (k < idle.size() && t.bytes > free) free += tenants[idle[k++]]->bytes;
      if (t.bytes > free) return false;
      idle.resize(k);
    }
    if (evict) *evict = std::move(idle);
    if (need) *need = r.tenant;
    return true;
  }
  return false;
}

// a host buffer of at least `words` words: a free pinned one of at most twice that, else a new
// one (worker thread)
void SketchServer::reserve_spill(uint64_t bytes) {
  std::lock_guard<std::mutex> lk(mu);
  if (arena || bytes == 0) return;
  void* p = nullptr;
  const size_t words = (size_t)((bytes + 7) / 8);
  gs_throw_if(gs_host_alloc(words * sizeof(uint64_t), &p));
  arena = static_cast<uint64_t*>(p);
  arena_words = words;
  arena_free.emplace(0, words);
}

SketchServer::SpillBuf SketchServer::take_spill_buffer(size_t words) {
  SpillBuf b;
  b.words = words;
  // first fit in the reserved arena
  for (auto it = arena_free.begin(); it != arena_free.end(); ++it) {
    if (it->second < words) continue;
    const size_t off = it->first, len = it->second;
    arena_free.erase(it);
    if (len > words) arena_free.emplace(off + words, len - words);
    b.p = arena + off;
    b.cap = words;
    b.pinned = true;
    b.in_arena = true;
    return b;
  }
  if (!cfg.pinned_spill) {
    b.p = new uint64_t[words];  // uninitialised: the export writes every word
    b.cap = words;
    return b;
  }
  b.pinned = true;
  auto it = pinned_free.lower_bound(words);
  if (it != pinned_free.end() && it->first <= 2 * words) {
    b.cap = it->first;
    b.p = it->second;
    pinned_free.erase(it);
    return b;
  }
  void* p = nullptr;
  gs_throw_if(gs_host_alloc(words * sizeof(uint64_t), &p));
  b.p = static_cast<uint64_t*>(p);
  b.cap = words;
  return b;
}

void SketchServer::give_back(SpillBuf& b) {
  if (b.p && b.in_arena) {
    size_t off = (size_t)(b.p - arena), len = b.cap;
    auto next = arena_free.lower_bound(off);
    if (next != arena_free.end() && off + len == next->first) {  // merge with the range after
      len += next->second;
      next = arena_free.erase(next);
    }
    if (next != arena_free.begin()) {  // and with the one before
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        len += prev->second;
        arena_free.erase(prev);
      }
    }
    arena_free.emplace(off, len);
  } else if (b.p) {
    if (b.pinned) pinned_free.emplace(b.cap, b.p);
    else delete[] b.p;
  }
  b = SpillBuf();
}

// spills `evict` to host memory, then gives `need` a store (restored from its spill, or new)
void SketchServer::make_resident(uint32_t need, const std::vector<uint32_t>& evict) {
  if (!evict.empty() && group) {
    // drop the group before the stores it refers to (once no forests are being computed)
    std::unique_lock<std::shared_mutex> glk(group_mu);
    gs_throw_if(gs_group_destroy(group));
    group = nullptr;
    group_members.clear();
  }
  for (uint32_t i : evict) {
    const double t_start = now_s();
    Tenant& t = *tenants[i];
    uint64_t bytes = 0;
    gs_throw_if(gs_store_snapshot_bytes(t.store, &bytes));
    const size_t words = (size_t)((bytes + 7) / 8);
    SpillBuf buf = take_spill_buffer(words);
    int rc = gs_store_snapshot(t.store, buf.p);
    if (rc != GS_OK) give_back(buf);
    gs_throw_if(rc);
    if (t.stream) gs_throw_if(gs_stream_destroy(t.stream));
    gs_throw_if(gs_store_destroy(t.store));
    std::lock_guard<std::mutex> lk(mu);
    t.stream = nullptr;
    t.store = nullptr;
    t.spill = buf;
    t.spilled = true;
    t.admitted = false;
    admitted -= t.bytes;
    num_evictions++;
    moves.spill_s += now_s() - t_start;
    moves.spill_bytes += words * sizeof(uint64_t);
  }
  const double t_start = now_s();
  Tenant& t = *tenants[need];
  const uint32_t samples =
      (uint32_t)Sketch::calc_cc_samples(t.spec.num_vertices, t.spec.sketches_factor);
  gs_store* st = nullptr;
  gs_stream* stream = nullptr;
  gs_throw_if(gs_store_create(t.spec.num_vertices, t.spec.seed, samples, 1, 0, &st));
  int rc = GS_OK;
  if (t.spilled)
    rc = gs_store_restore(st, t.spill.p);
  if (rc == GS_OK && !cfg.batching) rc = gs_stream_create(st, &stream);
  if (rc != GS_OK) {
    gs_store_destroy(st);  // the spill stays, for a later attempt
    gs_throw_if(rc);
  }
  std::lock_guard<std::mutex> lk(mu);
  if (t.spilled) {
    moves.restore_s += now_s() - t_start;
    moves.restore_bytes += t.spill.words * sizeof(uint64_t);
    give_back(t.spill);
    t.spilled = false;
    num_restores++;
  } else {
    waiting.erase(std::find(waiting.begin(), waiting.end(), need));
  }
  t.store = st;
  t.stream = stream;
  t.admitted = true;
  admitted += t.bytes;
}

void SketchServer::ensure_group() {
  std::vector<uint32_t> members;
  std::vector<gs_store*> stores;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (uint32_t i = 0; i < tenants.size(); i++)
      if (tenants[i]->admitted && !tenants[i]->removed) {
        members.push_back(i);
        stores.push_back(tenants[i]->store);
      }
  }
  if (group && members == group_members) return;
  const double t_build = now_s();
  if (group) {
    std::unique_lock<std::shared_mutex> glk(group_mu);  // no forests being computed
    gs_throw_if(gs_group_destroy(group));
  }
  group = nullptr;
  group_members = members;
  if (!members.empty()) {
    gs_throw_if(gs_group_create(stores.data(), (uint32_t)stores.size(), &group));
    // a step's host batches are copied while the kernels of the launches before them run; the
    // default launch grows with the group's vertex count (to 2^27 edges), which would leave a
    // step's copies and kernels one after the other, so the loop keeps launches at
    // cfg.launch_edges (2^23: a 2^25-update step in four launches, the first three of them
    // aggregated while the next one's payloads are copied)
    gs_stream* gsm = nullptr;
    gs_throw_if(gs_group_stream(group, &gsm));
    gs_tuning tun{};
    tun.launch_edges = cfg.launch_edges;
    gs_throw_if(gs_stream_set_tuning(gsm, &tun));
    // buffers and query scratch now, not in the first step that updates or queries
    gs_throw_if(gs_group_prepare(group));
  }
  std::lock_guard<std::mutex> lk(mu);
  phases.group_s += now_s() - t_build;
  phases.group_builds++;
}

void SketchServer::finish(Request& r, RequestResult res) {
  res.id = r.id;
  res.tenant = r.tenant;
  res.kind = r.kind;
  res.arrival_s = r.arrival_s;
  res.done_s = now_s();
  res.step = num_steps.load();
  std::lock_guard<std::mutex> lk(mu);
  release_payload_locked(r);
  tenants[r.tenant]->queued--;
  answered++;
  done.push_back(std::move(res));
}

void SketchServer::apply_updates(std::vector<Request*>& upd, bool may_wait) {
  if (upd.empty()) return;
  const double t_start = now_s();
  if (cfg.batching) {
    ensure_group();
    const size_t G = group_members.size();
    std::vector<size_t> slot(tenants.size(), (size_t)-1);
    for (size_t i = 0; i < G; i++) slot[group_members[i]] = i;
    gs_stream* gsm = nullptr;
    gs_throw_if(gs_group_stream(group, &gsm));
    release_inflight(gsm);
    // each payload by DMA as it arrived (member ids; rebase and range check on the device),
    // enqueued back to back without waiting for the copies: a pool buffer stays out of the pool
    // until they have drained
    std::vector<uint32_t> members(upd.size());
    std::vector<const uint32_t*> payloads(upd.size());
    std::vector<uint64_t> counts(upd.size());
    for (size_t i = 0; i < upd.size(); i++) {
      members[i] = (uint32_t)slot[upd[i]->tenant];
      payloads[i] = upd[i]->pairs;
      counts[i] = upd[i]->nedges;
    }
    gs_throw_if(gs_group_submit_member_batches_async(group, (uint32_t)upd.size(), members.data(),
                                                     payloads.data(), counts.data()));
    {
      std::lock_guard<std::mutex> lk(mu);
      for (Request* r : upd)
        if (r->owned) {
          payload_inflight.emplace_back(r->owned_words, r->owned);
          r->owned = nullptr;
        }
    }
    // the launches are not waited for while more requests are queued (a query waits for its
    // tenant's store); the last step before the queue runs dry waits, and so reports pairs the
    // device dropped (out of range) in error()
    bool more = !may_wait;
    if (!more) {
      std::lock_guard<std::mutex> lk(mu);
      more = !queue.empty();
    }
    gs_throw_if(more ? gs_flush_async(gsm) : gs_flush(gsm));
  } else {
    for (Request* r : upd) {
      Tenant& t = *tenants[r->tenant];
      gs_throw_if(gs_submit_edge_pairs(t.stream, r->pairs, r->nedges));
      gs_throw_if(gs_flush(t.stream));
    }
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    phases.update_s += now_s() - t_start;
  }
  for (Request* r : upd) {
    RequestResult res;
    res.updates = r->nedges;
    finish(*r, std::move(res));
  }
}

void SketchServer::release_inflight(gs_stream* gsm) {
  {
    std::lock_guard<std::mutex> lk(mu);
    if (payload_inflight.empty()) return;
  }
  gs_throw_if(gs_stream_wait_copies(gsm));
  std::lock_guard<std::mutex> lk(mu);
  for (auto& b : payload_inflight) payload_free.emplace(b.first, b.second);
  payload_inflight.clear();
}

void SketchServer::run_step(std::vector<Request>& step) {
  const uint64_t step_no = ++num_steps;
  for (auto& r : step) tenants[r.tenant]->last_step = step_no;
  if (cfg.batching) ensure_group();
  std::vector<Request*> upd_q, upd_o, cc, mc;
  std::vector<char> queried(tenants.size(), 0);
  for (auto& r : step)
    if (r.kind != RequestKind::UPDATE) queried[r.tenant] = 1;
  for (auto& r : step) {
    if (r.kind == RequestKind::UPDATE) (queried[r.tenant] ? upd_q : upd_o).push_back(&r);
    else (r.kind == RequestKind::CC_QUERY ? cc : mc).push_back(&r);
  }
  // 1. the updates the step's queries depend on: those of the queried tenants
  apply_updates(upd_q);
  // with batching and connectivity queries: a fence after those, then every other tenant's
  // updates are launched at once and run beside the batched query (its own stream, waiting for
  // the fence only)
  bool others_applied = false;
  if (cfg.batching && !cc.empty()) {
    gs_throw_if(gs_group_fence(group));
    apply_updates(upd_o, false);
    others_applied = true;
  }
  // 2. connectivity queries: one batched Boruvka over the queried tenants
  const double t_cc = now_s();
  if (!cc.empty()) {
    if (cfg.batching) {
      const size_t G = group_members.size();
      std::vector<uint8_t> mask(G, 0);
      std::vector<size_t> slot(tenants.size(), (size_t)-1);
      for (size_t i = 0; i < G; i++) slot[group_members[i]] = i;
      for (Request* r : cc) mask[slot[r->tenant]] = 1;
      uint32_t total = 0, base = 0;
      gs_throw_if(gs_group_info(group, 0, &base, &total));
      if (cc_labels.size() < total) cc_labels.resize(total);  // kept across steps
      std::vector<uint32_t>& labels = cc_labels;
      std::vector<uint32_t> ncc(G), rounds(G);
      std::vector<int32_t> status(G);
      const int rc = gs_group_query_cc(group, mask.data(), labels.data(), ncc.data(),
                                       rounds.data(), status.data());
      if (rc != GS_OK && rc != GS_ERR_OUT_OF_SAMPLES) gs_throw_if(rc);
      for (Request* r : cc) {
        const size_t i = slot[r->tenant];
        gs_throw_if(gs_group_info(group, (uint32_t)i, &base, nullptr));
        const uint32_t n = tenants[r->tenant]->spec.num_vertices;
        RequestResult res;
        res.status = status[i];
        res.num_cc = ncc[i];
        res.partition_hash = label_hash(labels.data() + base, n);
        if (cfg.keep_labels) res.labels.assign(labels.begin() + base, labels.begin() + base + n);
        finish(*r, std::move(res));
      }
    } else {
      for (Request* r : cc) {
        Tenant& t = *tenants[r->tenant];
        const uint32_t n = t.spec.num_vertices;
        std::vector<uint32_t> labels(n);
        uint32_t ncc = 0, rounds = 0;
        RequestResult res;
        res.status = gs_query_cc(t.store, labels.data(), &ncc, &rounds, nullptr, 0, nullptr);
        if (res.status != GS_OK && res.status != GS_ERR_OUT_OF_SAMPLES) gs_throw_if(res.status);
        res.num_cc = ncc;
        res.partition_hash = label_hash(labels.data(), n);
        if (cfg.keep_labels) res.labels = std::move(labels);
        finish(*r, std::move(res));
      }
    }
  }
  const double t_mc = now_s();
  // 3. min-cut queries: k spanning forests on the device and the exact cut of their union (a host
  //    computation of up to tens of milliseconds). With batching both go to the cutter thread and
  //    the tenant's later requests wait for it (its store must not change under the forests);
  //    otherwise the forests are computed here and only the cut goes to the cutter.
  for (Request* r : mc) {
    Tenant& t = *tenants[r->tenant];
    const uint32_t n = t.spec.num_vertices, k = std::max<uint32_t>(1, r->k);
    if (cfg.batching) {
      std::lock_guard<std::mutex> lk(mu);
      t.mc_busy = true;
      step_handed.push_back(r->id);
      cut_jobs.push_back(CutJob{*r, n, k, {}, 0, t.store});
      cut_cv.notify_one();
      continue;
    }
    std::vector<uint32_t> edges(2 * (size_t)k * n);
    std::vector<uint64_t> counts(k);
    RequestResult res;
    res.status = gs_k_spanning_forests(t.store, k, edges.data(), (uint64_t)k * n, counts.data());
    if (res.status == GS_OK) {
      uint64_t m = 0;
      for (uint64_t c : counts) m += c;
      std::lock_guard<std::mutex> lk(mu);
      step_handed.push_back(r->id);
      cut_jobs.push_back(CutJob{*r, n, k, std::move(edges), m});
      cut_cv.notify_one();
      continue;
    }
    if (res.status != GS_ERR_OUT_OF_SAMPLES) gs_throw_if(res.status);
    finish(*r, std::move(res));
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    if (!cc.empty()) {
      phases.cc_s += t_mc - t_cc;
      phases.cc_batches += cfg.batching ? 1 : cc.size();
    }
    phases.mincut_s += now_s() - t_mc;
    phases.mincut_queries += mc.size();
  }
  // 4. every other tenant's updates (unless they went in before the queries)
  if (!others_applied) apply_updates(upd_o);
}

void SketchServer::cut_loop() {
  for (;;) {
    CutJob job;
    {
      std::unique_lock<std::mutex> lk(mu);
      cut_cv.wait(lk, [&] { return cut_stop || !cut_jobs.empty(); });
      if (cut_jobs.empty()) return;
      job = std::move(cut_jobs.front());
      cut_jobs.pop_front();
    }
    RequestResult res;
    if (job.store) {
      // the forests, on the tenant's store: its sync drains the group's stream, which the worker
      // must not destroy meanwhile
      std::shared_lock<std::shared_mutex> glk(group_mu);
      job.edges.resize(2 * (size_t)job.k * job.n);
      std::vector<uint64_t> counts(job.k);
      res.status = gs_k_spanning_forests(job.store, job.k, job.edges.data(),
                                         (uint64_t)job.k * job.n, counts.data());
      job.m = 0;
      for (uint64_t c : counts) job.m += c;
      if (res.status != GS_OK) {
        std::lock_guard<std::mutex> lk(mu);
        if (res.status != GS_ERR_OUT_OF_SAMPLES) last_error = gs_last_error();
      }
    }
    uint64_t value = 0;
    if (res.status == GS_OK) {
      res.status = gs_min_cut(job.n, job.edges.data(), job.m, &value, nullptr);
      if (res.status == GS_OK) res.mincut = std::min<uint64_t>(value, job.k);
    }
    const uint32_t tenant = job.req.tenant;
    const bool forests = job.store != nullptr;
    finish(job.req, std::move(res));
    if (forests) {
      {
        std::lock_guard<std::mutex> lk(mu);
        tenants[tenant]->mc_busy = false;
      }
      cv.notify_all();  // the tenant's later requests can run
    }
    idle_cv.notify_all();
  }
}

// The requests of the next step. Plain: the first runnable request. Batching: for each runnable
// tenant (admitted, no min-cut on the cutter), in the order of its oldest queued request, its
// run -- its queued requests up to and including its first query, so that a query sees exactly
// the updates that arrived before it -- taken whole while the step holds at most
// max_step_updates updates (a run that alone exceeds it is cut there). Whole runs rather than
// slices of every tenant's queue: a launch reads and writes the sketch buckets its updates
// touch, about as many per vertex for a few updates as for many, so each tenant's updates should
// reach the device in as few launches as its queries allow.
void SketchServer::take_step_locked(std::vector<Request>& step) {
  auto runnable = [&](uint32_t t) { return tenants[t]->admitted && !tenants[t]->mc_busy; };
  if (!cfg.batching) {
    for (auto it = queue.begin(); it != queue.end(); ++it)
      if (runnable(it->tenant)) {
        step.push_back(std::move(*it));
        queue.erase(it);
        return;
      }
    return;
  }
  // each runnable tenant's run: positions in the queue (ascending) and updates
  std::vector<std::vector<size_t>> runs(tenants.size());
  std::vector<uint64_t> run_updates(tenants.size(), 0);
  std::vector<char> closed(tenants.size(), 0);
  std::vector<uint32_t> order;  // tenants by their oldest queued request
  size_t pos = 0;
  for (auto it = queue.begin(); it != queue.end(); ++it, ++pos) {
    const uint32_t t = it->tenant;
    if (!runnable(t) || closed[t]) continue;
    if (runs[t].empty()) order.push_back(t);
    runs[t].push_back(pos);
    if (it->kind == RequestKind::UPDATE) run_updates[t] += it->nedges;
    else closed[t] = 1;  // the run ends at the tenant's first query
  }
  std::vector<char> take(queue.size(), 0);
  uint64_t nupd = 0;
  for (uint32_t t : order) {
    if (nupd + run_updates[t] <= cfg.max_step_updates) {
      for (size_t p : runs[t]) take[p] = 1;
      nupd += run_updates[t];
    } else if (nupd == 0) {  // the first run alone exceeds the cap: its prefix up to the cap
      for (size_t p : runs[t]) {
        const Request& r = queue[p];
        if (nupd > 0 && r.kind == RequestKind::UPDATE && nupd + r.nedges > cfg.max_step_updates)
          break;
        take[p] = 1;
        if (r.kind == RequestKind::UPDATE) nupd += r.nedges;
      }
      break;
    }
    if (nupd >= cfg.max_step_updates) break;  // (else a smaller run of a later tenant may fit)
  }
  std::deque<Request> rest;
  for (size_t p = 0; p < queue.size(); p++)
    (take[p] ? step.emplace_back(std::move(queue[p])) : rest.emplace_back(std::move(queue[p])));
  queue.swap(rest);
}

void SketchServer::loop() {
  for (;;) {
    std::vector<Request> step;
    std::vector<uint32_t> release, evict;
    uint32_t need = UINT32_MAX;
    {
      std::unique_lock<std::mutex> lk(mu);
      auto runnable = [&] {
        for (auto& r : queue)
          if (tenants[r.tenant]->admitted && !tenants[r.tenant]->mc_busy) return true;
        return false;
      };
      auto removable = [&] {
        for (auto& t : tenants)
          if (t->removing && !t->removed && t->queued == 0) return true;
        return false;
      };
      cv.wait(lk, [&] {
        return stop || runnable() || removable() ||
               (!waiting.empty() && admitted + tenants[waiting.front()]->bytes <= cfg.memory_budget_bytes) ||
               plan_room_locked(nullptr, nullptr);
      });
      if (stop && queue.empty()) return;
      for (uint32_t i = 0; i < tenants.size(); i++)
        if (tenants[i]->removing && !tenants[i]->removed && tenants[i]->queued == 0) {
          tenants[i]->removed = true;
          release.push_back(i);
        }
      if (release.empty()) admit_locked();
      bool room = release.empty() && plan_room_locked(&evict, &need);
      // with batching, making a waiting tenant resident (spills, a restore) and serving the
      // resident tenants take turns, so neither waits for all of the other
      if (room && cfg.batching && runnable()) {
        step_turn = !step_turn;
        if (step_turn) {
          room = false;
          evict.clear();
          need = UINT32_MAX;
        }
      }
      if (release.empty() && !room) take_step_locked(step);
      busy = true;
      step_handed.clear();
    }
    if (!release.empty()) {
      // drop the group before the stores it refers to (once no forests are being computed)
      if (group) {
        std::unique_lock<std::shared_mutex> glk(group_mu);
        gs_throw_if(gs_group_destroy(group));
      }
      group = nullptr;
      group_members.clear();
      std::lock_guard<std::mutex> lk(mu);
      for (uint32_t i : release) {
        Tenant& t = *tenants[i];
        if (t.stream) gs_stream_destroy(t.stream);
        if (t.store) gs_store_destroy(t.store);
        t.stream = nullptr;
        t.store = nullptr;
        if (t.admitted) admitted -= t.bytes;
        t.admitted = false;
        give_back(t.spill);
        t.spilled = false;
        auto w = std::find(waiting.begin(), waiting.end(), i);
        if (w != waiting.end()) waiting.erase(w);
      }
      admit_locked();
    } else if (need != UINT32_MAX) {
      try {
        make_resident(need, evict);
      } catch (const std::exception& e) {
        // a tenant that cannot be made resident fails its pending requests
        std::lock_guard<std::mutex> lk(mu);
        last_error = e.what();
        for (auto it = queue.begin(); it != queue.end();) {
          if (it->tenant != need) {
            ++it;
            continue;
          }
          fail_locked(*it, GS_ERR_HIP);
          it = queue.erase(it);
        }
      }
    } else {
      try {
        run_step(step);
      } catch (const std::exception& e) {
        // a device error answers every unanswered request of the step with the error
        std::lock_guard<std::mutex> lk(mu);
        last_error = e.what();
        for (auto& r : step) {
          // answered already, or answered by the cutter
          bool answered_already =
              std::find(step_handed.begin(), step_handed.end(), r.id) != step_handed.end();
          for (auto& d : done)
            if (d.id == r.id) answered_already = true;
          if (answered_already) {
            release_payload_locked(r);
            continue;
          }
          fail_locked(r, GS_ERR_HIP);
        }
      }
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      busy = false;
    }
    idle_cv.notify_all();
  }
}
