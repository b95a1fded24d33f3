// SketchServer: many tenant graphs kept resident in device memory and served from one request
// queue. Tenants are admitted while their sketch memory fits the budget (later ones wait until
// memory is released). Requests (update batches, connectivity queries, min-cut queries) are
// taken in arrival order; with batching on, one scheduling step gathers every pending request of
// every admitted tenant up to and including that tenant's first pending query, applies all
// gathered updates of the queried tenants in one device submission (a store group), sets a
// fence, launches every other tenant's gathered updates, and answers all gathered connectivity
// queries with one batched Boruvka that waits for the fence only (so it runs beside those other
// updates); min-cut queries go to a second thread, which computes their forests and cuts (the
// tenant's later requests wait for it, the others go on). Update payloads stay in pinned host
// memory from arrival: each goes by DMA straight to the device, where the member's base is added
// and the ids range-checked (no host pass), and a step's launches are not waited for. A query
// therefore sees exactly the updates of its tenant that arrived before it, so every answer is
// the one the tenant would get alone. With batching off, requests are served one at a time on
// per-tenant streams (the plain single-stream path). No counterpart in the reference, whose
// algorithm objects each hold one graph and are driven by one stream.
#pragma once
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <vector>

#include "../gpusketch.h"
#include "types.h"

struct ServerConfig {
  uint64_t memory_budget_bytes = 200ull * 1000 * 1000 * 1000;  // sketch memory of admitted tenants
  uint64_t max_step_updates = 1ull << 25;  // updates per device submission when batching
  uint32_t launch_edges = 1u << 23;  // edges per launch of the group's ingest pipeline
  bool batching = true;
  // a tenant with pending requests that does not fit the budget makes room by spilling idle
  // resident tenants (least recently served first) to host memory; a spilled tenant is restored
  // bit-exact when its next request comes
  bool evict_idle = true;
  // spills go to pinned host buffers, kept for reuse by later spills once their tenant is restored
  // (else to pageable memory, allocated per spill)
  bool pinned_spill = true;
  bool keep_labels = false;  // keep the labels of every CC answer (tests)
};

struct TenantSpec {
  node_id_t num_vertices = 0;
  uint64_t seed = 0;
  double sketches_factor = 1;
};

enum class RequestKind { UPDATE, CC_QUERY, MINCUT_QUERY };

struct RequestResult {
  uint64_t id = 0;
  uint32_t tenant = 0;
  RequestKind kind = RequestKind::UPDATE;
  uint64_t updates = 0;      // UPDATE: edges in the batch
  double arrival_s = 0;      // seconds since the server started
  double done_s = 0;
  int status = 0;            // gs_status of the answer (e.g. out of samples)
  uint32_t num_cc = 0;       // CC_QUERY
  uint64_t partition_hash = 0;  // CC_QUERY: label_hash of the labels
  uint64_t step = 0;         // the scheduling step that was running when it was answered
  uint64_t mincut = 0;       // MINCUT_QUERY: the cut if below k, else k
  std::vector<node_id_t> labels;  // CC_QUERY with keep_labels
};

// The hash of a label array in RequestResult::partition_hash: FNV-1a over the 32-bit labels in
// four interleaved lanes (label i into lane i mod 4, so the four multiply chains overlap), the
// lanes then folded in order by FNV-1a.
inline uint64_t label_hash(const uint32_t* p, size_t n) {
  const uint64_t prime = 1099511628211ull;
  uint64_t h[4] = {1469598103934665603ull, 1469598103934665603ull ^ 1,
                   1469598103934665603ull ^ 2, 1469598103934665603ull ^ 3};
  size_t i = 0;
  for (; i + 4 <= n; i += 4)
    for (int l = 0; l < 4; l++) h[l] = (h[l] ^ p[i + l]) * prime;
  for (int l = 0; i < n; i++, l++) h[l] = (h[l] ^ p[i]) * prime;
  uint64_t out = 1469598103934665603ull ^ n;
  for (int l = 0; l < 4; l++) out = (out ^ h[l]) * prime;
  return out;
}

class SketchServer {
 public:
  explicit SketchServer(ServerConfig cfg);
  ~SketchServer();
  SketchServer(const SketchServer&) = delete;
  SketchServer& operator=(const SketchServer&) = delete;

  // register a tenant; it is admitted once its sketch memory fits the budget
  uint32_t add_tenant(const TenantSpec& spec);
  // release a tenant after its queued requests; its memory admits waiting tenants
  void remove_tenant(uint32_t tenant);
  bool is_admitted(uint32_t tenant);
  // resident, or spilled to host memory
  bool has_sketch(uint32_t tenant);
  // wait until every waiting tenant that fits the budget has been admitted
  void wait_admissions();
  // pins `bytes` of host memory now for spilled tenants (carved per spill, returned on restore),
  // so that evictions do not pin memory while requests wait; spills beyond it pin their own
  void reserve_spill(uint64_t bytes);
  uint64_t admitted_bytes();

  // enqueue (thread-safe, returns the request id); pairs are copied (into a pinned buffer of the
  // server's pool)
  uint64_t submit_update(uint32_t tenant, const uint32_t* pairs, uint64_t nedges);
  // the same without the copy: `pairs` must be pinned (gs_host_alloc) and stay valid and
  // unchanged until the request is answered (drain / results)
  uint64_t submit_update_pinned(uint32_t tenant, const uint32_t* pairs, uint64_t nedges);
  uint64_t submit_cc_query(uint32_t tenant);
  uint64_t submit_mincut_query(uint32_t tenant, uint32_t k);

  // wait until every request submitted so far is answered
  void drain();
  std::vector<RequestResult> results();
  // reference-ordered buckets of a tenant (after drain), for checks
  std::vector<uint64_t> export_tenant(uint32_t tenant);
  double now_s() const;
  uint64_t steps() const { return num_steps.load(); }
  uint64_t evictions() const { return num_evictions.load(); }
  uint64_t restores() const { return num_restores.load(); }
  // seconds spent spilling / restoring tenants and the bytes moved (after drain)
  struct MoveStats {
    double spill_s = 0, restore_s = 0;
    uint64_t spill_bytes = 0, restore_bytes = 0;
  };
  MoveStats move_stats() {
    std::lock_guard<std::mutex> lk(mu);
    return moves;
  }
  // worker seconds per phase of the steps (after drain): submitting updates, answering
  // connectivity queries, computing min-cut forests
  struct PhaseStats {
    double update_s = 0, cc_s = 0, mincut_s = 0;
    uint64_t cc_batches = 0, mincut_queries = 0;
    double group_s = 0;       // building store groups (batching: after membership changes)
    uint64_t group_builds = 0;
  };
  PhaseStats phase_stats() {
    std::lock_guard<std::mutex> lk(mu);
    return phases;
  }
  // message of the last device error that failed a step ("" if none)
  std::string error() {
    std::lock_guard<std::mutex> lk(mu);
    return last_error;
  }

 private:
  struct Request {
    uint64_t id = 0;
    uint32_t tenant = 0;
    RequestKind kind = RequestKind::UPDATE;
    uint32_t k = 0;
    const uint32_t* pairs = nullptr;  // UPDATE: 2 * nedges ids, pinned
    uint64_t nedges = 0;
    double arrival_s = 0;
    uint32_t* owned = nullptr;  // the pool buffer holding `pairs` (submit_update), or null
    size_t owned_words = 0;     // its capacity
  };
  struct SpillBuf {
    uint64_t* p = nullptr;
    size_t words = 0;  // in use
    size_t cap = 0;    // allocated
    bool pinned = false;
    bool in_arena = false;  // a range of the reserved arena
  };
  struct Tenant {
    TenantSpec spec;
    uint64_t bytes = 0;
    gs_store* store = nullptr;
    gs_stream* stream = nullptr;  // plain mode
    bool admitted = false, removing = false, removed = false;
    bool mc_busy = false;         // a min-cut query of it is on the cutter thread (batching)
    uint64_t queued = 0;          // requests not yet answered
    bool spilled = false;         // buckets held in `spill` (reference order), no store
    SpillBuf spill;
    uint64_t last_step = 0;       // last step that served the tenant
  };

  void loop();
  void admit_locked();
  bool plan_room_locked(std::vector<uint32_t>* evict, uint32_t* need);
  void make_resident(uint32_t need, const std::vector<uint32_t>& evict);
  SpillBuf take_spill_buffer(size_t words);
  void give_back(SpillBuf& b);
  void run_step(std::vector<Request>& step);
  // may_wait: the launches may be waited for when no request is queued (reporting dropped pairs)
  void apply_updates(std::vector<Request*>& upd, bool may_wait = true);
  void release_inflight(gs_stream* gsm);
  void take_step_locked(std::vector<Request>& step);
  void ensure_group();
  void finish(Request& r, RequestResult res);
  uint64_t enqueue(Request r);
  void release_payload_locked(Request& r);  // mu held
  // answers a request that failed (mu held)
  void fail_locked(Request& r, int status);
  void cut_loop();

  // a min-cut query handed to the cutter thread: with `store`, the cutter first computes its k
  // forests on the device (batching mode: the tenant's later requests wait, the worker goes on
  // with the others); otherwise the forests are `edges` and only the exact cut is left
  struct CutJob {
    Request req;
    uint32_t n, k;
    std::vector<uint32_t> edges;
    uint64_t m;
    gs_store* store = nullptr;
  };

  ServerConfig cfg;
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv, idle_cv;
  std::deque<Request> queue;
  std::vector<std::unique_ptr<Tenant>> tenants;
  std::vector<uint32_t> waiting;  // tenants not yet admitted, in arrival order
  uint64_t admitted = 0;
  uint64_t next_id = 0, answered = 0;
  bool stop = false;
  bool busy = false;
  bool step_turn = false;  // batching: a step's turn next (it alternates with residency work)
  std::vector<RequestResult> done;
  std::atomic<uint64_t> num_steps{0}, num_evictions{0}, num_restores{0};
  MoveStats moves;
  PhaseStats phases;
  std::multimap<size_t, uint64_t*> pinned_free;  // pinned spill buffers by capacity (words)
  // the reserved spill arena (reserve_spill): its free ranges, offset -> words, merged on return
  uint64_t* arena = nullptr;
  size_t arena_words = 0;
  std::map<size_t, size_t> arena_free;
  std::multimap<size_t, uint32_t*> payload_free;  // pinned payload buffers by capacity (words)
  // pool buffers submitted without waiting for their DMA (batching): back to payload_free once
  // the group stream's copies have drained (release_inflight)
  std::vector<std::pair<size_t, uint32_t*>> payload_inflight;
  std::string last_error;
  // store group over the admitted tenants (batching mode), rebuilt when membership changes
  gs_group* group = nullptr;
  std::vector<uint32_t> group_members;
  std::vector<uint32_t> cc_labels;  // labels buffer of the batched queries (worker only)
  std::thread worker;
  // host min-cut jobs, run by `cutter` while the worker goes on with the device
  std::deque<CutJob> cut_jobs;
  std::vector<uint64_t> step_handed;  // ids of the running step's requests given to the cutter
  bool cut_stop = false;
  std::condition_variable cut_cv;
  std::thread cutter;
  // held shared by the cutter while it computes forests on a member store of the group (whose
  // stream its sync drains) and exclusively by the worker while it destroys the group
  std::shared_mutex group_mu;
};
