// Serving benchmark and mix-versus-alone check for SketchServer.
//   serve_bench [--tenants 64] [--requests 1000] [--rate 2000] [--budget-gb 200] [--seed 7]
//               [--mc-tenants 8] [--k 4] [--min-n 4096] [--max-n 262144]
//               [--max-edges 1000000] [--alone-check 0|1] [--evict 0|1]
//               [--max-step-updates N] [--pinned-spill 0|1] [--verify 0|1]
//               [--cc-frac 0.01] [--mc-frac 0.005] [--pinned-payloads 0|1] [--dump PREFIX]
//               [--reserve-spill 0|1]
// Tenants: n log-uniform in [min-n, max-n]. Requests arrive as a Poisson process of the given
// rate (requests/s): a cc-frac share of connectivity queries on random tenants, mc-frac of
// min-cut queries (k) on the first mc-tenants tenants, the rest update batches of log-uniform
// 1k..max-edges edges, 5% of
// which delete an earlier inserted edge of the tenant. The same trace runs twice: plain (one
// request at a time on per-tenant streams) and served (continuous batching). Prints one JSON
// line: updates/s and p50/p99 latencies per mode, and whether every answer and every final
// sketch is identical in both modes (and, with --alone-check, for a tenant served alone). With a
// budget below the tenants' total, --evict 1 (the default) spills idle tenants to host memory to
// serve the others (into pinned memory reserved before the clock starts, about 1.25x the
// tenants' bytes beyond the budget, unless --reserve-spill 0); --evict 0 serves only the
// tenants admitted before the first request.
// --verify 1 replays the trace on the host and checks every answer of both modes against the
// exact graph: each CC answer against a union-find over the tenant's net graph at that point of
// the trace (component count and labels), each min-cut answer against min(k, exact minimum cut).
// With --pinned-payloads 1 (the default) the update payloads are placed in pinned host memory
// before the clock starts and handed over without a copy (SketchServer::submit_update_pinned),
// as a server that receives requests into pinned buffers would; with 0 the server copies each
// payload at arrival (submit_update). Both modes of a run take the payloads the same way.
// --dump PREFIX writes every answered request of each mode to PREFIX_plain.csv and
// PREFIX_served.csv (trace index, kind, tenant, edges, arrival and answer times, latency, step).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <string>
#include <thread>
#include <unordered_set>

#include "gpusketch.h"
#include "gpusketch/sketch_server.h"

struct Req {
  double t;  // arrival offset (s)
  uint32_t tenant;
  RequestKind kind;
  std::vector<uint32_t> pairs;
  const uint32_t* pinned = nullptr;  // the same pairs in the trace's pinned arena
};

static double g_cc_frac = 0.01, g_mc_frac = 0.005;
static bool g_pinned_payloads = true;

struct Trace {
  std::vector<TenantSpec> tenants;
  std::vector<Req> reqs;
  uint64_t total_updates = 0;
};

static Trace make_trace(int ntenants, int nreq, double rate, uint64_t seed, int mc_tenants,
                        uint32_t min_n, uint32_t max_n, uint64_t max_edges) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  Trace tr;
  for (int i = 0; i < ntenants; i++) {
    const double lg = std::log2((double)min_n) + U(rng) * (std::log2((double)max_n) - std::log2((double)min_n));
    TenantSpec s;
    s.num_vertices = (node_id_t)std::floor(std::pow(2.0, lg));
    s.seed = 0xC0DE0000ull + 7919ull * i;
    tr.tenants.push_back(s);
  }
  std::vector<std::vector<uint32_t>> history(ntenants);
  double t = 0;
  std::exponential_distribution<double> gap(rate > 0 ? rate : 1.0);
  for (int q = 0; q < nreq; q++) {
    if (rate > 0) t += gap(rng);
    Req r;
    r.t = t;
    const double x = U(rng);
    if (x < g_cc_frac) {
      r.kind = RequestKind::CC_QUERY;
      r.tenant = (uint32_t)(rng() % ntenants);
    } else if (x < g_cc_frac + g_mc_frac && mc_tenants > 0) {
      r.kind = RequestKind::MINCUT_QUERY;
      r.tenant = (uint32_t)(rng() % std::min(mc_tenants, ntenants));
    } else {
      r.kind = RequestKind::UPDATE;
      r.tenant = (uint32_t)(rng() % ntenants);
      const uint32_t n = tr.tenants[r.tenant].num_vertices;
      const uint64_t m = (uint64_t)std::floor(std::exp(std::log(1000.0) + U(rng) * (std::log((double)max_edges) - std::log(1000.0))));
      auto& h = history[r.tenant];
      r.pairs.resize(2 * m);
      for (uint64_t e = 0; e < m; e++) {
        uint32_t u, v;
        if (!h.empty() && U(rng) < 0.05) {  // delete an earlier inserted edge
          const size_t j = rng() % (h.size() / 2);
          u = h[2 * j];
          v = h[2 * j + 1];
        } else {
          do {
            u = (uint32_t)(rng() % n);
            v = (uint32_t)(rng() % n);
          } while (u == v);
          h.push_back(u);
          h.push_back(v);
        }
        r.pairs[2 * e] = u;
        r.pairs[2 * e + 1] = v;
      }
      tr.total_updates += m;
    }
    tr.reqs.push_back(std::move(r));
  }
  return tr;
}

struct RunStats {
  double makespan = 0;
  double upd_per_s = 0;
  std::map<std::string, std::pair<double, double>> p50p99;  // kind -> (p50 ms, p99 ms)
  uint64_t steps = 0, evictions = 0, restores = 0;
  SketchServer::MoveStats moves;
  SketchServer::PhaseStats phases;
  std::map<uint64_t, RequestResult> by_index;  // trace index -> result
  std::vector<std::vector<uint64_t>> exports;
  std::string error;
};

static double pct(std::vector<double> v, double p) {
  if (v.empty()) return 0;
  std::sort(v.begin(), v.end());
  const double idx = p / 100.0 * (v.size() - 1);
  const size_t lo = (size_t)std::floor(idx), hi = (size_t)std::ceil(idx);
  return v[lo] + (v[hi] - v[lo]) * (idx - lo);
}

static uint64_t g_max_step_updates = 0;  // 0: the server's default
static uint32_t g_launch_edges = 0;  // 0: the server's default
static std::string g_dump;                 // --dump prefix ("" = none)

static void dump_csv(const RunStats& s, const char* mode) {
  if (g_dump.empty()) return;
  const std::string path = g_dump + "_" + mode + ".csv";
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) {
    std::fprintf(stderr, "[serve] cannot write %s\n", path.c_str());
    return;
  }
  std::fprintf(f, "index,kind,tenant,edges,arrival_ms,done_ms,latency_ms,step\n");
  for (auto& kv : s.by_index) {
    const RequestResult& r = kv.second;
    const char* kind = r.kind == RequestKind::UPDATE ? "update"
                       : r.kind == RequestKind::CC_QUERY ? "cc" : "mincut";
    std::fprintf(f, "%llu,%s,%u,%llu,%.4f,%.4f,%.4f,%llu\n", (unsigned long long)kv.first, kind,
                 r.tenant, (unsigned long long)r.updates, r.arrival_s * 1e3, r.done_s * 1e3,
                 (r.done_s - r.arrival_s) * 1e3, (unsigned long long)r.step);
  }
  std::fclose(f);
}
static bool g_evict = true, g_pinned_spill = true, g_reserve_spill = true;

static RunStats run(const Trace& tr, bool batching, uint64_t budget, uint32_t k,
                    const std::vector<uint32_t>& only_tenants, bool keep_exports) {
  ServerConfig cfg;
  cfg.batching = batching;
  cfg.memory_budget_bytes = budget;
  if (g_max_step_updates) cfg.max_step_updates = g_max_step_updates;
  if (g_launch_edges) cfg.launch_edges = g_launch_edges;
  cfg.evict_idle = g_evict;
  cfg.pinned_spill = g_pinned_spill;
  SketchServer srv(cfg);
  std::vector<int> tid(tr.tenants.size(), -1);
  for (size_t i = 0; i < tr.tenants.size(); i++)
    if (only_tenants.empty() ||
        std::find(only_tenants.begin(), only_tenants.end(), (uint32_t)i) != only_tenants.end())
      tid[i] = (int)srv.add_tenant(tr.tenants[i]);
  if (g_evict && g_reserve_spill) {
    // pinned spill space for the tenants beyond the budget (with slack for fragmentation),
    // reserved before the clock starts as a server would at start-up
    uint64_t total = 0;
    for (size_t i = 0; i < tr.tenants.size(); i++) {
      if (tid[i] < 0) continue;
      gs_geometry g{};
      if (gs_geometry_for(tr.tenants[i].num_vertices, tr.tenants[i].sketches_factor, 1, &g) == GS_OK)
        total += g.device_bytes + 64;
    }
    if (total > budget) srv.reserve_spill((total - budget) + (total - budget) / 4);
  }
  srv.wait_admissions();  // the clock then measures requests only
  std::map<uint64_t, uint64_t> id_to_index;
  const double start = srv.now_s();
  for (size_t q = 0; q < tr.reqs.size(); q++) {
    const Req& r = tr.reqs[q];
    if (tid[r.tenant] < 0 || (!g_evict && !srv.is_admitted((uint32_t)tid[r.tenant]))) continue;
    const double due = start + r.t;
    double now = srv.now_s();
    if (due > now) std::this_thread::sleep_for(std::chrono::duration<double>(due - now));
    uint64_t id = 0;
    const uint32_t t = (uint32_t)tid[r.tenant];
    if (r.kind == RequestKind::UPDATE)
      id = r.pinned ? srv.submit_update_pinned(t, r.pinned, r.pairs.size() / 2)
                    : srv.submit_update(t, r.pairs.data(), r.pairs.size() / 2);
    else if (r.kind == RequestKind::CC_QUERY) id = srv.submit_cc_query(t);
    else id = srv.submit_mincut_query(t, k);
    id_to_index[id] = q;
  }
  srv.drain();
  RunStats st;
  st.error = srv.error();
  auto res = srv.results();
  double last = start, first = 1e300;
  std::map<std::string, std::vector<double>> lat;
  uint64_t upd = 0;
  for (auto& r : res) {
    last = std::max(last, r.done_s);
    first = std::min(first, r.arrival_s);
    const char* kind = r.kind == RequestKind::UPDATE ? "update"
                       : r.kind == RequestKind::CC_QUERY ? "cc_query" : "mincut_query";
    lat[kind].push_back((r.done_s - r.arrival_s) * 1e3);
    upd += r.updates;
    st.by_index[id_to_index[r.id]] = r;
  }
  st.makespan = last - std::min(first, start);
  st.upd_per_s = upd / st.makespan;
  for (auto& kv : lat) st.p50p99[kv.first] = {pct(kv.second, 50), pct(kv.second, 99)};
  st.steps = srv.steps();
  st.evictions = srv.evictions();
  st.restores = srv.restores();
  st.moves = srv.move_stats();
  st.phases = srv.phase_stats();
  if (keep_exports)
    for (size_t i = 0; i < tr.tenants.size(); i++)
      if (tid[i] >= 0 && srv.has_sketch((uint32_t)tid[i]))
        st.exports.push_back(srv.export_tenant((uint32_t)tid[i]));
  return st;
}

static bool same_answers(const RunStats& a, const RunStats& b, std::string* why) {
  for (auto& kv : b.by_index) {
    auto it = a.by_index.find(kv.first);
    if (it == a.by_index.end()) continue;
    const RequestResult &x = it->second, &y = kv.second;
    if (x.status != y.status || x.num_cc != y.num_cc || x.partition_hash != y.partition_hash ||
        x.mincut != y.mincut) {
      *why = "request " + std::to_string(kv.first) + " differs";
      return false;
    }
  }
  return true;
}


static uint32_t uf_find(std::vector<uint32_t>& p, uint32_t x) {
  while (p[x] != x) x = p[x] = p[p[x]];
  return x;
}

struct Verified {
  uint64_t cc = 0, mincut = 0, mismatches = 0;
  std::string first;
};

// The net graph of every tenant replayed in trace order; at each query, the exact answer from a
// union-find over it (labels: smallest vertex of each component) and, for min-cut queries,
// min(k, the exact global minimum cut) -- compared with what each run answered.
static Verified verify_answers(const Trace& tr, const std::vector<const RunStats*>& runs, uint32_t k) {
  Verified out;
  std::vector<std::unordered_set<uint64_t>> live(tr.tenants.size());
  for (size_t q = 0; q < tr.reqs.size(); q++) {
    const Req& r = tr.reqs[q];
    auto& L = live[r.tenant];
    if (r.kind == RequestKind::UPDATE) {
      for (size_t e = 0; 2 * e < r.pairs.size(); e++) {
        const uint32_t u = r.pairs[2 * e], v = r.pairs[2 * e + 1];
        const uint64_t key = u < v ? ((uint64_t)u << 32 | v) : ((uint64_t)v << 32 | u);
        if (!L.erase(key)) L.insert(key);
      }
      continue;
    }
    const uint32_t n = tr.tenants[r.tenant].num_vertices;
    std::vector<uint32_t> p(n);
    for (uint32_t v = 0; v < n; v++) p[v] = v;
    for (uint64_t key : L) {
      uint32_t a = uf_find(p, (uint32_t)(key >> 32)), b = uf_find(p, (uint32_t)key);
      if (a != b) p[std::max(a, b)] = std::min(a, b);  // the smaller id stays the root
    }
    std::vector<uint32_t> lab(n);
    uint32_t ncc = 0;
    for (uint32_t v = 0; v < n; v++) {
      lab[v] = uf_find(p, v);
      ncc += lab[v] == v;
    }
    uint64_t cut = 0;
    if (r.kind == RequestKind::MINCUT_QUERY && ncc == 1) {
      std::vector<uint32_t> pairs;
      pairs.reserve(2 * L.size());
      for (uint64_t key : L) {
        pairs.push_back((uint32_t)(key >> 32));
        pairs.push_back((uint32_t)key);
      }
      uint64_t value = 0;
      if (gs_min_cut(n, pairs.data(), L.size(), &value, nullptr) != GS_OK) {
        out.mismatches++;
        if (out.first.empty()) out.first = "exact min cut failed: " + std::string(gs_last_error());
        continue;
      }
      cut = std::min<uint64_t>(value, k);
    }
    const uint64_t hash = label_hash(lab.data(), n);
    for (const RunStats* rs : runs) {
      auto it = rs->by_index.find(q);
      if (it == rs->by_index.end()) continue;  // not served in this run (tenant not admitted)
      const RequestResult& a = it->second;
      bool ok = a.status == 0;
      if (r.kind == RequestKind::CC_QUERY) ok = ok && a.num_cc == ncc && a.partition_hash == hash;
      else ok = ok && a.mincut == cut;
      if (!ok) {
        out.mismatches++;
        if (out.first.empty())
          out.first = "request " + std::to_string(q) + " (tenant " + std::to_string(r.tenant) +
                      ") answered differently from the exact graph";
      }
    }
    (r.kind == RequestKind::CC_QUERY ? out.cc : out.mincut)++;
  }
  return out;
}

static std::string stats_json(const RunStats& s) {
  std::string j = "{\"updates_per_s\": " + std::to_string(s.upd_per_s) +
                  ", \"makespan_s\": " + std::to_string(s.makespan) +
                  ", \"steps\": " + std::to_string(s.steps) +
                  ", \"evictions\": " + std::to_string(s.evictions) +
                  ", \"restores\": " + std::to_string(s.restores) +
                  ", \"spill_gb\": " + std::to_string(s.moves.spill_bytes / 1e9) +
                  ", \"spill_s\": " + std::to_string(s.moves.spill_s) +
                  ", \"restore_gb\": " + std::to_string(s.moves.restore_bytes / 1e9) +
                  ", \"restore_s\": " + std::to_string(s.moves.restore_s) +
                  ", \"worker_s\": {\"updates\": " + std::to_string(s.phases.update_s) +
                  ", \"cc_queries\": " + std::to_string(s.phases.cc_s) +
                  ", \"mincut_forests\": " + std::to_string(s.phases.mincut_s) +
                  ", \"cc_batches\": " + std::to_string(s.phases.cc_batches) +
                  ", \"group_builds\": " + std::to_string(s.phases.group_builds) +
                  ", \"group_build_s\": " + std::to_string(s.phases.group_s) + "}";
  for (auto& kv : s.p50p99)
    j += ", \"" + kv.first + "_ms\": {\"p50\": " + std::to_string(kv.second.first) +
         ", \"p99\": " + std::to_string(kv.second.second) + "}";
  return j + "}";
}

int main(int argc, char** argv) {
  int ntenants = 64, nreq = 1000, mc = 8, alone = 1, verify = 0;
  double rate = 2000, budget_gb = 200;
  uint64_t seed = 7, max_edges = 1000000;
  uint32_t k = 4, min_n = 4096, max_n = 262144;
  for (int i = 1; i + 1 < argc; i += 2) {
    std::string a = argv[i];
    const char* v = argv[i + 1];
    if (a == "--tenants") ntenants = std::atoi(v);
    else if (a == "--requests") nreq = std::atoi(v);
    else if (a == "--rate") rate = std::atof(v);
    else if (a == "--budget-gb") budget_gb = std::atof(v);
    else if (a == "--seed") seed = std::stoull(v);
    else if (a == "--mc-tenants") mc = std::atoi(v);
    else if (a == "--k") k = (uint32_t)std::atoi(v);
    else if (a == "--min-n") min_n = (uint32_t)std::atoi(v);
    else if (a == "--max-n") max_n = (uint32_t)std::atoi(v);
    else if (a == "--max-edges") max_edges = std::stoull(v);
    else if (a == "--alone-check") alone = std::atoi(v);
    else if (a == "--max-step-updates") g_max_step_updates = std::stoull(v);
    else if (a == "--launch-edges") g_launch_edges = (uint32_t)std::stoull(v);
    else if (a == "--evict") g_evict = std::atoi(v) != 0;
    else if (a == "--pinned-spill") g_pinned_spill = std::atoi(v) != 0;
    else if (a == "--verify") verify = std::atoi(v);
    else if (a == "--cc-frac") g_cc_frac = std::atof(v);
    else if (a == "--mc-frac") g_mc_frac = std::atof(v);
    else if (a == "--pinned-payloads") g_pinned_payloads = std::atoi(v) != 0;
    else if (a == "--dump") g_dump = v;
    else if (a == "--reserve-spill") g_reserve_spill = std::atoi(v) != 0;
    else {
      std::fprintf(stderr, "unknown option %s\n", a.c_str());
      return 2;
    }
  }
  const uint64_t budget = (uint64_t)(budget_gb * 1e9);
  auto t0 = std::chrono::steady_clock::now();
  Trace tr = make_trace(ntenants, nreq, rate, seed, mc, min_n, max_n, max_edges);
  std::fprintf(stderr, "[serve] trace: %d tenants, %zu requests, %llu updates (%.1fs)\n",
               ntenants, tr.reqs.size(), (unsigned long long)tr.total_updates,
               std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  void* arena = nullptr;
  if (g_pinned_payloads && tr.total_updates) {
    if (gs_host_alloc(tr.total_updates * 2 * sizeof(uint32_t), &arena) != GS_OK) {
      std::fprintf(stderr, "pinned arena: %s\n", gs_last_error());
      return 2;
    }
    uint32_t* w = static_cast<uint32_t*>(arena);
    for (auto& r : tr.reqs) {
      if (r.kind != RequestKind::UPDATE) continue;
      std::memcpy(w, r.pairs.data(), r.pairs.size() * sizeof(uint32_t));
      r.pinned = w;
      w += r.pairs.size();
    }
  }
  RunStats plain = run(tr, false, budget, k, {}, true);
  std::fprintf(stderr, "[serve] plain: %.3e updates/s over %.3fs\n", plain.upd_per_s, plain.makespan);
  dump_csv(plain, "plain");
  RunStats served = run(tr, true, budget, k, {}, true);
  std::fprintf(stderr, "[serve] served: %.3e updates/s over %.3fs in %llu steps\n",
               served.upd_per_s, served.makespan, (unsigned long long)served.steps);
  dump_csv(served, "served");
  std::string why;
  bool same = same_answers(plain, served, &why);
  bool same_sketch = plain.exports == served.exports;
  bool alone_ok = true;
  if (alone) {
    // the first tenant with a query, served alone, must answer exactly as in the mix
    uint32_t t = 0;
    for (auto& r : tr.reqs)
      if (r.kind != RequestKind::UPDATE) {
        t = r.tenant;
        break;
      }
    RunStats solo = run(tr, true, budget, k, {t}, false);
    alone_ok = same_answers(served, solo, &why);
  }
  uint64_t nq = 0;
  for (auto& r : tr.reqs) nq += r.kind != RequestKind::UPDATE;
  std::string vjson = "null";
  bool verified_ok = true;
  if (verify) {
    const auto v0 = std::chrono::steady_clock::now();
    const Verified v = verify_answers(tr, {&plain, &served}, k);
    verified_ok = v.mismatches == 0;
    std::fprintf(stderr, "[serve] verify: %llu CC and %llu min-cut answers per mode against the "
                 "exact graphs, %llu mismatches (%.1fs)\n", (unsigned long long)v.cc,
                 (unsigned long long)v.mincut, (unsigned long long)v.mismatches,
                 std::chrono::duration<double>(std::chrono::steady_clock::now() - v0).count());
    vjson = "{\"cc_answers\": " + std::to_string(v.cc) + ", \"mincut_answers\": " +
            std::to_string(v.mincut) + ", \"mismatches\": " + std::to_string(v.mismatches) +
            ", \"first\": \"" + v.first + "\"}";
  }
  uint64_t ncc = 0;
  for (auto& r : tr.reqs) ncc += r.kind == RequestKind::CC_QUERY;
  if (arena) gs_host_free(arena);
  std::printf("{\"tenants\": %d, \"requests\": %zu, \"queries\": %llu, \"cc_queries\": %llu, "
              "\"pinned_payloads\": %s, \"updates\": %llu, "
              "\"rate_req_per_s\": %g, \"budget_gb\": %g, \"k\": %u, \"plain\": %s, "
              "\"served\": %s, \"answers_identical\": %s, \"sketches_identical\": %s, "
              "\"alone_equals_mix\": %s, \"verified\": %s, \"error\": \"%s\"}\n",
              ntenants, tr.reqs.size(), (unsigned long long)nq, (unsigned long long)ncc,
              g_pinned_payloads ? "true" : "false", (unsigned long long)tr.total_updates, rate,
              budget_gb, k,
              stats_json(plain).c_str(), stats_json(served).c_str(), same ? "true" : "false",
              same_sketch ? "true" : "false", alone_ok ? "true" : "false", vjson.c_str(),
              (plain.error + served.error).c_str());
  return (same && same_sketch && alone_ok && verified_ok && plain.error.empty() &&
          served.error.empty()) ? 0 : 1;
}
